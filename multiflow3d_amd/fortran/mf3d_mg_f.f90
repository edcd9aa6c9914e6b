! mf3d_mg_f.f90 -- ISO_C_BINDING layer over include/mf3d_mg.h.
!
! Lets a Fortran host -- MultiFlow3D's time loop (flosol.for:209-211 calls
! newsolv_mg) -- drive the gfx950 pressure multigrid with its own arrays:
! each dom(ib)%su / %sup / %p / %pp / %dens is passed by address (c_loc), in
! its own layout (column-major, ttc_i x ttc_j x ttc_k, pl ghost layers).
! The derived types mirror the C structs field by field (bind(c)).
module mf3d_mg_f
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: MF3D_OK = 0, MF3D_NOT_CONVERGED = 1, MF3D_DIVERGED = 2
  integer(c_int), parameter, public :: MF3D_FP64 = 0, MF3D_MIXED = 1
  integer(c_int), parameter, public :: MF3D_COARSE_REFERENCE = 0, MF3D_COARSE_FACE = 1
  integer(c_int), parameter, public :: MF3D_OUTER_DEFECT = 0, MF3D_OUTER_GCR = 1

  ! one dom(ib) (module_multidata.for:16-28,36,116)
  type, bind(c), public :: mf3d_block_desc
    integer(c_int32_t) :: n(3)       ! iep-isp+1, jep-jsp+1, kep-ksp+1
    integer(c_int32_t) :: ghost      ! pl
    integer(c_int32_t) :: ld(3)      ! ttc_i, ttc_j, ttc_k
    real(c_double)     :: h(3)       ! dx, dy, dz
    integer(c_int32_t) :: nbr(6)     ! iprev, inext, jprev, jnext, kprev, knext
    integer(c_int32_t) :: per(6)     ! per_ip, per_in, per_jp, per_jn, per_kp, per_kn
    integer(c_int32_t) :: bc(6)      ! bc_west .. bc_top
    integer(c_int32_t) :: rdiv = 1   ! rdiv(dom_id(ib)); n and h at the block's own resolution
  end type

  ! the MG scalars of control.cin (initial.for:20-26)
  type, bind(c), public :: mf3d_mg_cfg
    integer(c_int32_t) :: nblocks, levels, smoother, nu_pre, nu_post, maxcy
    real(c_double)     :: eps
    integer(c_int32_t) :: min_cycles, precision, variable_coef, device
    integer(c_int32_t) :: coarse_op = 0, krylov = 0  ! reference hierarchy, defect correction
    integer(c_int32_t) :: coarse_solve = 0           ! direct in mixed precision, smoothing in FP64
    integer(c_int32_t) :: differencing = 1, pl_ex = 0 ! control.cin's differencing and pl_ex
  end type

  public :: mf3d_mg_create, mf3d_mg_set_density, mf3d_mg_vcycle, mf3d_mg_solve
  public :: mf3d_mg_newsolv, mf3d_mg_heaviside, mf3d_mg_destroy, mf3d_mg_error

  interface
    integer(c_int) function mf3d_mg_create(cfg, blocks, out) bind(c, name='mf3d_mg_create')
      import :: c_int, c_ptr, mf3d_mg_cfg, mf3d_block_desc
      type(mf3d_mg_cfg), intent(in) :: cfg
      type(mf3d_block_desc), intent(in) :: blocks(*)
      type(c_ptr), intent(out) :: out
    end function

    integer(c_int) function mf3d_mg_set_density(mg, dens) bind(c, name='mf3d_mg_set_density')
      import :: c_int, c_ptr
      type(c_ptr), value :: mg
      type(c_ptr), intent(in) :: dens(*)
    end function

    integer(c_int) function mf3d_mg_vcycle(mg, rhs, pp) bind(c, name='mf3d_mg_vcycle')
      import :: c_int, c_ptr
      type(c_ptr), value :: mg
      type(c_ptr), intent(in) :: rhs(*), pp(*)
    end function

    integer(c_int) function mf3d_mg_solve(mg, f, p, rmax, iters, hist) bind(c, name='mf3d_mg_solve')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: mg
      type(c_ptr), intent(in) :: f(*), p(*)
      real(c_double), intent(out) :: rmax
      integer(c_int32_t), intent(out) :: iters
      real(c_double), intent(out) :: hist(*)
    end function

    ! newsolv_mg in the velocity form: dom(ib)%ustar/%vstar/%wstar in,
    ! %u/%v/%w/%p in and out, %su (or %sup) out (c_null_ptr entries or a
    ! null array pointer skip it); eps, dt, alfapr as in control.cin
    integer(c_int) function mf3d_mg_newsolv(mg, ustar, vstar, wstar, u, v, w, p, dt, alfapr, eps, su, &
                                            rmax, iters, hist) bind(c, name='mf3d_mg_newsolv')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: mg
      type(c_ptr), intent(in) :: ustar(*), vstar(*), wstar(*), u(*), v(*), w(*), p(*)
      real(c_double), value :: dt, alfapr, eps
      type(c_ptr), value :: su
      real(c_double), intent(out) :: rmax
      integer(c_int32_t), intent(out) :: iters
      real(c_double), intent(out) :: hist(*)
    end function

    ! heaviside (lsm.for:1456-1509) + bound_lsm(18/19), then coeff from the
    ! new density: dom(ib)%phi in; dom(ib)%dens and %mu out over the
    ! interior and the first ghost layer (c_loc of an array of their
    ! addresses, or c_null_ptr to skip); densl, densg, mul, mug as in
    ! control.cin
    integer(c_int) function mf3d_mg_heaviside(mg, phi, densl, densg, mul, mug, dens, mu) &
        bind(c, name='mf3d_mg_heaviside')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: mg
      type(c_ptr), intent(in) :: phi(*)
      real(c_double), value :: densl, densg, mul, mug
      type(c_ptr), value :: dens, mu
    end function

    subroutine mf3d_mg_destroy(mg) bind(c, name='mf3d_mg_destroy')
      import :: c_ptr
      type(c_ptr), value :: mg
    end subroutine

    type(c_ptr) function c_last_error(mg) bind(c, name='mf3d_mg_last_error')
      import :: c_ptr
      type(c_ptr), value :: mg
    end function

    integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! mf3d_mg_last_error as a Fortran string
  function mf3d_mg_error(mg) result(msg)
    type(c_ptr), intent(in) :: mg
    character(len=:), allocatable :: msg
    type(c_ptr) :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    s = c_last_error(mg)
    n = int(c_strlen(s))
    call c_f_pointer(s, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

end module mf3d_mg_f
