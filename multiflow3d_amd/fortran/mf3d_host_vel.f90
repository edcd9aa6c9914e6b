! mf3d_host_vel.f90 -- a minimal Fortran host for the velocity form of
! newsolv_mg on the gfx950 library.
!
! It holds the fields the way MultiFlow3D does -- per block, column-major
! arrays with pl = 2 ghost layers (initial.for:199-204): ustar, vstar,
! wstar, u, v, w, p, su, dens (phi, mu) -- and replaces the whole
! newsolv_mg call (newsolv_mg.for:2-102) with one mf3d_mg_newsolv through
! the ISO_C_BINDING module, passing each block's arrays by address.  With
! lsm = 2 the density comes from a level set first, as the reference's
! step does (flosol.for:114-117): mf3d_mg_heaviside replaces heaviside
! (lsm.for:1456-1509) and the coeff of newsolv_mg.for:12-14.  The outer
! ghost layer starts as a marker: newsolv's exchanges write parts of it
! as the reference's do (whole periodic planes, 3-deep corner and edge
! cubes), heaviside leaves it alone.
!
! Usage: mf3d_host_vel <control file>
!   nbx nby nbz levels lsm maxcy
!   eps dt alfapr
!   (lsm = 2 only) densl densg mul mug
!   bc(6)
!   ex(nbx) / ey(nby) / ez(nbz)
!   dx dy dz
!   input file: per block ustar, vstar, wstar, p (and dens under lsm = 1,
!     phi under lsm = 2), each the block's one-ghost-layer box (n+2)^3,
!     f64, i fastest
!   output file
! Output (f64 stream): iters, rmax(0:iters), then per block the whole
! arrays (pl layers) of p, u, v, w, the interior of su, (lsm = 2) the
! one-layer boxes of dens and mu, and the count of outer-layer entries of
! dens and mu that changed.
program mf3d_host_vel
  use iso_c_binding
  use mf3d_mg_f
  implicit none
  integer, parameter :: pl = 2
  real(c_double), parameter :: marker = -7.0d0
  type blockfield
    real(c_double), allocatable :: ustar(:,:,:), vstar(:,:,:), wstar(:,:,:)
    real(c_double), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), p(:,:,:), su(:,:,:), dens(:,:,:)
    real(c_double), allocatable :: phi(:,:,:), mu(:,:,:)
  end type
  type(blockfield), allocatable, target :: dom(:)
  type(mf3d_block_desc), allocatable :: desc(:)
  type(mf3d_mg_cfg) :: cfg
  type(c_ptr) :: mg
  type(c_ptr), allocatable, target :: pu(:), pv(:), pw(:), pus(:), pvs(:), pws(:), pp(:), psu(:), pd(:)
  type(c_ptr), allocatable, target :: pphi(:), pmu(:)
  integer :: nbx, nby, nbz, lsmf, maxcy, levels, bcs(6), a, b, c, ib, nb, u, rc, n(3)
  integer, allocatable :: ex(:), ey(:), ez(:)
  real(c_double) :: eps, dt, alfapr, h(3), rmax, densl, densg, mul, mug
  real(c_double), allocatable :: hist(:)
  integer(c_int32_t) :: iters
  integer(8) :: changed
  character(len=1024) :: ctl, fin, fout

  call get_command_argument(1, ctl)
  open(newunit=u, file=trim(ctl), status='old')
  read(u, *) nbx, nby, nbz, levels, lsmf, maxcy
  read(u, *) eps, dt, alfapr
  if (lsmf == 2) read(u, *) densl, densg, mul, mug
  read(u, *) bcs
  allocate(ex(nbx), ey(nby), ez(nbz))
  read(u, *) ex
  read(u, *) ey
  read(u, *) ez
  read(u, *) h
  read(u, '(a)') fin
  read(u, '(a)') fout
  close(u)

  nb = nbx * nby * nbz
  allocate(dom(nb), desc(nb), hist(maxcy + 1))
  allocate(pu(nb), pv(nb), pw(nb), pus(nb), pvs(nb), pws(nb), pp(nb), psu(nb), pd(nb), pphi(nb), pmu(nb))
  do c = 0, nbz - 1
    do b = 0, nby - 1
      do a = 0, nbx - 1
        ib = bid(a, b, c) + 1
        n = [ex(a + 1), ey(b + 1), ez(c + 1)]
        call alloc3(dom(ib)%ustar, n); call alloc3(dom(ib)%vstar, n); call alloc3(dom(ib)%wstar, n)
        call alloc3(dom(ib)%u, n); call alloc3(dom(ib)%v, n); call alloc3(dom(ib)%w, n)
        call alloc3(dom(ib)%p, n); call alloc3(dom(ib)%su, n); call alloc3(dom(ib)%dens, n)
        call alloc3(dom(ib)%phi, n); call alloc3(dom(ib)%mu, n)
        desc(ib)%n = n
        desc(ib)%ghost = pl
        desc(ib)%ld = n + 2*pl
        desc(ib)%h = h
        desc(ib)%nbr = [bid(a-1,b,c), bid(a+1,b,c), bid(a,b-1,c), bid(a,b+1,c), bid(a,b,c-1), bid(a,b,c+1)]
        desc(ib)%per = -1
        if (bcs(1) == 5 .and. a == 0 .and. nbx > 1) desc(ib)%per(1) = bid(nbx-1, b, c)
        if (bcs(2) == 5 .and. a == nbx-1 .and. nbx > 1) desc(ib)%per(2) = bid(0, b, c)
        if (bcs(3) == 5 .and. b == 0 .and. nby > 1) desc(ib)%per(3) = bid(a, nby-1, c)
        if (bcs(4) == 5 .and. b == nby-1 .and. nby > 1) desc(ib)%per(4) = bid(a, 0, c)
        if (bcs(5) == 5 .and. c == 0 .and. nbz > 1) desc(ib)%per(5) = bid(a, b, nbz-1)
        if (bcs(6) == 5 .and. c == nbz-1 .and. nbz > 1) desc(ib)%per(6) = bid(a, b, 0)
        desc(ib)%bc = bcs
      end do
    end do
  end do

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  do ib = 1, nb
    call read_box(u, dom(ib)%ustar)
    call read_box(u, dom(ib)%vstar)
    call read_box(u, dom(ib)%wstar)
    call read_box(u, dom(ib)%p)
    if (lsmf == 1) call read_box(u, dom(ib)%dens)
    if (lsmf == 2) call read_box(u, dom(ib)%phi)
    ! u, v, w start as the predicted velocities, as in the reference's step
    dom(ib)%u = dom(ib)%ustar
    dom(ib)%v = dom(ib)%vstar
    dom(ib)%w = dom(ib)%wstar
  end do
  close(u)

  cfg%nblocks = nb
  cfg%levels = levels
  cfg%smoother = 3
  cfg%nu_pre = 2
  cfg%nu_post = 2
  cfg%maxcy = maxcy
  cfg%eps = eps
  cfg%min_cycles = 2
  cfg%precision = MF3D_FP64
  cfg%variable_coef = merge(1, 0, lsmf >= 1)
  cfg%device = 0
  rc = mf3d_mg_create(cfg, desc, mg)
  if (rc /= MF3D_OK) then
    print *, 'mf3d_mg_create failed ', rc, ' ', mf3d_mg_error(mg)
    stop 1
  end if
  do ib = 1, nb
    pus(ib) = c_loc(dom(ib)%ustar); pvs(ib) = c_loc(dom(ib)%vstar); pws(ib) = c_loc(dom(ib)%wstar)
    pu(ib) = c_loc(dom(ib)%u); pv(ib) = c_loc(dom(ib)%v); pw(ib) = c_loc(dom(ib)%w)
    pp(ib) = c_loc(dom(ib)%p); psu(ib) = c_loc(dom(ib)%su); pd(ib) = c_loc(dom(ib)%dens)
    pphi(ib) = c_loc(dom(ib)%phi); pmu(ib) = c_loc(dom(ib)%mu)
  end do
  if (lsmf == 2) then   ! flosol.for:116 heaviside, then newsolv_mg.for:12-14 coeff
    rc = mf3d_mg_heaviside(mg, pphi, densl, densg, mul, mug, c_loc(pd), c_loc(pmu))
    if (rc /= MF3D_OK) then
      print *, 'heaviside failed ', rc, ' ', mf3d_mg_error(mg)
      stop 1
    end if
  end if
  if (lsmf == 1) then   ! newsolv_mg.for:12-14: coeff every call under LSM
    rc = mf3d_mg_set_density(mg, pd)
    if (rc /= MF3D_OK) then
      print *, 'coeff failed ', rc, ' ', mf3d_mg_error(mg)
      stop 1
    end if
  end if
  rc = mf3d_mg_newsolv(mg, pus, pvs, pws, pu, pv, pw, pp, dt, alfapr, eps, c_loc(psu), rmax, iters, hist)
  if (rc == MF3D_NOT_CONVERGED) then
    print *, 'not converged!! ', maxcy, rmax           ! newsolv_mg.for:60
  else if (rc == MF3D_DIVERGED) then
    print *, 'BIG RMAX!! STOP!!!!!!!!', rmax            ! newsolv_mg.for:66
  else if (rc /= MF3D_OK) then
    print *, 'newsolv failed ', rc, ' ', mf3d_mg_error(mg)
    stop 1
  end if
  print '(a,i4,a,es12.5)', ' iter', iters, ' rmax:', rmax
  call mf3d_mg_destroy(mg)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) dble(iters)
  write(u) hist(1:iters + 1)
  do ib = 1, nb
    call write_box(u, dom(ib)%p, pl)
    call write_box(u, dom(ib)%u, pl)
    call write_box(u, dom(ib)%v, pl)
    call write_box(u, dom(ib)%w, pl)
    call write_box(u, dom(ib)%su, 0)
    changed = 0
    if (lsmf == 2) then
      call write_box(u, dom(ib)%dens, 1)
      call write_box(u, dom(ib)%mu, 1)
      changed = changed + outer_changed(dom(ib)%dens) + outer_changed(dom(ib)%mu)
    end if
    write(u) dble(changed)
  end do
  close(u)

contains

  integer function bid(a, b, c)
    integer, intent(in) :: a, b, c
    if (a < 0 .or. a >= nbx .or. b < 0 .or. b >= nby .or. c < 0 .or. c >= nbz) then
      bid = -1
    else
      bid = a + nbx * (b + nby * c)
    end if
  end function

  subroutine alloc3(x, n)
    real(c_double), allocatable, intent(inout) :: x(:,:,:)
    integer, intent(in) :: n(3)
    allocate(x(n(1) + 2*pl, n(2) + 2*pl, n(3) + 2*pl))
    x = marker
  end subroutine

  ! the one-ghost-layer box (indices pl .. n+pl+1) of a pl-ghost array
  subroutine read_box(unit, x)
    integer, intent(in) :: unit
    real(c_double), intent(inout) :: x(:,:,:)
    integer :: m(3)
    m = shape(x)
    read(unit) x(pl:m(1)-pl+1, pl:m(2)-pl+1, pl:m(3)-pl+1)
  end subroutine

  subroutine write_box(unit, x, ring)
    integer, intent(in) :: unit, ring
    real(c_double), intent(in) :: x(:,:,:)
    integer :: m(3)
    m = shape(x)
    write(unit) x(pl+1-ring:m(1)-pl+ring, pl+1-ring:m(2)-pl+ring, pl+1-ring:m(3)-pl+ring)
  end subroutine

  ! entries of the outermost ghost layer that no longer hold the marker
  integer(8) function outer_changed(x)
    real(c_double), intent(in) :: x(:,:,:)
    integer :: i, j, k, m(3)
    m = shape(x)
    outer_changed = 0
    do k = 1, m(3)
      do j = 1, m(2)
        do i = 1, m(1)
          if (i == 1 .or. j == 1 .or. k == 1 .or. i == m(1) .or. j == m(2) .or. k == m(3)) then
            if (x(i, j, k) /= marker) outer_changed = outer_changed + 1
          end if
        end do
      end do
    end do
  end function

end program mf3d_host_vel
