! mf3d_host.f90 -- a minimal Fortran host for the gfx950 pressure multigrid.
!
! It holds the pressure fields the way MultiFlow3D does -- per block,
! column-major arrays with pl ghost layers (initial.for:199-204) -- and
! replaces the solver call of newsolv_mg (newsolv_mg.for:23-58) with one
! mf3d_mg_solve through the ISO_C_BINDING module, passing each block's
! arrays by address.  Used to check the Fortran boundary end to end.
!
! Usage: mf3d_host <control file>   (the same control format as the
! oracle's reference driver: grid, levels, cycles, tolerance, boundary
! codes, extents, spacing, RHS file, density file or '-', output file)
! Output (f64 stream): ncycles, rho(0:ncycles), rmax, p interiors.
program mf3d_host
  use iso_c_binding
  use mf3d_mg_f
  implicit none
  integer, parameter :: pl = 2
  type blockfield
    real(c_double), allocatable :: su(:,:,:), p(:,:,:), dens(:,:,:)
  end type
  type(blockfield), allocatable, target :: dom(:)
  type(mf3d_block_desc), allocatable :: desc(:)
  type(mf3d_mg_cfg) :: cfg
  type(c_ptr) :: mg
  type(c_ptr), allocatable :: fptr(:), pptr(:), dptr(:)
  integer :: nbx, nby, nbz, lsmf, maxcy, mincy, nupre, nupost, levels
  integer :: bcs(6), a, b, c, ib, nb, u, rc, n(3)
  integer, allocatable :: ex(:), ey(:), ez(:)
  real(c_double) :: tol, h(3), rmax
  real(c_double), allocatable :: hist(:)
  integer(c_int32_t) :: iters
  character(len=1024) :: ctl, frhs, fdens, fout

  call get_command_argument(1, ctl)
  open(newunit=u, file=trim(ctl), status='old')
  read(u, *) nbx, nby, nbz, levels, lsmf, maxcy, mincy, nupre, nupost
  read(u, *) tol
  read(u, *) bcs
  allocate(ex(nbx), ey(nby), ez(nbz))
  read(u, *) ex
  read(u, *) ey
  read(u, *) ez
  read(u, *) h
  read(u, '(a)') frhs
  read(u, '(a)') fdens
  read(u, '(a)') fout
  close(u)

  nb = nbx * nby * nbz
  allocate(dom(nb), desc(nb), fptr(nb), pptr(nb), dptr(nb), hist(maxcy + 1))
  do c = 0, nbz - 1
    do b = 0, nby - 1
      do a = 0, nbx - 1
        ib = bid(a, b, c) + 1
        n = [ex(a + 1), ey(b + 1), ez(c + 1)]
        allocate(dom(ib)%su(n(1) + 2*pl, n(2) + 2*pl, n(3) + 2*pl))
        allocate(dom(ib)%p(n(1) + 2*pl, n(2) + 2*pl, n(3) + 2*pl))
        allocate(dom(ib)%dens(n(1) + 2*pl, n(2) + 2*pl, n(3) + 2*pl))
        dom(ib)%su = 0.0d0
        dom(ib)%p = -7.0d0          ! ghosts keep whatever the host had
        dom(ib)%dens = 1.0d0
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

  call read_interiors(trim(frhs), 1)
  if (lsmf == 1) call read_interiors(trim(fdens), 2)

  cfg%nblocks = nb
  cfg%levels = levels
  cfg%smoother = 3
  cfg%nu_pre = nupre
  cfg%nu_post = nupost
  cfg%maxcy = maxcy
  cfg%eps = tol
  cfg%min_cycles = mincy
  cfg%precision = MF3D_FP64
  cfg%variable_coef = lsmf
  cfg%device = 0
  rc = mf3d_mg_create(cfg, desc, mg)
  if (rc /= MF3D_OK) then
    print *, 'mf3d_mg_create failed ', rc, ' ', mf3d_mg_error(mg)
    stop 1
  end if
  do ib = 1, nb
    fptr(ib) = c_loc(dom(ib)%su)
    pptr(ib) = c_loc(dom(ib)%p)
    dptr(ib) = c_loc(dom(ib)%dens)
  end do
  if (lsmf == 1) then
    rc = mf3d_mg_set_density(mg, dptr)
    if (rc /= MF3D_OK) then
      print *, 'coeff failed ', rc, ' ', mf3d_mg_error(mg)
      stop 1
    end if
  end if
  rc = mf3d_mg_solve(mg, fptr, pptr, rmax, iters, hist)
  if (rc == MF3D_NOT_CONVERGED) then
    print *, 'not converged!! ', maxcy, rmax           ! newsolv_mg.for:60
  else if (rc /= MF3D_OK) then
    print *, 'solve failed ', rc, ' ', mf3d_mg_error(mg)
    stop 1
  end if
  print '(a,i4,a,es12.5)', ' iter', iters, ' rmax:', rmax
  call mf3d_mg_destroy(mg)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) dble(iters)
  write(u) hist(1:iters + 1)
  write(u) rmax
  do ib = 1, nb
    n = desc(ib)%n
    write(u) dom(ib)%p(pl+1:pl+n(1), pl+1:pl+n(2), pl+1:pl+n(3))
  end do
  close(u)

contains

  integer function bid(ia, ja, ka)
    integer, intent(in) :: ia, ja, ka
    if (ia < 0 .or. ia >= nbx .or. ja < 0 .or. ja >= nby .or. ka < 0 .or. ka >= nbz) then
      bid = -1
    else
      bid = ia + nbx * (ja + nby * ka)
    end if
  end function

  subroutine read_interiors(fname, which)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: which
    integer :: uu, jb, m(3)
    open(newunit=uu, file=fname, access='stream', form='unformatted', status='old')
    do jb = 1, nb
      m = desc(jb)%n
      if (which == 1) then
        read(uu) dom(jb)%su(pl+1:pl+m(1), pl+1:pl+m(2), pl+1:pl+m(3))
      else
        read(uu) dom(jb)%dens(pl+1:pl+m(1), pl+1:pl+m(2), pl+1:pl+m(3))
      end if
    end do
    close(uu)
  end subroutine

end program mf3d_host
