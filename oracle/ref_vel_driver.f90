! ref_vel_driver.f90 -- harness that runs the reference's newsolv_mg loop in
! its own velocity form: calvel, calmas, mgkcyc, the pressure update,
! exchange(4) and pbound, all compiled unchanged from the reference sources
! (press.for, exchange.for, exchange_bc.for, mgsolver.for, exchangepp.for),
! in the order of newsolv_mg.for:12-58.  boundu/v/w, which newsolv_mg calls
! after the loop, and the ppref reduction (commented out there) are left out.
!
! TEST INFRASTRUCTURE ONLY: it produces golden vectors in the build
! container.  All blocks live on one MPI rank.
!
! lsm = 2 runs heaviside (lsm.for:1456-1509; bound_lsm(18/19) and
! exchange(18/19) after it, bounds_lsm.for:2-104) on a level set phi before
! newsolv_mg, as the reference's time step does (flosol.for:114-117), and
! writes dens and mu as well.  lsm = 3 only fills the ghosts of a given
! cell field through bound_lsm(19) (the viscosity's path) and writes it.
! lsm = 4 runs like lsm = 0, then writes the reference's checkpoint with
! its own tecbin (post.for:496-548) -- tecbin####.bin in the working
! directory, with the mean and viscosity fields filled from the solution --
! and final_ctime.dat in the reference's format (flosol.for:296-299), and
! appends to the output the x, y, z and 16 field arrays it checkpointed.
!
! Usage: mg_ref_vel <control file>
!   line 1: nbx nby nbz levels lsm maxcy nu_pre nu_post
!   line 2: eps (absolute: rmax = max|div u| dV, newsolv_mg.for:56),
!           mg_itrsch (1 lexicographic GS, 2 ADI-TDMA, 3 red-black GS),
!           pl_ex, LMR and differencing (control.cin; the ghost width is
!           pl = LMR + pl_ex, localparameters.for:17-24)
!   line 3: bc_w bc_e bc_s bc_n bc_b bc_t
!   lines 4-6: block extents along x, y, z
!   line 7: dx dy dz
!   line 8: dt alfapr
!   (lsm = 2 only) densl densg mul mug
!   line 9: input file: per block, full ghosted arrays (pl layers) of ustar,
!           vstar, wstar, p and dens (lsm = 1) or phi (lsm = 2), in that
!           order; lsm = 3: per block the field alone
!   line 10: output file
! Output (f64 stream): iterations, rmax(0:maxcy), then per block the full
! ghosted arrays of p, u, v, w and su (sup under lsm), then (lsm = 2) dens
! and mu.  lsm = 3: per block the field after bound_lsm(19).
program mg_ref_vel_driver
  use vars
  use multidata
  use mpi
  implicit none
  integer :: nbx, nby, nbz, lsmf, ia, ib2, ic2, id, ib, i, j, k, l
  integer :: bcs(6), nn(3), it, nits
  integer, allocatable :: ex(:), ey(:), ez(:)
  double precision :: gdx, gdy, gdz
  double precision, allocatable :: rmh(:)
  character(len=1024) :: ctl, fin, fout
  integer :: mi, mj, mk, nwork, tot, u, uu, ierr2

  call MPI_INIT(ierr2)
  MPI_FLT = MPI_DOUBLE_PRECISION
  call get_command_argument(1, ctl)
  open(newunit=u, file=trim(ctl), status='old')
  read(u, *) nbx, nby, nbz, ngrid_input, lsmf, maxcy, irestr, iproln
  read(u, *) eps, mg_itrsch, pl_ex, LMR, differencing
  read(u, *) bcs
  allocate(ex(nbx), ey(nby), ez(nbz))
  read(u, *) ex
  read(u, *) ey
  read(u, *) ez
  read(u, *) gdx, gdy, gdz
  read(u, *) dt, alfapr
  if (lsmf == 2) read(u, *) densl, densg, mul, mug
  read(u, '(a)') fin
  read(u, '(a)') fout
  close(u)

  num_domains = nbx * nby * nbz
  nbp = num_domains
  ngrd_gl = ngrid_input
  L_LSM = (lsmf == 1 .or. lsmf == 2)
  ntime = 0; ctime = 0.0; forcn = 0.0; qstpn = 0.0; count = 0; ntav1_count = 0.0; ntav2_count = 0.0
  pl = LMR + pl_ex
  rdivmax = 1
  g_dx = gdx; g_dy = gdy; g_dz = gdz
  allocate(dom(nbp), dom_id(nbp), dom_ad(0:num_domains-1), dom_indid(0:num_domains-1))
  allocate(rdiv(0:num_domains-1), rdv(0:num_domains-1, ngrd_gl))
  rdiv = 1
  rdv = 1
  PERIODIC = any(bcs == 5)
  if (lsmf == 2 .and. PERIODIC) then
    ! exchange_bc.for:105,422 select no array for op 18 in the pack loops
    print *, 'heaviside with periodic faces is not run: exchange_bc(18) has no array'
    stop 2
  end if
  ngg = 0

  do ic2 = 0, nbz - 1
  do ib2 = 0, nby - 1
  do ia = 0, nbx - 1
    id = bid(ia, ib2, ic2)
    ib = id + 1
    dom_id(ib) = id
    dom_ad(id) = 0
    dom_indid(id) = ib
    nn = (/ ex(ia+1), ey(ib2+1), ez(ic2+1) /)
    dom(ib)%isp = 1 + pl; dom(ib)%iep = nn(1) + pl
    dom(ib)%jsp = 1 + pl; dom(ib)%jep = nn(2) + pl
    dom(ib)%ksp = 1 + pl; dom(ib)%kep = nn(3) + pl
    dom(ib)%ttc_i = nn(1) + 2*pl; dom(ib)%ttc_j = nn(2) + 2*pl; dom(ib)%ttc_k = nn(3) + 2*pl
    dom(ib)%ngrid = ngrd_gl
    dom(ib)%dx = gdx; dom(ib)%dy = gdy; dom(ib)%dz = gdz
    dom(ib)%bc_west = bcs(1); dom(ib)%bc_east = bcs(2)
    dom(ib)%bc_south = bcs(3); dom(ib)%bc_north = bcs(4)
    dom(ib)%bc_bottom = bcs(5); dom(ib)%bc_top = bcs(6)
    dom(ib)%iprev = bid(ia-1, ib2, ic2); dom(ib)%inext = bid(ia+1, ib2, ic2)
    dom(ib)%jprev = bid(ia, ib2-1, ic2); dom(ib)%jnext = bid(ia, ib2+1, ic2)
    dom(ib)%kprev = bid(ia, ib2, ic2-1); dom(ib)%knext = bid(ia, ib2, ic2+1)
    dom(ib)%corprev1 = bid(ia-1, ib2-1, ic2-1); dom(ib)%corprev2 = bid(ia-1, ib2+1, ic2-1)
    dom(ib)%corprev3 = bid(ia+1, ib2+1, ic2-1); dom(ib)%corprev4 = bid(ia+1, ib2-1, ic2-1)
    dom(ib)%cornext1 = bid(ia+1, ib2+1, ic2+1); dom(ib)%cornext2 = bid(ia+1, ib2-1, ic2+1)
    dom(ib)%cornext3 = bid(ia-1, ib2-1, ic2+1); dom(ib)%cornext4 = bid(ia-1, ib2+1, ic2+1)
    dom(ib)%edgprev1 = bid(ia-1, ib2, ic2-1); dom(ib)%edgprev2 = bid(ia, ib2+1, ic2-1)
    dom(ib)%edgprev3 = bid(ia+1, ib2, ic2-1); dom(ib)%edgprev4 = bid(ia, ib2-1, ic2-1)
    dom(ib)%edgprev5 = bid(ia-1, ib2-1, ic2); dom(ib)%edgprev6 = bid(ia-1, ib2+1, ic2)
    dom(ib)%edgnext1 = bid(ia+1, ib2, ic2+1); dom(ib)%edgnext2 = bid(ia, ib2-1, ic2+1)
    dom(ib)%edgnext3 = bid(ia-1, ib2, ic2+1); dom(ib)%edgnext4 = bid(ia, ib2+1, ic2+1)
    dom(ib)%edgnext5 = bid(ia+1, ib2+1, ic2); dom(ib)%edgnext6 = bid(ia+1, ib2-1, ic2)
    dom(ib)%per_ip = -1; dom(ib)%per_in = -1; dom(ib)%per_jp = -1
    dom(ib)%per_jn = -1; dom(ib)%per_kp = -1; dom(ib)%per_kn = -1
    if (bcs(1) == 5 .and. ia == 0 .and. nbx > 1) dom(ib)%per_ip = bid(nbx-1, ib2, ic2)
    if (bcs(2) == 5 .and. ia == nbx-1 .and. nbx > 1) dom(ib)%per_in = bid(0, ib2, ic2)
    if (bcs(3) == 5 .and. ib2 == 0 .and. nby > 1) dom(ib)%per_jp = bid(ia, nby-1, ic2)
    if (bcs(4) == 5 .and. ib2 == nby-1 .and. nby > 1) dom(ib)%per_jn = bid(ia, 0, ic2)
    if (bcs(5) == 5 .and. ic2 == 0 .and. nbz > 1) dom(ib)%per_kp = bid(ia, ib2, nbz-1)
    if (bcs(6) == 5 .and. ic2 == nbz-1 .and. nbz > 1) dom(ib)%per_kn = bid(ia, ib2, 0)
    ! velocity ranges (localparameters.for:104-131,185-190)
    dom(ib)%isu = dom(ib)%isp; dom(ib)%ieu = dom(ib)%iep
    dom(ib)%jsu = dom(ib)%jsp; dom(ib)%jeu = dom(ib)%jep
    dom(ib)%ksu = dom(ib)%ksp; dom(ib)%keu = dom(ib)%kep
    dom(ib)%isv = dom(ib)%isp; dom(ib)%iev = dom(ib)%iep
    dom(ib)%jsv = dom(ib)%jsp; dom(ib)%jev = dom(ib)%jep
    dom(ib)%ksv = dom(ib)%ksp; dom(ib)%kev = dom(ib)%kep
    dom(ib)%isw = dom(ib)%isp; dom(ib)%iew = dom(ib)%iep
    dom(ib)%jsw = dom(ib)%jsp; dom(ib)%jew = dom(ib)%jep
    dom(ib)%ksw = dom(ib)%ksp; dom(ib)%kew = dom(ib)%kep
    if (dom(ib)%inext < 0 .and. dom(ib)%bc_east /= 5) dom(ib)%ieu = dom(ib)%ieu - 1
    if (dom(ib)%jnext < 0 .and. dom(ib)%bc_north /= 5) dom(ib)%jev = dom(ib)%jev - 1
    if (dom(ib)%knext < 0 .and. dom(ib)%bc_top /= 5) dom(ib)%kew = dom(ib)%kew - 1
    ngg = max(ngg, dom(ib)%ttc_i*dom(ib)%ttc_j, dom(ib)%ttc_i*dom(ib)%ttc_k, &
              dom(ib)%ttc_j*dom(ib)%ttc_k)
  end do
  end do
  end do
  nge = maxval((/ maxval(ex), maxval(ey), maxval(ez) /) + 2*pl) * (pl+1)**2
  ngc = (pl+1)**3

  do ib = 1, nbp
    mi = dom(ib)%ttc_i; mj = dom(ib)%ttc_j; mk = dom(ib)%ttc_k
    allocate(dom(ib)%p(mi,mj,mk), dom(ib)%pp(mi,mj,mk), dom(ib)%su(mi,mj,mk), &
             dom(ib)%sup(mi,mj,mk), dom(ib)%u(mi,mj,mk), dom(ib)%v(mi,mj,mk), &
             dom(ib)%w(mi,mj,mk), dom(ib)%ustar(mi,mj,mk), dom(ib)%vstar(mi,mj,mk), &
             dom(ib)%wstar(mi,mj,mk))
    dom(ib)%p = 0.0; dom(ib)%pp = 0.0; dom(ib)%su = 0.0; dom(ib)%sup = 0.0
    allocate(dom(ib)%sendb_m1(ngg), dom(ib)%sendb_p1(ngg), dom(ib)%recvb_m1(ngg), &
             dom(ib)%recvb_p1(ngg), dom(ib)%sendb_m2(ngg), dom(ib)%sendb_p2(ngg), &
             dom(ib)%recvb_m2(ngg), dom(ib)%recvb_p2(ngg), dom(ib)%sendb_m3(ngg), &
             dom(ib)%sendb_p3(ngg), dom(ib)%recvb_m3(ngg), dom(ib)%recvb_p3(ngg))
    allocate(dom(ib)%sc1m(ngc), dom(ib)%sc1p(ngc), dom(ib)%rc1m(ngc), dom(ib)%rc1p(ngc), &
             dom(ib)%sc2m(ngc), dom(ib)%sc2p(ngc), dom(ib)%rc2m(ngc), dom(ib)%rc2p(ngc), &
             dom(ib)%sc3m(ngc), dom(ib)%sc3p(ngc), dom(ib)%rc3m(ngc), dom(ib)%rc3p(ngc), &
             dom(ib)%sc4m(ngc), dom(ib)%sc4p(ngc), dom(ib)%rc4m(ngc), dom(ib)%rc4p(ngc))
    allocate(dom(ib)%se1m(nge), dom(ib)%se1p(nge), dom(ib)%re1m(nge), dom(ib)%re1p(nge), &
             dom(ib)%se2m(nge), dom(ib)%se2p(nge), dom(ib)%re2m(nge), dom(ib)%re2p(nge), &
             dom(ib)%se3m(nge), dom(ib)%se3p(nge), dom(ib)%re3m(nge), dom(ib)%re3p(nge), &
             dom(ib)%se4m(nge), dom(ib)%se4p(nge), dom(ib)%re4m(nge), dom(ib)%re4p(nge), &
             dom(ib)%se5m(nge), dom(ib)%se5p(nge), dom(ib)%re5m(nge), dom(ib)%re5p(nge), &
             dom(ib)%se6m(nge), dom(ib)%se6p(nge), dom(ib)%re6m(nge), dom(ib)%re6p(nge))
    ! message tags and the scalar-exchange staging buffer, as initial.for:192-194,246
    do i = 1, 26
      dom(ib)%tg(i) = i*10**5 + dom_id(ib)
    end do
    allocate(dom(ib)%stfcinf(6,pl,ngg))
    dom(ib)%stfcinf = 0.0
    allocate(dom(ib)%faz(ngrd_gl))
    nwork = 0
    do l = 1, ngrd_gl
      mi = (dom(ib)%iep-dom(ib)%isp+1)/2**(l-1)+2
      mj = (dom(ib)%jep-dom(ib)%jsp+1)/2**(l-1)+2
      mk = (dom(ib)%kep-dom(ib)%ksp+1)/2**(l-1)+2
      nwork = nwork + mi*mj*mk + 8*(mi-2)*(mj-2)*(mk-2)
    end do
    allocate(dom(ib)%cof(nwork))
    dom(ib)%cof = 0.0
    allocate(dom(ib)%ijkp_lsm(0:dom(ib)%ngrid))
    dom(ib)%ijkp_lsm = 0
    dom(ib)%ijkp_lsm(1) = (dom(ib)%ttc_i-2*pl)*(dom(ib)%ttc_j-2*pl)*(dom(ib)%ttc_k-2*pl)
    do l = 2, dom(ib)%ngrid
      mi = (dom(ib)%iep-dom(ib)%isp+1)/2**(l-1)+2
      mj = (dom(ib)%jep-dom(ib)%jsp+1)/2**(l-1)+2
      mk = (dom(ib)%kep-dom(ib)%ksp+1)/2**(l-1)+2
      dom(ib)%ijkp_lsm(l) = dom(ib)%ijkp_lsm(l-1) + (mi-2)*(mj-2)*(mk-2)
    end do
    tot = dom(ib)%ijkp_lsm(dom(ib)%ngrid)
    dom(ib)%tot = tot
    allocate(dom(ib)%dens_mg(max(tot, 1)))
    allocate(dom(ib)%dens(dom(ib)%ttc_i, dom(ib)%ttc_j, dom(ib)%ttc_k))
    dom(ib)%dens = 1.0
    if (lsmf >= 2) then
      allocate(dom(ib)%phi(dom(ib)%ttc_i, dom(ib)%ttc_j, dom(ib)%ttc_k), &
               dom(ib)%h_phi(dom(ib)%ttc_i, dom(ib)%ttc_j, dom(ib)%ttc_k), &
               dom(ib)%mu(dom(ib)%ttc_i, dom(ib)%ttc_j, dom(ib)%ttc_k))
      dom(ib)%h_phi = 0.0
      dom(ib)%mu = 0.0
    end if
  end do

  if (lsmf == 3) then
    open(newunit=uu, file=trim(fin), access='stream', form='unformatted', status='old')
    do ib = 1, nbp
      read(uu) dom(ib)%mu
    end do
    close(uu)
    call bound_lsm(19)
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    do ib = 1, nbp
      write(u) dom(ib)%mu
    end do
    close(u)
    call MPI_FINALIZE(ierr2)
    stop
  end if

  open(newunit=uu, file=trim(fin), access='stream', form='unformatted', status='old')
  do ib = 1, nbp
    read(uu) dom(ib)%ustar
    read(uu) dom(ib)%vstar
    read(uu) dom(ib)%wstar
    read(uu) dom(ib)%p
    if (lsmf == 1) read(uu) dom(ib)%dens
    if (lsmf == 2) read(uu) dom(ib)%phi
    dom(ib)%u = dom(ib)%ustar
    dom(ib)%v = dom(ib)%vstar
    dom(ib)%w = dom(ib)%wstar
  end do
  close(uu)

  allocate(rmh(0:maxcy))
  rmh = 0.0
  if (lsmf == 2) call heaviside
  ! newsolv_mg.for:12-21
  call coeff
  call calvel
  call calmas
  rmh(0) = rmax
  do ib = 1, nbp
    dom(ib)%pp = 0.0
  end do
  nits = maxcy
  ! newsolv_mg.for:23-58
  do it = 1, maxcy
    call mgkcyc
    do ib = 1, nbp
      do k = dom(ib)%ksp-1, dom(ib)%kep+1
        do i = dom(ib)%isp-1, dom(ib)%iep+1
          do j = dom(ib)%jsp-1, dom(ib)%jep+1
            dom(ib)%p(i,j,k) = (dom(ib)%p(i,j,k) + (dom(ib)%pp(i,j,k) - 0.0d0))
            dom(ib)%pp(i,j,k) = 0.0
          end do
        end do
      end do
    end do
    call exchange(4)
    call pbound
    call calvel
    call calmas
    rmh(it) = rmax
    if (rmax .lt. eps .and. it .gt. 1) then
      nits = it
      exit
    end if
  end do

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) dble(nits)
  write(u) rmh
  do ib = 1, nbp
    write(u) dom(ib)%p
    write(u) dom(ib)%u
    write(u) dom(ib)%v
    write(u) dom(ib)%w
    if (L_LSM) then
      write(u) dom(ib)%sup
    else
      write(u) dom(ib)%su
    end if
    if (lsmf == 2) then
      write(u) dom(ib)%dens
      write(u) dom(ib)%mu
    end if
  end do
  if (lsmf == 4) then
    do ib = 1, nbp
      mi = dom(ib)%ttc_i; mj = dom(ib)%ttc_j; mk = dom(ib)%ttc_k
      allocate(dom(ib)%x(mi), dom(ib)%y(mj), dom(ib)%z(mk))
      allocate(dom(ib)%pm(mi,mj,mk), dom(ib)%ppm(mi,mj,mk), dom(ib)%vis(mi,mj,mk), &
               dom(ib)%um(mi,mj,mk), dom(ib)%uum(mi,mj,mk), dom(ib)%vm(mi,mj,mk), &
               dom(ib)%vvm(mi,mj,mk), dom(ib)%wm(mi,mj,mk), dom(ib)%wwm(mi,mj,mk), &
               dom(ib)%uvm(mi,mj,mk), dom(ib)%uwm(mi,mj,mk), dom(ib)%vwm(mi,mj,mk))
      do i = 1, mi
        dom(ib)%x(i) = (i - 1 - pl + 0.5d0) * gdx + 0.25d0 * dom_id(ib)
      end do
      do j = 1, mj
        dom(ib)%y(j) = (j - 1 - pl + 0.5d0) * gdy
      end do
      do k = 1, mk
        dom(ib)%z(k) = (k - 1 - pl + 0.5d0) * gdz
      end do
      dom(ib)%pm = 0.5d0 * dom(ib)%p; dom(ib)%ppm = dom(ib)%p * dom(ib)%p
      dom(ib)%vis = 1.0d-6 + 1.0d-3 * abs(dom(ib)%u)
      dom(ib)%um = 0.5d0 * dom(ib)%u; dom(ib)%uum = dom(ib)%u * dom(ib)%u
      dom(ib)%vm = 0.5d0 * dom(ib)%v; dom(ib)%vvm = dom(ib)%v * dom(ib)%v
      dom(ib)%wm = 0.5d0 * dom(ib)%w; dom(ib)%wwm = dom(ib)%w * dom(ib)%w
      dom(ib)%uvm = dom(ib)%u * dom(ib)%v; dom(ib)%uwm = dom(ib)%u * dom(ib)%w
      dom(ib)%vwm = dom(ib)%v * dom(ib)%w
    end do
    call tecbin(nits)
    ntime = nits; ctime = 0.0123456789d0 * nits; forcn = -1.5d0; qstpn = 2.25d0
    count = 7; ntav1_count = 3.0d0; ntav2_count = 12345.678901d0
    open(newunit=uu, file='final_ctime.dat')
    write(uu, '(i8,3F15.6)') ntime, ctime, forcn, qstpn, count, ntav1_count, ntav2_count
    close(uu)
    do ib = 1, nbp
      write(u) dom(ib)%x
      write(u) dom(ib)%y
      write(u) dom(ib)%z
      write(u) dom(ib)%p
      write(u) dom(ib)%pm
      write(u) dom(ib)%ppm
      write(u) dom(ib)%vis
      write(u) dom(ib)%u
      write(u) dom(ib)%um
      write(u) dom(ib)%uum
      write(u) dom(ib)%v
      write(u) dom(ib)%vm
      write(u) dom(ib)%vvm
      write(u) dom(ib)%w
      write(u) dom(ib)%wm
      write(u) dom(ib)%wwm
      write(u) dom(ib)%uvm
      write(u) dom(ib)%uwm
      write(u) dom(ib)%vwm
    end do
  end if
  close(u)
  call MPI_FINALIZE(ierr2)

contains

  integer function bid(a, b, c)
    integer, intent(in) :: a, b, c
    if (a < 0 .or. a >= nbx .or. b < 0 .or. b >= nby .or. c < 0 .or. c >= nbz) then
      bid = -1
    else
      bid = a + nbx * (b + nby * c)
    end if
  end function

end program
