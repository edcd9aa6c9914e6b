! ref_driver.f90 -- harness that runs the reference's own multigrid
! (coeff, mgkcyc and the exchangepp halo code, compiled unchanged from the
! reference sources) on a hand-built block grid, in the Poisson form of the
! newsolv_mg outer loop, and writes residual histories and pressure fields.
!
! TEST INFRASTRUCTURE ONLY: it produces golden vectors in the build
! container.  All blocks live on one MPI rank, so exchangepp takes its
! same-rank branches and no MPI routine is ever called.
!
! Usage: mg_ref <control file>
!   line 1: nbx nby nbz levels lsm maxcy min_cycles nu_pre nu_post
!   line 2: tol mg_itrsch (1 lexicographic GS, 2 ADI-TDMA, 3 red-black GS)
!   line 3: bc_w bc_e bc_s bc_n bc_b bc_t
!   line 4: x extents of the nbx block columns
!   line 5: y extents (nby values)
!   line 6: z extents (nbz values)
!   line 7: dx dy dz
!   line 8: rhs file  (f64, per block interior, i fastest; blocks x fastest)
!   line 9: density file or '-'   (same layout)
!   line 10: output file
!   line 11 (optional): rdiv of every block (blocks x fastest); a block's
!            extents are its column's extents x rdiv, its spacing dx / rdiv
! Output (f64 stream): ncycles, rho(0:maxcy), rmax(0:maxcy), then per block
! the interior of p, then per block pp after cycle 1 including one ghost
! layer, then the seconds the solve took (the loop, without input reading).
program mg_ref_driver
  use vars
  use multidata
  use mpi
  implicit none
  type fieldp
    double precision, pointer :: a(:,:,:)
  end type
  integer :: nbx, nby, nbz, lsmf, mincy, ia, ib2, ic2, id, ib, i, j, k, l
  integer :: bcs(6), nn(3), it, ncyc
  integer, allocatable :: ex(:), ey(:), ez(:)
  double precision :: tol, gdx, gdy, gdz, nf, nr, rm
  double precision, allocatable :: rho(:), rmx(:)
  character(len=1024) :: ctl, frhs, fdens, fout
  type(fieldp), allocatable :: f(:), pp1(:)
  integer :: mi, mj, mk, nwork, pre, tot, u, ios, maxr, gl, r
  integer(8) :: clk0, clk1, clkrate
  integer, allocatable :: rdivs(:)

  call get_command_argument(1, ctl)
  open(newunit=u, file=trim(ctl), status='old')
  read(u, *) nbx, nby, nbz, ngrid_input, lsmf, maxcy, mincy, irestr, iproln
  read(u, *) tol, mg_itrsch
  read(u, *) bcs
  allocate(ex(nbx), ey(nby), ez(nbz))
  read(u, *) ex
  read(u, *) ey
  read(u, *) ez
  read(u, *) gdx, gdy, gdz
  read(u, '(a)') frhs
  read(u, '(a)') fdens
  read(u, '(a)') fout
  num_domains = nbx * nby * nbz
  allocate(rdivs(num_domains))
  read(u, *, iostat=ios) rdivs
  if (ios /= 0) rdivs = 1
  close(u)

  nbp = num_domains
  ! global level count and per-level ratios as localparameters.for:26-54
  ! computes them (ngrd_gl = ngrid_input + log2(max rdiv))
  maxr = maxval(rdivs)
  ngrd_gl = ngrid_input + nint(log(real(maxr)) / log(2.0))
  L_LSM = (lsmf == 1)
  pl = 1
  g_dx = gdx; g_dy = gdy; g_dz = gdz
  allocate(dom(nbp), dom_id(nbp), dom_ad(0:num_domains-1), dom_indid(0:num_domains-1))
  allocate(rdiv(0:num_domains-1), rdv(0:num_domains-1, ngrd_gl))
  do id = 0, num_domains - 1
    rdiv(id) = rdivs(id + 1)
  end do
  rdv = 0
  PERIODIC = any(bcs == 5)
  ngg = 0

  do ic2 = 0, nbz - 1
  do ib2 = 0, nby - 1
  do ia = 0, nbx - 1
    id = bid(ia, ib2, ic2)
    ib = id + 1
    dom_id(ib) = id
    dom_ad(id) = 0
    dom_indid(id) = ib
    r = rdiv(id)
    nn = (/ ex(ia+1), ey(ib2+1), ez(ic2+1) /) * r
    dom(ib)%isp = 1 + pl; dom(ib)%iep = nn(1) + pl
    dom(ib)%jsp = 1 + pl; dom(ib)%jep = nn(2) + pl
    dom(ib)%ksp = 1 + pl; dom(ib)%kep = nn(3) + pl
    dom(ib)%ttc_i = nn(1) + 2*pl; dom(ib)%ttc_j = nn(2) + 2*pl; dom(ib)%ttc_k = nn(3) + 2*pl
    dom(ib)%ngrid = ngrid_input + nint(log(real(r)) / log(2.0))
    do l = 1, ngrd_gl
      if (l > dom(ib)%ngrid) then
        rdv(id, l) = 1
      else if (l <= ngrid_input) then
        rdv(id, l) = r
      else
        rdv(id, l) = rdv(id, l-1) / 2
      end if
    end do
    ! localparameters.for:255-257
    dom(ib)%dx = gdx / real(r); dom(ib)%dy = gdy / real(r); dom(ib)%dz = gdz / real(r)
    dom(ib)%bc_west = bcs(1); dom(ib)%bc_east = bcs(2)
    dom(ib)%bc_south = bcs(3); dom(ib)%bc_north = bcs(4)
    dom(ib)%bc_bottom = bcs(5); dom(ib)%bc_top = bcs(6)
    ! face neighbours (alloc_dom.for:456-616)
    dom(ib)%iprev = bid(ia-1, ib2, ic2); dom(ib)%inext = bid(ia+1, ib2, ic2)
    dom(ib)%jprev = bid(ia, ib2-1, ic2); dom(ib)%jnext = bid(ia, ib2+1, ic2)
    dom(ib)%kprev = bid(ia, ib2, ic2-1); dom(ib)%knext = bid(ia, ib2, ic2+1)
    ! corners (alloc_dom.for:619-750): offsets as matched there
    dom(ib)%corprev1 = bid(ia-1, ib2-1, ic2-1); dom(ib)%corprev2 = bid(ia-1, ib2+1, ic2-1)
    dom(ib)%corprev3 = bid(ia+1, ib2+1, ic2-1); dom(ib)%corprev4 = bid(ia+1, ib2-1, ic2-1)
    dom(ib)%cornext1 = bid(ia+1, ib2+1, ic2+1); dom(ib)%cornext2 = bid(ia+1, ib2-1, ic2+1)
    dom(ib)%cornext3 = bid(ia-1, ib2-1, ic2+1); dom(ib)%cornext4 = bid(ia-1, ib2+1, ic2+1)
    ! edges (alloc_dom.for:752-968)
    dom(ib)%edgprev1 = bid(ia-1, ib2, ic2-1); dom(ib)%edgprev2 = bid(ia, ib2+1, ic2-1)
    dom(ib)%edgprev3 = bid(ia+1, ib2, ic2-1); dom(ib)%edgprev4 = bid(ia, ib2-1, ic2-1)
    dom(ib)%edgprev5 = bid(ia-1, ib2-1, ic2); dom(ib)%edgprev6 = bid(ia-1, ib2+1, ic2)
    dom(ib)%edgnext1 = bid(ia+1, ib2, ic2+1); dom(ib)%edgnext2 = bid(ia, ib2-1, ic2+1)
    dom(ib)%edgnext3 = bid(ia-1, ib2, ic2+1); dom(ib)%edgnext4 = bid(ia, ib2+1, ic2+1)
    dom(ib)%edgnext5 = bid(ia+1, ib2+1, ic2); dom(ib)%edgnext6 = bid(ia+1, ib2-1, ic2)
    ! periodic partners (alloc_dom.for:972-1150); a block is never its own partner
    dom(ib)%per_ip = -1; dom(ib)%per_in = -1; dom(ib)%per_jp = -1
    dom(ib)%per_jn = -1; dom(ib)%per_kp = -1; dom(ib)%per_kn = -1
    if (bcs(1) == 5 .and. ia == 0 .and. nbx > 1) dom(ib)%per_ip = bid(nbx-1, ib2, ic2)
    if (bcs(2) == 5 .and. ia == nbx-1 .and. nbx > 1) dom(ib)%per_in = bid(0, ib2, ic2)
    if (bcs(3) == 5 .and. ib2 == 0 .and. nby > 1) dom(ib)%per_jp = bid(ia, nby-1, ic2)
    if (bcs(4) == 5 .and. ib2 == nby-1 .and. nby > 1) dom(ib)%per_jn = bid(ia, 0, ic2)
    if (bcs(5) == 5 .and. ic2 == 0 .and. nbz > 1) dom(ib)%per_kp = bid(ia, ib2, nbz-1)
    if (bcs(6) == 5 .and. ic2 == nbz-1 .and. nbz > 1) dom(ib)%per_kn = bid(ia, ib2, 0)
    ngg = max(ngg, dom(ib)%ttc_i*dom(ib)%ttc_j, dom(ib)%ttc_i*dom(ib)%ttc_k, &
              dom(ib)%ttc_j*dom(ib)%ttc_k)
  end do
  end do
  end do
  nge = (maxval((/ maxval(ex), maxval(ey), maxval(ez) /)) * maxr + 2*pl) * (pl+1)**2
  ngc = (pl+1)**3

  allocate(f(nbp), pp1(nbp))
  do ib = 1, nbp
    mi = dom(ib)%ttc_i; mj = dom(ib)%ttc_j; mk = dom(ib)%ttc_k
    allocate(dom(ib)%p(mi,mj,mk), dom(ib)%pp(mi,mj,mk), dom(ib)%su(mi,mj,mk), &
             dom(ib)%sup(mi,mj,mk), f(ib)%a(mi,mj,mk), pp1(ib)%a(mi,mj,mk))
    dom(ib)%p = 0.0; dom(ib)%pp = 0.0; dom(ib)%su = 0.0; dom(ib)%sup = 0.0
    f(ib)%a = 0.0; pp1(ib)%a = 0.0
    allocate(dom(ib)%sendb_m1(ngg), dom(ib)%sendb_p1(ngg), dom(ib)%recvb_m1(ngg), &
             dom(ib)%recvb_p1(ngg), dom(ib)%sendb_m2(ngg), dom(ib)%sendb_p2(ngg), &
             dom(ib)%recvb_m2(ngg), dom(ib)%recvb_p2(ngg), dom(ib)%sendb_m3(ngg), &
             dom(ib)%sendb_p3(ngg), dom(ib)%recvb_m3(ngg), dom(ib)%recvb_p3(ngg))
    ! defined start for the receive buffers (a 2:1 fine side reads entries no
    ! sender fills)
    dom(ib)%recvb_m1 = 0.0; dom(ib)%recvb_p1 = 0.0
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
    ! MG workspace sizing (localparameters.for:223-237)
    allocate(dom(ib)%faz(ngrd_gl))
    nwork = 0
    do l = 1, ngrd_gl
      gl = min(l, dom(ib)%ngrid)
      mi = (dom(ib)%iep-dom(ib)%isp+1)/2**(gl-1)+2
      mj = (dom(ib)%jep-dom(ib)%jsp+1)/2**(gl-1)+2
      mk = (dom(ib)%kep-dom(ib)%ksp+1)/2**(gl-1)+2
      nwork = nwork + mi*mj*mk + 8*(mi-2)*(mj-2)*(mk-2)
    end do
    allocate(dom(ib)%cof(nwork))
    dom(ib)%cof = 0.0
    ! LSM workspace (initial.for:257-278); dens_mg is sized here, not from an
    ! uninitialised tot
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
  end do

  call read_blocks(trim(frhs), f)
  if (L_LSM) call read_dens(trim(fdens))

  call coeff

  allocate(rho(0:maxcy), rmx(0:maxcy))
  rho = 0.0; rmx = 0.0
  call system_clock(clk0, clkrate)
  call residual(nf, rm, .true.)
  call residual(nr, rm, .false.)
  if (nf == 0.0) nf = 1.0
  rho(0) = nr / nf
  rmx(0) = rm
  ncyc = maxcy
  do it = 1, maxcy
    do ib = 1, nbp
      dom(ib)%pp = 0.0
    end do
    call mgkcyc
    if (it == 1) then
      do ib = 1, nbp
        pp1(ib)%a = dom(ib)%pp
      end do
    end if
    do ib = 1, nbp
      do k = dom(ib)%ksp, dom(ib)%kep
        do j = dom(ib)%jsp, dom(ib)%jep
          do i = dom(ib)%isp, dom(ib)%iep
            dom(ib)%p(i,j,k) = dom(ib)%p(i,j,k) + dom(ib)%pp(i,j,k)
          end do
        end do
      end do
    end do
    call residual(nr, rm, .false.)
    rho(it) = nr / nf
    rmx(it) = rm
    if (rho(it) <= tol .and. it >= mincy) then
      ncyc = it
      exit
    end if
  end do

  call system_clock(clk1)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) dble(ncyc)
  write(u) rho
  write(u) rmx
  do ib = 1, nbp
    write(u) dom(ib)%p(dom(ib)%isp:dom(ib)%iep, dom(ib)%jsp:dom(ib)%jep, dom(ib)%ksp:dom(ib)%kep)
  end do
  do ib = 1, nbp
    write(u) pp1(ib)%a
  end do
  write(u) dble(clk1 - clk0) / dble(clkrate)   ! seconds of the solve (after the input is read)
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

  subroutine read_blocks(fname, dst)
    character(len=*), intent(in) :: fname
    type(fieldp), intent(inout) :: dst(:)
    integer :: uu, jb
    open(newunit=uu, file=fname, access='stream', form='unformatted', status='old')
    do jb = 1, nbp
      read(uu) dst(jb)%a(dom(jb)%isp:dom(jb)%iep, dom(jb)%jsp:dom(jb)%jep, dom(jb)%ksp:dom(jb)%kep)
    end do
    close(uu)
  end subroutine

  subroutine read_dens(fname)
    character(len=*), intent(in) :: fname
    integer :: uu, jb
    open(newunit=uu, file=fname, access='stream', form='unformatted', status='old')
    do jb = 1, nbp
      read(uu) dom(jb)%dens(dom(jb)%isp:dom(jb)%iep, dom(jb)%jsp:dom(jb)%jep, &
                            dom(jb)%ksp:dom(jb)%kep)
    end do
    close(uu)
  end subroutine

  ! r = f - A p on level 1: p (with its ghost layer) goes into the level-1
  ! phi slot of cof, the reference's exchangepp(1) refreshes the ghosts, and
  ! the residual is formed as in mgrestr (mgsolver.for:702-709).  r goes
  ! into su (or sup under LSM), where mgkcyc picks up its RHS.  With
  ! want_f the norms of f itself are returned instead.
  subroutine residual(l2, rmax_out, want_f)
    double precision, intent(out) :: l2, rmax_out
    logical, intent(in) :: want_f
    integer :: jb, ni, nj, nk, nij, ijk, ijkp, w, incr
    double precision :: s, r, a
    l2 = 0.0
    rmax_out = 0.0
    do jb = 1, nbp
      ni = dom(jb)%ttc_i; nj = dom(jb)%ttc_j; nk = dom(jb)%ttc_k
      do k = 1, nk
        do i = 1, ni
          do j = 1, nj
            ijk = (k-1)*ni*nj + (i-1)*nj + j
            dom(jb)%cof(ijk) = dom(jb)%p(i,j,k)
          end do
        end do
      end do
    end do
    call exchangepp(1)
    do jb = 1, nbp
      ni = dom(jb)%ttc_i; nj = dom(jb)%ttc_j; nk = dom(jb)%ttc_k
      nij = ni*nj
      incr = (ni-2)*(nj-2)*(nk-2)
      do k = 2, nk-1
        do j = 2, nj-1
          do i = 2, ni-1
            ijkp = (k-1)*nij + (i-1)*nj + j
            w = dom(jb)%faz(1) + (k-2)*(ni-2)*(nj-2) + (i-2)*(nj-2) + (j-1)
            if (want_f) then
              r = f(jb)%a(i,j,k)
            else
              s = dom(jb)%cof(w)*dom(jb)%cof(ijkp-nj) + dom(jb)%cof(w+incr)*dom(jb)%cof(ijkp+nj) &
                + dom(jb)%cof(w+2*incr)*dom(jb)%cof(ijkp-1) + dom(jb)%cof(w+3*incr)*dom(jb)%cof(ijkp+1) &
                + dom(jb)%cof(w+4*incr)*dom(jb)%cof(ijkp-nij) + dom(jb)%cof(w+5*incr)*dom(jb)%cof(ijkp+nij) &
                + dom(jb)%cof(w+6*incr)*dom(jb)%cof(ijkp)
              r = f(jb)%a(i,j,k) - s
              if (L_LSM) then
                dom(jb)%sup(i,j,k) = r
              else
                dom(jb)%su(i,j,k) = r
              end if
            end if
            l2 = l2 + r*r
            if (L_LSM) then
              a = abs(r)
            else
              a = abs(r*dom(jb)%dx*dom(jb)%dy*dom(jb)%dz)
            end if
            if (a > rmax_out) rmax_out = a
          end do
        end do
      end do
    end do
    l2 = sqrt(l2)
  end subroutine

end program
