! ref3_harness.f90 -- C-callable harness around the reference's own 3-D
! Fortran (TEST INFRASTRUCTURE ONLY; built into oracle/_ref, never shipped).
!
! stepgrid, step3, flux3, inlinelimiter, flag2refine and the level update
! around the step (bound, filpatch, intfil, prefilp, bc3amr, saveqc, cstore,
! qad, fluxsv, fluxad, upbnd, update, prepf, prepc, setuse, ...) are
! compiled unmodified from the reference's src/3d by oracle/Makefile
! (target ref3).  rpn3/rpt3/rptt3 are the physics plugin the application
! links (flux3.f, qad.f); here they forward to the oracle's solvers.

subroutine rpn3(ixyz,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,auxl,auxr,wave,s,amdq,apdq)
  use iso_c_binding
  implicit none
  integer :: ixyz,maxm,meqn,mwaves,maux,mbc,mx
  real(kind=8) :: ql(*),qr(*),auxl(*),auxr(*),wave(*),s(*),amdq(*),apdq(*)
  interface
     subroutine orc_rpn3_byref(ixyz,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,auxl,auxr, &
          wave,s,amdq,apdq) bind(C, name="orc_rpn3_byref")
       import :: c_int, c_double
       integer(c_int) :: ixyz,maxm,meqn,mwaves,maux,mbc,mx
       real(c_double) :: ql(*),qr(*),auxl(*),auxr(*),wave(*),s(*),amdq(*),apdq(*)
     end subroutine orc_rpn3_byref
  end interface
  call orc_rpn3_byref(ixyz,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,auxl,auxr,wave,s,amdq,apdq)
end subroutine rpn3

subroutine rpt3(ixyz,icoor,imp,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1,aux2,aux3, &
     asdq,bmasdq,bpasdq)
  use iso_c_binding
  implicit none
  integer :: ixyz,icoor,imp,maxm,meqn,mwaves,maux,mbc,mx
  real(kind=8) :: ql(*),qr(*),aux1(*),aux2(*),aux3(*),asdq(*),bmasdq(*),bpasdq(*)
  interface
     subroutine orc_rpt3_byref(ixyz,icoor,imp,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1, &
          aux2,aux3,asdq,bmasdq,bpasdq) bind(C, name="orc_rpt3_byref")
       import :: c_int, c_double
       integer(c_int) :: ixyz,icoor,imp,maxm,meqn,mwaves,maux,mbc,mx
       real(c_double) :: ql(*),qr(*),aux1(*),aux2(*),aux3(*),asdq(*),bmasdq(*),bpasdq(*)
     end subroutine orc_rpt3_byref
  end interface
  call orc_rpt3_byref(ixyz,icoor,imp,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1,aux2,aux3, &
       asdq,bmasdq,bpasdq)
end subroutine rpt3

subroutine rptt3(ixyz,icoor,imp,impt,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1,aux2,aux3, &
     bsasdq,cmbsasdq,cpbsasdq)
  use iso_c_binding
  implicit none
  integer :: ixyz,icoor,imp,impt,maxm,meqn,mwaves,maux,mbc,mx
  real(kind=8) :: ql(*),qr(*),aux1(*),aux2(*),aux3(*),bsasdq(*),cmbsasdq(*),cpbsasdq(*)
  interface
     subroutine orc_rptt3_byref(ixyz,icoor,imp,impt,maxm,meqn,mwaves,maux,mbc,mx,ql,qr, &
          aux1,aux2,aux3,bsasdq,cmbsasdq,cpbsasdq) bind(C, name="orc_rptt3_byref")
       import :: c_int, c_double
       integer(c_int) :: ixyz,icoor,imp,impt,maxm,meqn,mwaves,maux,mbc,mx
       real(c_double) :: ql(*),qr(*),aux1(*),aux2(*),aux3(*),bsasdq(*),cmbsasdq(*),cpbsasdq(*)
     end subroutine orc_rptt3_byref
  end interface
  call orc_rptt3_byref(ixyz,icoor,imp,impt,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1,aux2,aux3, &
       bsasdq,cmbsasdq,cpbsasdq)
end subroutine rptt3

module ref3_harness
  use iso_c_binding
  use amr_module
  implicit none
contains

  subroutine ref3_set_method(meth, lim, nwaves, ncapa, fwaves, cfl_in, cflv1_in, ng) &
       bind(C, name="ref3_set_method")
    integer(c_int), intent(in) :: meth(7), lim(*)
    integer(c_int), value :: nwaves, ncapa, fwaves, ng
    real(c_double), value :: cfl_in, cflv1_in
    method(1:7) = meth(1:7)
    mwaves = nwaves
    if (allocated(mthlim)) deallocate(mthlim)
    allocate(mthlim(mwaves))
    mthlim(1:mwaves) = lim(1:mwaves)
    mcapa = ncapa
    use_fwaves = (fwaves /= 0)
    cfl = cfl_in
    cflv1 = cflv1_in
    nghost = ng
  end subroutine ref3_set_method

  ! stepgrid.f (3-D) on one patch, q and the six flux arrays
  ! (nvar, mx+2mbc, my+2mbc, mz+2mbc)
  subroutine ref3_stepgrid(nvar, mbc, mx, my, mz, q, dx, dy, dz, dt, fm, fp, gm, gp, hm, hp, &
       cflgrid, dtnew) bind(C, name="ref3_stepgrid")
    integer(c_int), value :: nvar, mbc, mx, my, mz
    real(c_double), value :: dx, dy, dz, dt
    real(c_double) :: q(*), fm(*), fp(*), gm(*), gp(*), hm(*), hp(*), cflgrid, dtnew
    real(kind=8) :: aux1(1)
    cfl_level = 0.d0
    call stepgrid(q, fm, fp, gm, gp, hm, hp, mx + 2*mbc, my + 2*mbc, mz + 2*mbc, mbc, dt, &
         dtnew, dx, dy, dz, nvar, 0.d0, 0.d0, 0.d0, 0.d0, 1, 0, aux1)
    cflgrid = cfl_level
  end subroutine ref3_stepgrid

  ! flag2refine.f (3-D) as spest.f:71-84 calls it: amrflags set to goodpt,
  ! then the interior flagged goodpt / badpt
  subroutine ref3_flag2refine(nvar, mbc, mx, my, mz, q, tolsp_in, amrflags) &
       bind(C, name="ref3_flag2refine")
    integer(c_int), value :: nvar, mbc, mx, my, mz
    real(c_double), value :: tolsp_in
    real(c_double) :: q(*), amrflags(*)
    real(kind=8) :: aux1(1)
    integer :: n
    n = (mx + 2*mbc) * (my + 2*mbc) * (mz + 2*mbc)
    amrflags(1:n) = goodpt
    call flag2refine(mx, my, mz, mbc, nvar, 0, 0.d0, 0.d0, 0.d0, 1.d0, 1.d0, 1.d0, 0.d0, &
         1, tolsp_in, q, aux1, amrflags, goodpt, badpt)
  end subroutine ref3_flag2refine

  ! ---------------------------------------------------------------------
  ! Hierarchy set-up as in the 2-D harness: node/rnode/alloc/level lists
  ! the way regrid leaves them (src/3d/amr_module.f90), then makeBndryList
  ! (init_bndryList.f).  Grid k of level l is node gbase(l)+k; its new/old
  ! solution blocks are consecutive in alloc.
  subroutine ref3_hier_setup(nlev, mxnest_in, nvar, ng, ngr, boxes, hx, hy, hz, pk, tm, &
       isz, jsz, ksz, ratx, raty, ratz, xlo, xhi, ylo, yhi, zlo, zhi, xper, yper, zper, bc) &
       bind(C, name="ref3_hier_setup")
    integer(c_int), value :: nlev, mxnest_in, nvar, ng
    integer(c_int), intent(in) :: ngr(nlev), boxes(*), isz(nlev), jsz(nlev), ksz(nlev)
    integer(c_int), intent(in) :: ratx(nlev), raty(nlev), ratz(nlev), bc(6)
    real(c_double), intent(in) :: hx(nlev), hy(nlev), hz(nlev), pk(nlev), tm(nlev)
    real(c_double), value :: xlo, xhi, ylo, yhi, zlo, zhi
    integer(c_int), value :: xper, yper, zper
    integer :: l, k, mptr, nb, ntot, loc, mitot, mjtot, mktot, mprev
    ntot = sum(ngr(1:nlev))
    maxgr = max(ntot + 10, 16)
    if (allocated(node)) deallocate(node)
    if (allocated(rnode)) deallocate(rnode)
    if (allocated(listOfGrids)) deallocate(listOfGrids)
    allocate(node(nsize, maxgr), rnode(rsize, maxgr), listOfGrids(maxgr))
    node = 0
    rnode = 0.d0
    memsize = 1
    nb = 0
    do l = 1, nlev
       do k = 1, ngr(l)
          mitot = boxes(6*nb+2) - boxes(6*nb+1) + 1 + 2*ng
          mjtot = boxes(6*nb+4) - boxes(6*nb+3) + 1 + 2*ng
          mktot = boxes(6*nb+6) - boxes(6*nb+5) + 1 + 2*ng
          memsize = memsize + 2*nvar*mitot*mjtot*mktot
          nb = nb + 1
       end do
    end do
    memsize = memsize + 8000000    ! room for registers and saveqc scratch
    if (allocated(alloc)) deallocate(alloc)
    allocate(alloc(memsize))
    alloc = 0.d0
    mxnest = mxnest_in
    nghost = ng
    xlower = xlo
    xupper = xhi
    ylower = ylo
    yupper = yhi
    zlower = zlo
    zupper = zhi
    xperdom = (xper /= 0)
    yperdom = (yper /= 0)
    zperdom = (zper /= 0)
    mthbc(1:6) = bc(1:6)
    lstart = 0
    numgrids = 0
    listStart = 0
    nb = 0
    loc = 2
    mptr = 0
    do l = 1, nlev
       hxposs(l) = hx(l)
       hyposs(l) = hy(l)
       hzposs(l) = hz(l)
       possk(l) = pk(l)
       iregsz(l) = isz(l)
       jregsz(l) = jsz(l)
       kregsz(l) = ksz(l)
       intratx(l) = ratx(l)
       intraty(l) = raty(l)
       intratz(l) = ratz(l)
       numgrids(l) = ngr(l)
       listStart(l) = mptr + 1
       mprev = 0
       do k = 1, ngr(l)
          mptr = mptr + 1
          listOfGrids(mptr) = mptr
          if (k == 1) lstart(l) = mptr
          if (mprev /= 0) node(levelptr, mprev) = mptr
          node(levelptr, mptr) = null
          mprev = mptr
          node(nestlevel, mptr) = l
          node(ndilo, mptr) = boxes(6*nb+1)
          node(ndihi, mptr) = boxes(6*nb+2)
          node(ndjlo, mptr) = boxes(6*nb+3)
          node(ndjhi, mptr) = boxes(6*nb+4)
          node(ndklo, mptr) = boxes(6*nb+5)
          node(ndkhi, mptr) = boxes(6*nb+6)
          rnode(cornxlo, mptr) = node(ndilo, mptr)*hx(l) + xlower
          rnode(cornxhi, mptr) = (node(ndihi, mptr)+1)*hx(l) + xlower
          rnode(cornylo, mptr) = node(ndjlo, mptr)*hy(l) + ylower
          rnode(cornyhi, mptr) = (node(ndjhi, mptr)+1)*hy(l) + ylower
          rnode(cornzlo, mptr) = node(ndklo, mptr)*hz(l) + zlower
          rnode(cornzhi, mptr) = (node(ndkhi, mptr)+1)*hz(l) + zlower
          rnode(timemult, mptr) = tm(l)
          mitot = node(ndihi, mptr) - node(ndilo, mptr) + 1 + 2*ng
          mjtot = node(ndjhi, mptr) - node(ndjlo, mptr) + 1 + 2*ng
          mktot = node(ndkhi, mptr) - node(ndklo, mptr) + 1 + 2*ng
          node(store1, mptr) = loc
          node(store2, mptr) = loc + nvar*mitot*mjtot*mktot
          node(storeaux, mptr) = 1
          node(bndListSt, mptr) = null
          node(bndListNum, mptr) = 0
          loc = loc + 2*nvar*mitot*mjtot*mktot
          nb = nb + 1
       end do
    end do
    listStart(nlev+1) = mptr + 1
    lenf = 1
    lfree(1,1) = loc
    lfree(1,2) = memsize - loc + 1
    lendim = loc
    if (allocated(bndList)) deallocate(bndList)
    call initBndryList()
    do l = 1, nlev
       call makeBndryList(l)
    end do
  end subroutine ref3_hier_setup

  integer function ref3_gridlen(mptr, nvar)
    integer :: mptr, nvar
    ref3_gridlen = nvar*(node(ndihi,mptr)-node(ndilo,mptr)+1+2*nghost) &
         *(node(ndjhi,mptr)-node(ndjlo,mptr)+1+2*nghost)*(node(ndkhi,mptr)-node(ndklo,mptr)+1+2*nghost)
  end function ref3_gridlen

  ! copy grid (level l, index k from 0) new (which=0) / old (1) to or from q
  subroutine ref3_hier_xfer(l, k, which, nvar, q, to_ref) bind(C, name="ref3_hier_xfer")
    integer(c_int), value :: l, k, which, nvar, to_ref
    real(c_double) :: q(*)
    integer :: mptr, loc, n
    mptr = listOfGrids(listStart(l) + k)
    loc = node(store1, mptr)
    if (which == 1) loc = node(store2, mptr)
    n = ref3_gridlen(mptr, nvar)
    if (to_ref /= 0) then
       alloc(loc:loc+n-1) = q(1:n)
    else
       q(1:n) = alloc(loc:loc+n-1)
    end if
  end subroutine ref3_hier_xfer

  ! bound.f (3-D) on grid k of level l at time t, in place on its new array
  subroutine ref3_bound(l, k, t, nvar) bind(C, name="ref3_bound")
    integer(c_int), value :: l, k, nvar
    real(c_double), value :: t
    integer :: mptr, mitot, mjtot, mktot
    mptr = listOfGrids(listStart(l) + k)
    mitot = node(ndihi,mptr)-node(ndilo,mptr)+1+2*nghost
    mjtot = node(ndjhi,mptr)-node(ndjlo,mptr)+1+2*nghost
    mktot = node(ndkhi,mptr)-node(ndklo,mptr)+1+2*nghost
    call bound(t, nvar, nghost, alloc(node(store1,mptr)), mitot, mjtot, mktot, mptr, &
         alloc(node(storeaux,mptr)), 0)
  end subroutine ref3_bound

  ! prepf (fine levels) and prepc (coarse levels), as regrid does
  subroutine ref3_prep_regs(nlev, nvar) bind(C, name="ref3_prep_regs")
    integer(c_int), value :: nlev, nvar
    integer :: l
    do l = 2, nlev
       call prepf(l, nvar, 0)
    end do
    do l = 1, nlev - 1
       call prepc(l, nvar)
    end do
  end subroutine ref3_prep_regs

  ! advanc.f (3-D) restated around the reference's own routines (advanc.f
  ! itself needs gauges_module, whose utility_module is absent)
  subroutine ref3_advanc(level, nvar, dtlevnew, cflout) bind(C, name="ref3_advanc")
    integer(c_int), value :: level, nvar
    real(c_double) :: dtlevnew, cflout
    integer :: j, mptr, mitot, mjtot, mktot
    real(kind=8) :: dtnew
    do j = 1, numgrids(level)
       mptr = listOfGrids(listStart(level)+j-1)
       mitot = node(ndihi,mptr) - node(ndilo,mptr) + 1 + 2*nghost
       mjtot = node(ndjhi,mptr) - node(ndjlo,mptr) + 1 + 2*nghost
       mktot = node(ndkhi,mptr) - node(ndklo,mptr) + 1 + 2*nghost
       call bound(rnode(timemult,mptr), nvar, nghost, alloc(node(store1,mptr)), mitot, mjtot, &
            mktot, mptr, alloc(node(storeaux,mptr)), 0)
    end do
    if (level+1 .le. mxnest) then
       if (lstart(level+1) .ne. null) call saveqc(level+1, nvar, 0)
    endif
    dtlevnew = rinfinity
    cfl_level = 0.d0
    do j = 1, numgrids(level)
       mptr = listOfGrids(listStart(level)+j-1)
       mitot = node(ndihi,mptr) - node(ndilo,mptr) + 1 + 2*nghost
       mjtot = node(ndjhi,mptr) - node(ndjlo,mptr) + 1 + 2*nghost
       mktot = node(ndkhi,mptr) - node(ndklo,mptr) + 1 + 2*nghost
       call harness3_par_advanc(mptr, mitot, mjtot, mktot, nvar, dtnew)
       dtlevnew = dmin1(dtlevnew, dtnew)
    end do
    cflout = cfl_level
  end subroutine ref3_advanc

  ! advanc.f (3-D) par_advanc without gauges and timers
  subroutine harness3_par_advanc(mptr, mitot, mjtot, mktot, nvar, dtnew)
    integer :: mptr, mitot, mjtot, mktot, nvar
    real(kind=8) :: dtnew
    real(kind=8), allocatable :: fp(:), fm(:), gp(:), gm(:), hp(:), hm(:)
    integer :: level, nx, ny, nz, locold, locnew, locaux, ntot, i, lenbc, locsvf, locsvq, locx1d
    real(kind=8) :: hx, hy, hz, delt, time, xlow, ylow, zlow
    ntot = nvar*mitot*mjtot*mktot
    allocate(fp(ntot), fm(ntot), gp(ntot), gm(ntot), hp(ntot), hm(ntot))
    fp = 0.d0
    fm = 0.d0
    gp = 0.d0
    gm = 0.d0
    hp = 0.d0
    hm = 0.d0
    level = node(nestlevel,mptr)
    hx = hxposs(level)
    hy = hyposs(level)
    hz = hzposs(level)
    delt = possk(level)
    nx = node(ndihi,mptr) - node(ndilo,mptr) + 1
    ny = node(ndjhi,mptr) - node(ndjlo,mptr) + 1
    nz = node(ndkhi,mptr) - node(ndklo,mptr) + 1
    time = rnode(timemult,mptr)
    locold = node(store2, mptr)
    locnew = node(store1, mptr)
    if (level .lt. mxnest) then
       do i = 1, ntot
          alloc(locold + i - 1) = alloc(locnew + i - 1)
       end do
    endif
    xlow = rnode(cornxlo,mptr) - nghost*hx
    ylow = rnode(cornylo,mptr) - nghost*hy
    zlow = rnode(cornzlo,mptr) - nghost*hz
    locaux = node(storeaux,mptr)
    if (node(ffluxptr,mptr) .ne. 0) then
       lenbc = 2*( (nx/intratx(level-1))*(nz/intratz(level-1)) &
            +(ny/intraty(level-1))*(nz/intratz(level-1)) &
            +(nx/intratx(level-1))*(ny/intraty(level-1)))
       locsvf = node(ffluxptr,mptr)
       locsvq = locsvf + nvar*lenbc
       locx1d = locsvq + nvar*lenbc
       call qad(alloc(locnew),mitot,mjtot,mktot,nvar,alloc(locsvf),alloc(locsvq),lenbc, &
            intratx(level-1),intraty(level-1),intratz(level-1),hx,hy,hz,0,alloc(locaux), &
            alloc(locx1d),delt,mptr)
    endif
    call stepgrid(alloc(locnew),fm,fp,gm,gp,hm,hp,mitot,mjtot,mktot,nghost,delt,dtnew, &
         hx,hy,hz,nvar,xlow,ylow,zlow,time,mptr,0,alloc(locaux))
    if (node(cfluxptr,mptr) .ne. 0) call fluxsv(mptr,fm,fp,gm,gp,hm,hp, &
         alloc(node(cfluxptr,mptr)),mitot,mjtot,mktot,nvar,listsp(level),delt,hx,hy,hz)
    if (node(ffluxptr,mptr) .ne. 0) then
       lenbc = 2*( (nx/intratx(level-1))*(nz/intratz(level-1)) &
            +(ny/intraty(level-1))*(nz/intratz(level-1)) &
            +(nx/intratx(level-1))*(ny/intraty(level-1)))
       locsvf = node(ffluxptr,mptr)
       call fluxad(fm,fp,gm,gp,hm,hp,alloc(locsvf),mptr,mitot,mjtot,mktot,nvar,lenbc, &
            intratx(level-1),intraty(level-1),intratz(level-1),nghost,delt,hx,hy,hz)
    endif
    rnode(timemult,mptr) = rnode(timemult,mptr)+delt
    deallocate(fp, fm, gp, gm, hp, hm)
  end subroutine harness3_par_advanc

  subroutine ref3_update(level, nvar) bind(C, name="ref3_update")
    integer(c_int), value :: level, nvar
    call update(level, nvar, 0)
  end subroutine ref3_update

  ! register block (svdflx then qc1d) of grid k (0-based) of fine level l
  subroutine ref3_get_reg(l, k, nvar, out) bind(C, name="ref3_get_reg")
    integer(c_int), value :: l, k, nvar
    real(c_double) :: out(*)
    integer :: mptr, lenbc, nx, ny, nz
    mptr = listOfGrids(listStart(l) + k)
    nx = node(ndihi,mptr) - node(ndilo,mptr) + 1
    ny = node(ndjhi,mptr) - node(ndjlo,mptr) + 1
    nz = node(ndkhi,mptr) - node(ndklo,mptr) + 1
    lenbc = 2*((nx/intratx(l-1))*(ny/intraty(l-1)) + (ny/intraty(l-1))*(nz/intratz(l-1)) &
         + (nz/intratz(l-1))*(nx/intratx(l-1)))
    out(1:2*nvar*lenbc) = alloc(node(ffluxptr,mptr):node(ffluxptr,mptr)+2*nvar*lenbc-1)
  end subroutine ref3_get_reg

  ! listbc of grid k (0-based) of coarse level l (6 integers per entry)
  subroutine ref3_get_listbc(l, k, nmax, out, n) bind(C, name="ref3_get_listbc")
    integer(c_int), value :: l, k, nmax
    integer(c_int) :: out(*), n
    integer :: mptr
    external copy_listbc3
    mptr = listOfGrids(listStart(l) + k)
    n = 0
    if (node(cfluxptr,mptr) .eq. 0) return
    call copy_listbc3(alloc(node(cfluxptr,mptr)), listsp(l), nmax, out, n)
  end subroutine ref3_get_listbc

  subroutine ref3_get_time(l, k, t) bind(C, name="ref3_get_time")
    integer(c_int), value :: l, k
    real(c_double) :: t
    t = rnode(timemult, listOfGrids(listStart(l) + k))
  end subroutine ref3_get_time

  ! spest.f (3-D) for every grid of `level` and, with frich, errest.f +
  ! errf1.f (whose flags errf1 leaves in component 1 of each grid's store2);
  ! domflags is dom1flags(0:isize+1, 0:jsize+1, 0:ksize+1), rflags the
  ! grids' store2 component-1 values one after the other.  spest.f itself
  ! does not compile here (its dummy argument t0 clashes with amr_module's
  ! t0, spest.f:4-6), so its per-grid steps are performed below with the
  ! reference's own bound, flag2refine and setflags (spest.f:14-99).
  subroutine ref3_flag_level(level, nvar, tolsp_in, tol_in, iorder_in, frich, domflags, rflags) &
       bind(C, name="ref3_flag_level")
    integer(c_int), value :: level, nvar, iorder_in, frich
    real(c_double), value :: tolsp_in, tol_in
    integer(c_int8_t) :: domflags(*)
    real(c_double) :: rflags(*)
    integer :: j, mptr, n, off, c, loc
    tolsp = tolsp_in
    tol = tol_in
    iorder = iorder_in
    flag_gradient = .true.
    flag_richardson = frich .ne. 0
    eprint = .false.
    edebug = .false.
    call harness3_spest(nvar, level, domflags, iregsz(level), jregsz(level), kregsz(level))
    if (frich .ne. 0) call errest(nvar, 0, level)
    off = 0
    do j = 1, numgrids(level)
       mptr = listOfGrids(listStart(level) + j - 1)
       n = ref3_gridlen(mptr, 1)
       if (frich .ne. 0) then
          loc = node(store2, mptr)
          do c = 1, n
             rflags(off + c) = alloc(loc + (c - 1) * nvar)
          end do
       end if
       call reclam(node(tempptr, mptr), n * nvar)
       off = off + n
    end do
  end subroutine ref3_flag_level

  subroutine harness3_spest(nvar, lcheck, dom1flags, isize, jsize, ksize)
    integer :: nvar, lcheck, isize, jsize, ksize
    integer(c_int8_t) :: dom1flags(0:isize+1, 0:jsize+1, 0:ksize+1)
    integer :: i, j, k, mptr, nx, ny, nz, mitot, mjtot, mktot, locnew, locbig, locamrflags
    integer, external :: igetsp
    real(kind=8) :: time, dx, dy, dz, xlow, ylow, zlow, aux1(1)
    external flag2refine
    do i = 1, isize
       do j = 1, jsize
          do k = 1, ksize
             dom1flags(i, j, k) = 0
          end do
       end do
    end do
    mptr = lstart(lcheck)
    do while (mptr .ne. 0)
       nx = node(ndihi,mptr) - node(ndilo,mptr) + 1
       ny = node(ndjhi,mptr) - node(ndjlo,mptr) + 1
       nz = node(ndkhi,mptr) - node(ndklo,mptr) + 1
       mitot = nx + 2*nghost
       mjtot = ny + 2*nghost
       mktot = nz + 2*nghost
       locnew = node(store1,mptr)
       time = rnode(timemult,mptr)
       dx = hxposs(lcheck)
       dy = hyposs(lcheck)
       dz = hzposs(lcheck)
       xlow = rnode(cornxlo,mptr) - nghost*dx
       ylow = rnode(cornylo,mptr) - nghost*dy
       zlow = rnode(cornzlo,mptr) - nghost*dz
       locbig = igetsp(mitot*mjtot*mktot*nvar)
       node(tempptr,mptr) = locbig
       do i = 1, mitot*mjtot*mktot*nvar
          alloc(locbig+i-1) = alloc(locnew+i-1)
       end do
       call bound(time,nvar,nghost,alloc(locbig),mitot,mjtot,mktot,mptr,aux1,0)
       locamrflags = igetsp(mitot*mjtot*mktot)
       do i = 1, mitot*mjtot*mktot
          alloc(locamrflags+i-1) = goodpt
       end do
       call flag2refine(nx,ny,nz,nghost,nvar,0,xlow,ylow,zlow,dx,dy,dz,time,lcheck,tolsp, &
            alloc(locbig),aux1,alloc(locamrflags),goodpt,badpt)
       call setflags(dom1flags,isize,jsize,ksize,alloc(locamrflags),1,mitot,mjtot,mktot,mptr)
       call reclam(locamrflags, mitot*mjtot*mktot)
       mptr = node(levelptr,mptr)
    end do
  end subroutine harness3_spest

end module ref3_harness

! listbc lives in the double heap as integers (prepc.f, setuse.f); read it
! through an integer dummy as fluxsv does
subroutine copy_listbc3(listbc, maxsp, nmax, out, n)
  implicit none
  integer :: maxsp, nmax, n, listbc(6, maxsp), out(6, *)
  integer :: e
  n = 0
  do e = 1, maxsp
     if (listbc(1, e) .eq. 0) exit
     n = n + 1
     if (n .le. nmax) out(:, n) = listbc(:, e)
  end do
end subroutine copy_listbc3
