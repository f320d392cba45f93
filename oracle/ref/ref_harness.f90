! ref_harness.f90 -- C-callable harness around the reference's own Fortran
! (TEST INFRASTRUCTURE ONLY; built into oracle/_ref, never shipped).
!
! The reference routines (step2, stepgrid, flag2refine2, bc2amr, qad, fluxsv,
! fluxad, upbnd, ...) are compiled unmodified from the reference tree by
! oracle/Makefile.  This file is the caller: it fills the amr_module state
! those routines read and exposes bind(C) entry points for the tests.
! rpn2/rpt2 are the physics plugin the reference expects the application to
! link (flux2.f:125, qad.f:129); here they forward to the oracle's solvers.

subroutine rpn2(ixy,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,auxl,auxr,wave,s,amdq,apdq)
  use iso_c_binding
  implicit none
  integer :: ixy,maxm,meqn,mwaves,maux,mbc,mx
  real(kind=8) :: ql(*),qr(*),auxl(*),auxr(*),wave(*),s(*),amdq(*),apdq(*)
  interface
     subroutine orc_rpn2_byref(ixy,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,auxl,auxr, &
          wave,s,amdq,apdq) bind(C, name="orc_rpn2_byref")
       import :: c_int, c_double
       integer(c_int) :: ixy,maxm,meqn,mwaves,maux,mbc,mx
       real(c_double) :: ql(*),qr(*),auxl(*),auxr(*),wave(*),s(*),amdq(*),apdq(*)
     end subroutine orc_rpn2_byref
  end interface
  call orc_rpn2_byref(ixy,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,auxl,auxr,wave,s,amdq,apdq)
end subroutine rpn2

subroutine rpt2(ixy,imp,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1,aux2,aux3,asdq,bmasdq,bpasdq)
  use iso_c_binding
  implicit none
  integer :: ixy,imp,maxm,meqn,mwaves,maux,mbc,mx
  real(kind=8) :: ql(*),qr(*),aux1(*),aux2(*),aux3(*),asdq(*),bmasdq(*),bpasdq(*)
  interface
     subroutine orc_rpt2_byref(ixy,imp,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1,aux2, &
          aux3,asdq,bmasdq,bpasdq) bind(C, name="orc_rpt2_byref")
       import :: c_int, c_double
       integer(c_int) :: ixy,imp,maxm,meqn,mwaves,maux,mbc,mx
       real(c_double) :: ql(*),qr(*),aux1(*),aux2(*),aux3(*),asdq(*),bmasdq(*),bpasdq(*)
     end subroutine orc_rpt2_byref
  end interface
  call orc_rpt2_byref(ixy,imp,maxm,meqn,mwaves,maux,mbc,mx,ql,qr,aux1,aux2,aux3,asdq,bmasdq,bpasdq)
end subroutine rpt2

module ref_harness
  use iso_c_binding
  use amr_module
  implicit none
  logical :: harness_parallel = .false.
  ! aux values per cell of the installed grids (ref_hier_set_aux)
  integer :: hnaux = 0
contains

  ! method(1:7), mthlim(1:mwaves), mcapa, use_fwaves, cfl_desired, cfl_max
  subroutine ref_set_method(meth, lim, nwaves, ncapa, fwaves, cfl_in, cflv1_in, ng) &
       bind(C, name="ref_set_method")
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
    dimensional_split = 0
  end subroutine ref_set_method

  ! step2.f90:7 on one patch; arrays are (meqn|maux, mx+2mbc, my+2mbc)
  subroutine ref_step2(meqn, maux, mbc, mx, my, qold, aux, dx, dy, dt, cflgrid, &
       fm, fp, gm, gp) bind(C, name="ref_step2")
    integer(c_int), value :: meqn, maux, mbc, mx, my
    real(c_double), value :: dx, dy, dt
    real(c_double) :: qold(meqn,1-mbc:mx+mbc,1-mbc:my+mbc)
    real(c_double) :: aux(max(maux,1),1-mbc:mx+mbc,1-mbc:my+mbc)
    real(c_double) :: fm(meqn,1-mbc:mx+mbc,1-mbc:my+mbc), fp(meqn,1-mbc:mx+mbc,1-mbc:my+mbc)
    real(c_double) :: gm(meqn,1-mbc:mx+mbc,1-mbc:my+mbc), gp(meqn,1-mbc:mx+mbc,1-mbc:my+mbc)
    real(c_double) :: cflgrid
    external rpn2, rpt2
    call step2(max(mx,my), meqn, maux, mbc, mx, my, qold, aux, dx, dy, dt, cflgrid, &
         fm, fp, gm, gp, rpn2, rpt2)
  end subroutine ref_step2

  ! stepgrid.f:24 on one patch (q updated in place)
  subroutine ref_stepgrid(nvar, naux, mbc, mx, my, q, aux, dx, dy, dt, time, &
       fm, fp, gm, gp, dtnew, cflgrid) bind(C, name="ref_stepgrid")
    integer(c_int), value :: nvar, naux, mbc, mx, my
    real(c_double), value :: dx, dy, dt, time
    real(c_double) :: q(nvar,mx+2*mbc,my+2*mbc), aux(max(naux,1),mx+2*mbc,my+2*mbc)
    real(c_double) :: fm(nvar,mx+2*mbc,my+2*mbc), fp(nvar,mx+2*mbc,my+2*mbc)
    real(c_double) :: gm(nvar,mx+2*mbc,my+2*mbc), gp(nvar,mx+2*mbc,my+2*mbc)
    real(c_double) :: dtnew, cflgrid
    real(c_double) :: xlow, ylow
    xlow = 0.d0
    ylow = 0.d0
    cfl_level = 0.d0
    call stepgrid(q, fm, fp, gm, gp, mx+2*mbc, my+2*mbc, mbc, dt, dtnew, dx, dy, &
         nvar, xlow, ylow, time, 1, naux, aux)
    cflgrid = cfl_level
  end subroutine ref_stepgrid

  ! The reference's OpenMP patch loop (advanc.f:106-129) over a batch of
  ! patches: stepgrid on each, dynamic schedule, one patch per chunk.
  ! q holds the patches back to back, each (nvar, mx+2mbc, my+2mbc).
  subroutine ref_stepgrid_batch(npatch, nvar, mbc, mxs, mys, offs, q, dx, dy, dt, &
       cflmax) bind(C, name="ref_stepgrid_batch")
    integer(c_int), value :: npatch, nvar, mbc
    integer(c_int), intent(in) :: mxs(npatch), mys(npatch)
    integer(c_int64_t), intent(in) :: offs(npatch)
    real(c_double) :: q(*)
    real(c_double), value :: dx, dy, dt
    real(c_double) :: cflmax
    integer :: k, mitot, mjtot
    real(kind=8) :: dtnew, aux1(1), cmax
    real(kind=8), allocatable :: fm(:), fp(:), gm(:), gp(:)
    cfl_level = 0.d0
    cmax = 0.d0
    !$omp parallel do schedule(dynamic,1) private(k,mitot,mjtot,fm,fp,gm,gp,dtnew,aux1) &
    !$omp& reduction(max:cmax)
    do k = 1, npatch
       mitot = mxs(k) + 2*mbc
       mjtot = mys(k) + 2*mbc
       allocate(fm(nvar*mitot*mjtot), fp(nvar*mitot*mjtot), gm(nvar*mitot*mjtot), &
            gp(nvar*mitot*mjtot))
       call stepgrid(q(offs(k)+1), fm, fp, gm, gp, mitot, mjtot, mbc, dt, dtnew, &
            dx, dy, nvar, 0.d0, 0.d0, 0.d0, k, 0, aux1)
       deallocate(fm, fp, gm, gp)
    end do
    !$omp end parallel do
    cflmax = max(cmax, cfl_level)
  end subroutine ref_stepgrid_batch

  ! ---------------------------------------------------------------------
  ! Hierarchy set-up: fills node/rnode/alloc/level lists the way regrid and
  ! makeGridList leave them (amr_module.f90:30-151), then makeBndryList
  ! (init_bndryList.f:50).  Grid k of level l gets node number
  ! gbase(l)+k; its new/old solution blocks are consecutive in alloc.
  subroutine ref_hier_setup(nlev, mxnest_in, nvar, ng, ngr, boxes, hx, hy, pk, tm, &
       isz, jsz, ratx, raty, xlo, xhi, ylo, yhi, xper, yper, bc) bind(C, name="ref_hier_setup")
    integer(c_int), value :: nlev, mxnest_in, nvar, ng
    integer(c_int), intent(in) :: ngr(nlev), boxes(*), isz(nlev), jsz(nlev), ratx(nlev), raty(nlev), bc(4)
    real(c_double), intent(in) :: hx(nlev), hy(nlev), pk(nlev), tm(nlev)
    real(c_double), value :: xlo, xhi, ylo, yhi
    integer(c_int), value :: xper, yper
    integer :: l, k, mptr, nb, ntot, loc, mitot, mjtot, mprev
    ntot = sum(ngr(1:nlev))
    maxgr = max(ntot + 10, 16)
    if (allocated(node)) deallocate(node)
    if (allocated(rnode)) deallocate(rnode)
    if (allocated(listOfGrids)) deallocate(listOfGrids)
    allocate(node(nsize, maxgr), rnode(rsize, maxgr), listOfGrids(maxgr))
    node = 0
    rnode = 0.d0
    ! solution storage
    memsize = 1
    nb = 0
    do l = 1, nlev
       do k = 1, ngr(l)
          mitot = boxes(4*nb+2) - boxes(4*nb+1) + 1 + 2*ng
          mjtot = boxes(4*nb+4) - boxes(4*nb+3) + 1 + 2*ng
          memsize = memsize + 2*nvar*mitot*mjtot
          nb = nb + 1
       end do
    end do
    memsize = memsize + 4000000    ! room for registers and saveqc scratch
    if (allocated(alloc)) deallocate(alloc)
    allocate(alloc(memsize))
    alloc = 0.d0
    mxnest = mxnest_in
    nghost = ng
    xlower = xlo
    xupper = xhi
    ylower = ylo
    yupper = yhi
    xperdom = (xper /= 0)
    yperdom = (yper /= 0)
    spheredom = .false.
    mthbc(1:4) = bc(1:4)
    lstart = 0
    numgrids = 0
    listStart = 0
    nb = 0
    loc = 2
    mptr = 0
    do l = 1, nlev
       hxposs(l) = hx(l)
       hyposs(l) = hy(l)
       possk(l) = pk(l)
       iregsz(l) = isz(l)
       jregsz(l) = jsz(l)
       intratx(l) = ratx(l)
       intraty(l) = raty(l)
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
          node(ndilo, mptr) = boxes(4*nb+1)
          node(ndihi, mptr) = boxes(4*nb+2)
          node(ndjlo, mptr) = boxes(4*nb+3)
          node(ndjhi, mptr) = boxes(4*nb+4)
          rnode(cornxlo, mptr) = node(ndilo, mptr)*hx(l) + xlower
          rnode(cornxhi, mptr) = (node(ndihi, mptr)+1)*hx(l) + xlower
          rnode(cornylo, mptr) = node(ndjlo, mptr)*hy(l) + ylower
          rnode(cornyhi, mptr) = (node(ndjhi, mptr)+1)*hy(l) + ylower
          rnode(timemult, mptr) = tm(l)
          mitot = node(ndihi, mptr) - node(ndilo, mptr) + 1 + 2*ng
          mjtot = node(ndjhi, mptr) - node(ndjlo, mptr) + 1 + 2*ng
          node(store1, mptr) = loc
          node(store2, mptr) = loc + nvar*mitot*mjtot
          node(storeaux, mptr) = 1
          node(bndListSt, mptr) = null
          node(bndListNum, mptr) = 0
          loc = loc + 2*nvar*mitot*mjtot
          nb = nb + 1
       end do
    end do
    listStart(nlev+1) = mptr + 1
    ! free list of the heap (igetsp.f) after the solution blocks
    lenf = 1
    lfree(1,1) = loc
    lfree(1,2) = memsize - loc + 1
    lendim = loc
    if (allocated(bndList)) deallocate(bndList)
    call initBndryList()
    do l = 1, nlev
       call makeBndryList(l)
    end do
  end subroutine ref_hier_setup

  ! copy grid (level l, index k from 0) new (which=0) / old (1) to or from q
  subroutine ref_hier_xfer(l, k, which, nvar, q, to_ref) bind(C, name="ref_hier_xfer")
    integer(c_int), value :: l, k, which, nvar, to_ref
    real(c_double) :: q(*)
    integer :: mptr, loc, n
    mptr = listOfGrids(listStart(l) + k)
    loc = node(store1, mptr)
    if (which == 1) loc = node(store2, mptr)
    n = nvar*(node(ndihi,mptr)-node(ndilo,mptr)+1+2*nghost)*(node(ndjhi,mptr)-node(ndjlo,mptr)+1+2*nghost)
    if (to_ref /= 0) then
       alloc(loc:loc+n-1) = q(1:n)
    else
       q(1:n) = alloc(loc:loc+n-1)
    end if
  end subroutine ref_hier_xfer

  ! bound.f90:51 on grid k of level l at time t, in place on its new array
  subroutine ref_bound(l, k, t, nvar) bind(C, name="ref_bound")
    integer(c_int), value :: l, k, nvar
    real(c_double), value :: t
    integer :: mptr, mitot, mjtot
    mptr = listOfGrids(listStart(l) + k)
    mitot = node(ndihi,mptr)-node(ndilo,mptr)+1+2*nghost
    mjtot = node(ndjhi,mptr)-node(ndjlo,mptr)+1+2*nghost
    call bound(t, nvar, nghost, alloc(node(store1,mptr)), mitot, mjtot, mptr, &
         alloc(node(storeaux,mptr)), hnaux)
  end subroutine ref_bound

  ! prepf (fine levels) and prepc (coarse levels), as regrid.f:81-82 does
  subroutine ref_prep_regs(nlev, nvar) bind(C, name="ref_prep_regs")
    integer(c_int), value :: nlev, nvar
    integer :: l
    do l = 2, nlev
       call prepf(l, nvar, hnaux)
    end do
    do l = 1, nlev - 1
       call prepc(l, nvar)
    end do
  end subroutine ref_prep_regs

  ! advanc.f:10-142 (the loop is restated here because advanc.f needs
  ! gauges_module, which needs the absent clawpack utility_module); every
  ! routine it calls is the reference's own
  ! OpenMP over the grids of a level as advanc.f:60-84 and :106-129 do
  ! (off by default so that parity runs are deterministic)
  subroutine ref_set_parallel(on) bind(C, name="ref_set_parallel")
    integer(c_int), value :: on
    harness_parallel = on .ne. 0
  end subroutine ref_set_parallel

  subroutine ref_advanc(level, nvar, dtlevnew, cflout) bind(C, name="ref_advanc")
    integer(c_int), value :: level, nvar
    real(c_double) :: dtlevnew, cflout
    integer :: j, mptr, mitot, mjtot
    real(kind=8) :: dtnew
    !$omp parallel do if(harness_parallel) schedule(dynamic,1) private(j,mptr,mitot,mjtot)
    do j = 1, numgrids(level)
       mptr = listOfGrids(listStart(level)+j-1)
       mitot = node(ndihi,mptr) - node(ndilo,mptr) + 1 + 2*nghost
       mjtot = node(ndjhi,mptr) - node(ndjlo,mptr) + 1 + 2*nghost
       call bound(rnode(timemult,mptr), nvar, nghost, alloc(node(store1,mptr)), mitot, mjtot, &
            mptr, alloc(node(storeaux,mptr)), hnaux)
    end do
    !$omp end parallel do
    if (level+1 .le. mxnest) then
       if (lstart(level+1) .ne. null) call saveqc(level+1, nvar, hnaux)
    endif
    dtlevnew = rinfinity
    cfl_level = 0.d0
    !$omp parallel do if(harness_parallel) schedule(dynamic,1) &
    !$omp& private(j,mptr,mitot,mjtot,dtnew) reduction(min:dtlevnew)
    do j = 1, numgrids(level)
       mptr = listOfGrids(listStart(level)+j-1)
       mitot = node(ndihi,mptr) - node(ndilo,mptr) + 1 + 2*nghost
       mjtot = node(ndjhi,mptr) - node(ndjlo,mptr) + 1 + 2*nghost
       call harness_par_advanc(mptr, mitot, mjtot, nvar, dtnew)
       dtlevnew = dmin1(dtlevnew, dtnew)
    end do
    !$omp end parallel do
    cflout = cfl_level
  end subroutine ref_advanc

  ! advanc.f:170-292 without gauges, timers and dimensional splitting
  subroutine harness_par_advanc(mptr, mitot, mjtot, nvar, dtnew)
    integer :: mptr, mitot, mjtot, nvar
    real(kind=8) :: dtnew
    real(kind=8), allocatable :: fp(:,:,:), fm(:,:,:), gp(:,:,:), gm(:,:,:)
    integer :: level, nx, ny, locold, locnew, locaux, ntot, i, lenbc, locsvf, locsvq, locx1d
    real(kind=8) :: hx, hy, delt, time, xlow, ylow
    allocate(fp(nvar,mitot,mjtot), fm(nvar,mitot,mjtot), gp(nvar,mitot,mjtot), gm(nvar,mitot,mjtot))
    level = node(nestlevel,mptr)
    hx = hxposs(level)
    hy = hyposs(level)
    delt = possk(level)
    nx = node(ndihi,mptr) - node(ndilo,mptr) + 1
    ny = node(ndjhi,mptr) - node(ndjlo,mptr) + 1
    time = rnode(timemult,mptr)
    locold = node(store2, mptr)
    locnew = node(store1, mptr)
    if (level .lt. mxnest) then
       ntot = mitot * mjtot * nvar
       do i = 1, ntot
          alloc(locold + i - 1) = alloc(locnew + i - 1)
       end do
    endif
    xlow = rnode(cornxlo,mptr) - nghost*hx
    ylow = rnode(cornylo,mptr) - nghost*hy
    locaux = node(storeaux,mptr)
    if (node(ffluxptr,mptr) .ne. 0) then
       lenbc = 2*(nx/intratx(level-1)+ny/intraty(level-1))
       locsvf = node(ffluxptr,mptr)
       locsvq = locsvf + nvar*lenbc
       locx1d = locsvq + nvar*lenbc
       call qad(alloc(locnew),mitot,mjtot,nvar,alloc(locsvf),alloc(locsvq),lenbc, &
            intratx(level-1),intraty(level-1),hx,hy,hnaux,alloc(locaux),alloc(locx1d),delt,mptr)
    endif
    call stepgrid(alloc(locnew),fm,fp,gm,gp,mitot,mjtot,nghost,delt,dtnew,hx,hy,nvar, &
         xlow,ylow,time,mptr,hnaux,alloc(locaux))
    if (node(cfluxptr,mptr) .ne. 0) call fluxsv(mptr,fm,fp,gm,gp,alloc(node(cfluxptr,mptr)), &
         mitot,mjtot,nvar,listsp(level),delt,hx,hy)
    if (node(ffluxptr,mptr) .ne. 0) then
       lenbc = 2*(nx/intratx(level-1)+ny/intraty(level-1))
       locsvf = node(ffluxptr,mptr)
       call fluxad(fm,fp,gm,gp,alloc(locsvf),mptr,mitot,mjtot,nvar,lenbc, &
            intratx(level-1),intraty(level-1),nghost,delt,hx,hy)
    endif
    rnode(timemult,mptr) = rnode(timemult,mptr)+delt
    deallocate(fp, fm, gp, gm)
  end subroutine harness_par_advanc

  subroutine ref_update(level, nvar) bind(C, name="ref_update")
    integer(c_int), value :: level, nvar
    call update(level, nvar, hnaux)
  end subroutine ref_update

  ! register block (svdflx then qc1d) of grid k (0-based) of fine level l
  subroutine ref_get_reg(l, k, nvar, out) bind(C, name="ref_get_reg")
    integer(c_int), value :: l, k, nvar
    real(c_double) :: out(*)
    integer :: mptr, lenbc, nx, ny
    mptr = listOfGrids(listStart(l) + k)
    nx = node(ndihi,mptr) - node(ndilo,mptr) + 1
    ny = node(ndjhi,mptr) - node(ndjlo,mptr) + 1
    lenbc = 2*(nx/intratx(l-1)+ny/intraty(l-1))
    out(1:2*nvar*lenbc) = alloc(node(ffluxptr,mptr):node(ffluxptr,mptr)+2*nvar*lenbc-1)
  end subroutine ref_get_reg

  ! listbc of grid k (0-based) of coarse level l: returns the entry count
  ! and copies up to nmax entries (5 integers each)
  subroutine ref_get_listbc(l, k, nmax, out, n) bind(C, name="ref_get_listbc")
    integer(c_int), value :: l, k, nmax
    integer(c_int) :: out(*), n
    integer :: mptr
    external copy_listbc
    mptr = listOfGrids(listStart(l) + k)
    n = 0
    if (node(cfluxptr,mptr) .eq. 0) return
    call copy_listbc(alloc(node(cfluxptr,mptr)), listsp(l), nmax, out, n)
  end subroutine ref_get_listbc

  ! flush the reference's list-directed output (unit 6) so a caller that
  ! redirected the process's stdout around a call gets it all there
  subroutine ref_flush() bind(C, name="ref_flush")
    flush(6)
  end subroutine ref_flush

  subroutine ref_set_time(l, t) bind(C, name="ref_set_time")
    integer(c_int), value :: l
    real(c_double), value :: t
    integer :: j
    do j = 1, numgrids(l)
       rnode(timemult, listOfGrids(listStart(l) + j - 1)) = t
    end do
  end subroutine ref_set_time

  subroutine ref_get_time(l, k, t) bind(C, name="ref_get_time")
    integer(c_int), value :: l, k
    real(c_double) :: t
    t = rnode(timemult, listOfGrids(listStart(l) + k))
  end subroutine ref_get_time

  ! flagger.f:131-164 for one grid without Richardson estimation:
  ! flags = UNSET, flagregions2, flag2refine2 on q (ghost cells set)
  subroutine ref_flag_grid(mx, my, mbuff, meqn, xlower, ylower, dx, dy, t, level, tolsp, &
                           q, nreg, regs, flags) bind(C, name="ref_flag_grid")
    use regions_module, only: num_regions, regions
    integer(c_int), value :: mx, my, mbuff, meqn, level, nreg
    real(c_double), value :: xlower, ylower, dx, dy, t, tolsp
    real(c_double) :: q(meqn, 1-2:mx+2, 1-2:my+2), regs(8, *)
    real(c_double) :: flags(1-mbuff:mx+mbuff, 1-mbuff:my+mbuff)
    real(c_double) :: aux(1, 1-2:mx+2, 1-2:my+2)
    integer :: m
    if (allocated(regions)) deallocate(regions)
    num_regions = nreg
    allocate(regions(max(nreg, 1)))
    do m = 1, nreg
       regions(m)%min_level = int(regs(1, m))
       regions(m)%max_level = int(regs(2, m))
       regions(m)%t_low = regs(3, m)
       regions(m)%t_hi = regs(4, m)
       regions(m)%x_low = regs(5, m)
       regions(m)%x_hi = regs(6, m)
       regions(m)%y_low = regs(7, m)
       regions(m)%y_hi = regs(8, m)
    end do
    aux = 0.d0
    flags = UNSET
    call flagregions2(mx, my, mbuff, xlower, ylower, dx, dy, level, t, flags)
    call flag2refine2(mx, my, 2, mbuff, meqn, 0, xlower, ylower, dx, dy, t, level, tolsp, &
                      q, aux, flags)
  end subroutine ref_flag_grid

  ! flagger.f:31-180 for every grid of `level` (flagregions2, flag2refine2
  ! when fgrad, errest when frich) with the module's tolerances set from
  ! the arguments; the flags of the grids, each (nx+2 mbuff, ny+2 mbuff),
  ! are copied to out one after the other and their storage returned
  subroutine ref_flag_level(level, nvar, start_time, tolsp_in, tol_in, iorder_in, ibuff_in, &
                            fgrad, frich, nreg, regs, out) bind(C, name="ref_flag_level")
    use regions_module, only: num_regions, regions
    integer(c_int), value :: level, nvar, iorder_in, ibuff_in, fgrad, frich, nreg
    real(c_double), value :: start_time, tolsp_in, tol_in
    real(c_double) :: regs(8, *), out(*)
    integer :: j, m, mptr, nx, ny, mbuff, n, off, loc
    if (allocated(regions)) deallocate(regions)
    num_regions = nreg
    allocate(regions(max(nreg, 1)))
    do m = 1, nreg
       regions(m)%min_level = int(regs(1, m))
       regions(m)%max_level = int(regs(2, m))
       regions(m)%t_low = regs(3, m)
       regions(m)%t_hi = regs(4, m)
       regions(m)%x_low = regs(5, m)
       regions(m)%x_hi = regs(6, m)
       regions(m)%y_low = regs(7, m)
       regions(m)%y_hi = regs(8, m)
    end do
    tolsp = tolsp_in
    tol = tol_in
    iorder = iorder_in
    ibuff = ibuff_in
    flag_gradient = fgrad .ne. 0
    flag_richardson = frich .ne. 0
    eprint = .false.
    edebug = .false.
    call flagger(nvar, 0, level, start_time)
    mbuff = max(nghost, ibuff + 1)
    off = 0
    do j = 1, numgrids(level)
       mptr = listOfGrids(listStart(level) + j - 1)
       nx = node(ndihi, mptr) - node(ndilo, mptr) + 1
       ny = node(ndjhi, mptr) - node(ndjlo, mptr) + 1
       n = (nx + 2*mbuff) * (ny + 2*mbuff)
       loc = node(storeflags, mptr)
       out(off+1:off+n) = alloc(loc:loc+n-1)
       call reclam(loc, n)
       if (flag_richardson) call reclam(node(tempptr, mptr), (nx + 2*nghost)*(ny + 2*nghost)*nvar)
       off = off + n
    end do
  end subroutine ref_flag_level

  ! aux arrays of every installed grid (node(storeaux), naux values per
  ! cell like store1) and their types (amr_module auxtype: 0 "center",
  ! 1 "xleft", 2 "yleft", 3 "capacity"); call after ref_hier_setup and
  ! before ref_prep_regs
  subroutine ref_hier_set_aux(naux_in, types) bind(C, name="ref_hier_set_aux")
    integer(c_int), value :: naux_in
    integer(c_int), intent(in) :: types(*)
    integer :: ma, l, mptr, mitot, mjtot
    integer, external :: igetsp
    hnaux = naux_in
    if (allocated(auxtype)) deallocate(auxtype)
    allocate(auxtype(max(naux_in, 1)))
    auxtype = "center"
    do ma = 1, naux_in
       select case (types(ma))
       case (1)
          auxtype(ma) = "xleft"
       case (2)
          auxtype(ma) = "yleft"
       case (3)
          auxtype(ma) = "capacity"
       end select
    end do
    if (naux_in == 0) return
    do l = 1, maxlv
       if (numgrids(l) == 0) cycle
       mptr = lstart(l)
       do while (mptr /= 0)
          mitot = node(ndihi, mptr) - node(ndilo, mptr) + 1 + 2*nghost
          mjtot = node(ndjhi, mptr) - node(ndjlo, mptr) + 1 + 2*nghost
          node(storeaux, mptr) = igetsp(naux_in*mitot*mjtot)
          mptr = node(levelptr, mptr)
       end do
    end do
  end subroutine ref_hier_set_aux

  ! copy the aux array (naux, mitot, mjtot) of grid k (from 0) of level l in
  subroutine ref_hier_put_aux(l, k, aux) bind(C, name="ref_hier_put_aux")
    integer(c_int), value :: l, k
    real(c_double), intent(in) :: aux(*)
    integer :: mptr, n
    mptr = listOfGrids(listStart(l) + k)
    n = hnaux*(node(ndihi,mptr)-node(ndilo,mptr)+1+2*nghost)*(node(ndjhi,mptr)-node(ndjlo,mptr)+1+2*nghost)
    alloc(node(storeaux,mptr):node(storeaux,mptr)+n-1) = aux(1:n)
  end subroutine ref_hier_put_aux

  ! filval.f90:3 for a new grid of `level` with box (ilo, ihi, jlo, jhi):
  ! the installed grids of `level` are the old ones, those of level-1 the
  ! coarse ones; corners as grdfit2.f:98-101 sets them, mic / mjc as
  ! gfixup.f:83-84.  val (nvar, mitot, mjtot) is updated in place.
  subroutine ref_filval(level, box, t, nvar, val) bind(C, name="ref_filval")
    integer(c_int), value :: level, nvar
    integer(c_int), intent(in) :: box(4)
    real(c_double), value :: t
    real(c_double) :: val(*)
    integer :: ilo, ihi, jlo, jhi, nx, ny, mitot, mjtot, mic, mjc
    real(kind=8) :: xl, xr, yb, yt, aux(1)
    ilo = box(1)
    ihi = box(2)
    jlo = box(3)
    jhi = box(4)
    nx = ihi - ilo + 1
    ny = jhi - jlo + 1
    mitot = nx + 2*nghost
    mjtot = ny + 2*nghost
    mic = nx/intratx(level-1) + 2
    mjc = ny/intraty(level-1) + 2
    xl = ilo*hxposs(level) + xlower
    xr = (ihi+1)*hxposs(level) + xlower
    yb = jlo*hyposs(level) + ylower
    yt = (jhi+1)*hyposs(level) + ylower
    call filval(val, mitot, mjtot, hxposs(level), hyposs(level), level, t, mic, mjc, &
         xl, xr, yb, yt, nvar, 0, ilo, ihi, jlo, jhi, aux, 0)
  end subroutine ref_filval

  ! the neighbour list makeBndryList built for grid k of `level`
  ! (init_bndryList.f:50-107), in its linked-list order, as 0-based indices
  integer(c_int) function ref_bndry_list(level, k, out) bind(C, name="ref_bndry_list")
    integer(c_int), value :: level, k
    integer(c_int) :: out(*)
    integer :: mptr, i, spot
    mptr = listOfGrids(listStart(level) + k)
    spot = node(bndListSt, mptr)
    do i = 1, node(bndListNum, mptr)
       out(i) = bndList(spot, gridNbor) - listStart(level)
       spot = bndList(spot, nextfree)
    end do
    ref_bndry_list = node(bndListNum, mptr)
  end function ref_bndry_list

end module ref_harness

! listbc lives in the double heap as integers (prepc.f:55-59, setuse.f:95);
! this external routine reads it through an integer dummy like fluxsv does
subroutine copy_listbc(listbc, maxsp, nmax, out, n)
  implicit none
  integer :: maxsp, nmax, n, listbc(5, maxsp), out(5, *)
  integer :: e
  n = 0
  do e = 1, maxsp
     if (listbc(1, e) .eq. 0) exit
     n = n + 1
     if (n .le. nmax) out(:, n) = listbc(:, e)
  end do
end subroutine copy_listbc
