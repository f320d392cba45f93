! amrgpu3_level_shim.f90 -- ISO_C_BINDING boundary of the 3-D level update.
!
! Drop-in for src/3d/advanc.f:4 and src/3d/update.f:4 (same names and
! arguments; tick.f calls them unchanged): every level's grids live on the
! GPU (include/amrgpu3.h, hierarchy section) and a level step is one call of
! amrgpu3_advance_level (bound, saveqc, qad, stepgrid, fluxsv, fluxad), the
! coarse-fine update one call of amrgpu3_update_level (upbnd, averaging).
! The driver calls amrgpu3_upload_hierarchy after regridding (and in
! setgrd) and amrgpu3_download_hierarchy before anything on the host reads
! alloc (valout, check, regrid flagging).  Uses amrgpu3_iface from
! amrgpu3_shim.f90 for the configuration and the error check.
module amrgpu3_level_iface
  use iso_c_binding
  implicit none
  interface
     integer(c_int) function amrgpu3_set_geometry(nlev, iregsz, jregsz, kregsz, ratx, raty, &
          ratz, hx, hy, hz, domain, mthbc) bind(C, name="amrgpu3_set_geometry")
       import :: c_int, c_double
       integer(c_int), value :: nlev
       integer(c_int), intent(in) :: iregsz(*), jregsz(*), kregsz(*), ratx(*), raty(*), ratz(*)
       real(c_double), intent(in) :: hx(*), hy(*), hz(*), domain(6)
       integer(c_int), intent(in) :: mthbc(6)
     end function
     integer(c_int) function amrgpu3_set_level(level, ngrids, boxes) bind(C, name="amrgpu3_set_level")
       import :: c_int
       integer(c_int), value :: level, ngrids
       integer(c_int), intent(in) :: boxes(*)
     end function
     integer(c_int) function amrgpu3_set_time(level, time, possk) bind(C, name="amrgpu3_set_time")
       import :: c_int, c_double
       integer(c_int), value :: level
       real(c_double), value :: time, possk
     end function
     integer(c_int) function amrgpu3_put_level(level, which, alloc_base, offsets) &
          bind(C, name="amrgpu3_put_level")
       import :: c_int, c_double, c_int64_t
       integer(c_int), value :: level, which
       real(c_double), intent(in) :: alloc_base(*)
       integer(c_int64_t), intent(in) :: offsets(*)
     end function
     integer(c_int) function amrgpu3_get_level(level, which, alloc_base, offsets) &
          bind(C, name="amrgpu3_get_level")
       import :: c_int, c_double, c_int64_t
       integer(c_int), value :: level, which
       real(c_double) :: alloc_base(*)
       integer(c_int64_t), intent(in) :: offsets(*)
     end function
     integer(c_int) function amrgpu3_advance_level(level, dtlevnew, cfl_level) &
          bind(C, name="amrgpu3_advance_level")
       import :: c_int, c_double
       integer(c_int), value :: level
       real(c_double) :: dtlevnew, cfl_level
     end function
     integer(c_int) function amrgpu3_update_level(level) bind(C, name="amrgpu3_update_level")
       import :: c_int
       integer(c_int), value :: level
     end function
     integer(c_int) function amrgpu3_conck_level(level, totmass) bind(C, name="amrgpu3_conck_level")
       import :: c_int, c_double
       integer(c_int), value :: level
       real(c_double) :: totmass
     end function
     integer(c_int) function amrgpu3_flag_level(level, tolsp, flag_richardson, tol, iorder, &
          nflag, nrich) bind(C, name="amrgpu3_flag_level")
       import :: c_int, c_double, c_int64_t
       integer(c_int), value :: level, flag_richardson, iorder
       real(c_double), value :: tolsp, tol
       integer(c_int64_t) :: nflag, nrich
     end function
     integer(c_int) function amrgpu3_get_level_flags(level, grid, amrflags, rflags) &
          bind(C, name="amrgpu3_get_level_flags")
       import :: c_int, c_double
       integer(c_int), value :: level, grid
       real(c_double) :: amrflags(*), rflags(*)
     end function
  end interface
end module amrgpu3_level_iface

! After regrid and in setgrd: hand every level's grids, both time levels and
! the times to the device.
subroutine amrgpu3_upload_hierarchy(nvar)
  use amr_module
  use amrgpu3_iface, only: amrgpu3_check
  use amrgpu3_level_iface
  implicit none
  integer, intent(in) :: nvar
  integer :: level, j, mptr, n
  integer(c_int), allocatable :: boxes(:)
  integer(c_int64_t), allocatable :: off1(:), off2(:)
  integer(c_int) :: isz(maxlv), jsz(maxlv), ksz(maxlv), rx(maxlv), ry(maxlv), rz(maxlv)
  integer(c_int) :: bc(6)
  real(c_double) :: dom(6)
  do level = 1, mxnest
     isz(level) = iregsz(level)
     jsz(level) = jregsz(level)
     ksz(level) = kregsz(level)
     ! the finest level has no ratio to a finer one (amr3.f90 reads the
     ! ratios of levels 1..mxnest-1 only)
     rx(level) = 1
     ry(level) = 1
     rz(level) = 1
     if (level .lt. mxnest) then
        rx(level) = intratx(level)
        ry(level) = intraty(level)
        rz(level) = intratz(level)
     end if
  end do
  dom = (/ xlower, xupper, ylower, yupper, zlower, zupper /)
  bc = mthbc(1:6)
  call amrgpu3_check(amrgpu3_set_geometry(mxnest, isz, jsz, ksz, rx, ry, rz, hxposs, hyposs, &
       hzposs, dom, bc), 'set_geometry')
  do level = 1, mxnest
     n = numgrids(level)
     allocate(boxes(6 * max(n, 1)), off1(max(n, 1)), off2(max(n, 1)))
     do j = 1, n
        mptr = listOfGrids(listStart(level) + j - 1)
        boxes(6*j-5:6*j) = (/ node(ndilo, mptr), node(ndihi, mptr), node(ndjlo, mptr), &
             node(ndjhi, mptr), node(ndklo, mptr), node(ndkhi, mptr) /)
        off1(j) = node(store1, mptr) - 1          ! 0-based offsets into alloc
        off2(j) = node(store2, mptr) - 1
     end do
     call amrgpu3_check(amrgpu3_set_level(level, n, boxes), 'set_level')
     if (n .gt. 0) then
        call amrgpu3_check(amrgpu3_put_level(level, 0, alloc, off1), 'put_level new')
        if (level .lt. mxnest) call amrgpu3_check(amrgpu3_put_level(level, 1, alloc, off2), &
             'put_level old')
        call amrgpu3_check(amrgpu3_set_time(level, rnode(timemult, listOfGrids(listStart(level))), &
             possk(level)), 'set_time')
     end if
     deallocate(boxes, off1, off2)
  end do
end subroutine amrgpu3_upload_hierarchy

! Before valout, check or regrid flagging: bring the new (and old) values of
! every level back into alloc.
subroutine amrgpu3_download_hierarchy(nvar)
  use amr_module
  use amrgpu3_iface, only: amrgpu3_check
  use amrgpu3_level_iface
  implicit none
  integer, intent(in) :: nvar
  integer :: level, j, n
  integer(c_int64_t), allocatable :: off1(:), off2(:)
  do level = 1, mxnest
     n = numgrids(level)
     if (n .eq. 0) cycle
     allocate(off1(n), off2(n))
     do j = 1, n
        off1(j) = node(store1, listOfGrids(listStart(level) + j - 1)) - 1
        off2(j) = node(store2, listOfGrids(listStart(level) + j - 1)) - 1
     end do
     call amrgpu3_check(amrgpu3_get_level(level, 0, alloc, off1), 'get_level new')
     if (level .lt. mxnest) call amrgpu3_check(amrgpu3_get_level(level, 1, alloc, off2), &
          'get_level old')
     deallocate(off1, off2)
  end do
end subroutine amrgpu3_download_hierarchy

! src/3d/advanc.f:4 -- same name and arguments.
subroutine advanc(level, nvar, dtlevnew, vtime, naux)
  use amr_module
  use amrgpu3_iface, only: amrgpu3_check
  use amrgpu3_level_iface
  implicit none
  integer, intent(in) :: level, nvar, naux
  real(kind=8), intent(out) :: dtlevnew
  logical, intent(in) :: vtime
  real(c_double) :: dtn, cfl_dev
  integer :: j, mptr, ncell
  call amrgpu3_check(amrgpu3_advance_level(level, dtn, cfl_dev), 'advanc')
  dtlevnew = dtn
  cfl_level = cfl_dev
  cflmax = dmax1(cflmax, cfl_level)
  do j = 1, numgrids(level)
     mptr = listOfGrids(listStart(level) + j - 1)
     rnode(timemult, mptr) = rnode(timemult, mptr) + possk(level)
     ncell = (node(ndihi,mptr) - node(ndilo,mptr) + 1) * (node(ndjhi,mptr) - node(ndjlo,mptr) + 1) &
          * (node(ndkhi,mptr) - node(ndklo,mptr) + 1)
     rvol = rvol + ncell
     rvoll(level) = rvoll(level) + ncell
  end do
end subroutine advanc

! src/3d/update.f:4 -- same name and arguments.
subroutine update(level, nvar, maux)
  use amrgpu3_iface, only: amrgpu3_check
  use amrgpu3_level_iface
  implicit none
  integer, intent(in) :: level, nvar, maux
  call amrgpu3_check(amrgpu3_update_level(level), 'update')
end subroutine update

! Replaces, in flglvl.f (3-D):64-67, the calls of spest and errest: the
! device flags every grid of lcheck from its resident solution (spest: bound
! at the grid's time and flag2refine; errest: copysol, bound with 2 nghost at
! time - possk, coarsen, the two stepgrids, errf1).  As spest does, the
! gradient flags go into dom1flags through setflags and a locbig array is
! left in node(tempptr) for bufnst to reclaim; as errf1 does, the
! Richardson flags go into component 1 of store2, where bufnst reads them.
! The device's grid j-1 of lcheck is listOfGrids(listStart(lcheck)+j-1),
! the order amrgpu3_upload_hierarchy uploads them in.
subroutine amrgpu3_spest_errest(nvar, naux, lcheck, dom1flags, isize, jsize, ksize)
  use amr_module
  use amrgpu3_iface, only: amrgpu3_check
  use amrgpu3_level_iface
  implicit none
  integer, intent(in) :: nvar, naux, lcheck, isize, jsize, ksize
  integer*1 :: dom1flags(isize+2, jsize+2, ksize+2)
  integer :: mptr, j, mitot, mjtot, mktot, n, i, loc2
  integer(c_int64_t) :: nflag, nrich
  real(c_double), allocatable :: fl(:), rf(:)
  integer, external :: igetsp
  call amrgpu3_check(amrgpu3_flag_level(lcheck, tolsp, merge(1, 0, flag_richardson), tol, &
       iorder, nflag, nrich), 'flag_level')
  do j = 1, numgrids(lcheck)
     mptr = listOfGrids(listStart(lcheck) + j - 1)
     mitot = node(ndihi, mptr) - node(ndilo, mptr) + 1 + 2 * nghost
     mjtot = node(ndjhi, mptr) - node(ndjlo, mptr) + 1 + 2 * nghost
     mktot = node(ndkhi, mptr) - node(ndklo, mptr) + 1 + 2 * nghost
     n = mitot * mjtot * mktot
     allocate(fl(n), rf(n))
     call amrgpu3_check(amrgpu3_get_level_flags(lcheck, j - 1, fl, rf), 'get_level_flags')
     node(tempptr, mptr) = igetsp(n * nvar)
     if (flag_gradient) call setflags(dom1flags, isize, jsize, ksize, fl, 1, mitot, mjtot, &
          mktot, mptr)
     if (flag_richardson) then
        loc2 = node(store2, mptr)
        do i = 1, n
           alloc(loc2 + (i - 1) * nvar) = rf(i)
        end do
     end if
     deallocate(fl, rf)
  end do
end subroutine amrgpu3_spest_errest
