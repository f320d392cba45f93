! amrgpu_shim.f90 -- Fortran side of the drop-in boundary: ISO_C_BINDING
! interfaces of include/amrgpu.h and replacements for the reference's
! advanc (advanc.f:10) and update (update.f:16) with the same names and
! arguments.  Compiled against the reference's amr_module; linked with
! libamrgpu.so in place of advanc.f and update.f (Makefile.amr_2d
! EXCLUDE_SOURCES).  The reference's driver (amr2, tick, setgrd, regrid,
! flagging, valout, conck, gauges output) runs unmodified:
!   * the first call initialises the library from amr_module and the
!     application's constants (amrgpu_app_physics, supplied by the
!     application next to its setprob and Riemann solvers);
!   * every call compares the hierarchy (grids, boxes and heap locations of
!     each level) with what the device holds and hands over the levels that
!     regrid or setgrd changed (the device keeps the registers of the
!     others, as the reference keeps their node(ffluxptr) blocks);
!   * the values the device computes (solutions and coarse-fine registers)
!     are written back into alloc at once, so that everything the driver
!     reads between steps (flagging, valout, conck, checkpoints, gauges)
!     sees the reference's heap as it would be.
module amrgpu_iface
  use iso_c_binding
  implicit none

  type, bind(C) :: amrgpu_config
     integer(c_int) :: solver, nvar, mwaves, naux, nghost
     integer(c_int) :: method(7), mthlim(8), mcapa, use_fwaves
     real(c_double) :: cfl, cflv1
     integer(c_int) :: mthbc(4)
     real(c_double) :: xlower, xupper, ylower, yupper
     real(c_double) :: ubar, vbar, rho, bulk, gamma
     integer(c_int) :: efix, precision
     integer(c_int) :: auxtype(8)
  end type amrgpu_config

  interface
     integer(c_int) function amrgpu_init(cfg) bind(C, name="amrgpu_init")
       import :: c_int, amrgpu_config
       type(amrgpu_config), intent(in) :: cfg
     end function
     integer(c_int) function amrgpu_last_error(buf, len) bind(C, name="amrgpu_last_error")
       import :: c_int, c_char
       character(kind=c_char) :: buf(*)
       integer(c_int), value :: len
     end function
     integer(c_int) function amrgpu_set_geometry(mxnest, iregsz, jregsz, intratx, intraty, &
          hxposs, hyposs) bind(C, name="amrgpu_set_geometry")
       import :: c_int, c_double
       integer(c_int), value :: mxnest
       integer(c_int), intent(in) :: iregsz(*), jregsz(*), intratx(*), intraty(*)
       real(c_double), intent(in) :: hxposs(*), hyposs(*)
     end function
     integer(c_int) function amrgpu_set_level(level, ngrids, ilo, ihi, jlo, jhi, hx, hy) &
          bind(C, name="amrgpu_set_level")
       import :: c_int, c_double
       integer(c_int), value :: level, ngrids
       integer(c_int), intent(in) :: ilo(*), ihi(*), jlo(*), jhi(*)
       real(c_double), value :: hx, hy
     end function
     integer(c_int) function amrgpu_set_corners(level, xlo, ylo) bind(C, name="amrgpu_set_corners")
       import :: c_int, c_double
       integer(c_int), value :: level
       real(c_double), intent(in) :: xlo(*), ylo(*)
     end function
     integer(c_int) function amrgpu_put_level(level, which, alloc_base, offsets) &
          bind(C, name="amrgpu_put_level")
       import :: c_int, c_double, c_int64_t
       integer(c_int), value :: level, which
       real(c_double), intent(in) :: alloc_base(*)
       integer(c_int64_t), intent(in) :: offsets(*)
     end function
     integer(c_int) function amrgpu_put_level_aux(level, alloc_base, offsets) &
          bind(C, name="amrgpu_put_level_aux")
       import :: c_int, c_double, c_int64_t
       integer(c_int), value :: level
       real(c_double), intent(in) :: alloc_base(*)
       integer(c_int64_t), intent(in) :: offsets(*)
     end function
     integer(c_int) function amrgpu_get_level(level, which, alloc_base, offsets) &
          bind(C, name="amrgpu_get_level")
       import :: c_int, c_double, c_int64_t
       integer(c_int), value :: level, which
       real(c_double) :: alloc_base(*)
       integer(c_int64_t), intent(in) :: offsets(*)
     end function
     integer(c_int) function amrgpu_bound_level(level, time) bind(C, name="amrgpu_bound_level")
       import :: c_int, c_double
       integer(c_int), value :: level
       real(c_double), value :: time
     end function
     integer(c_int) function amrgpu_set_time(level, time, possk) bind(C, name="amrgpu_set_time")
       import :: c_int, c_double
       integer(c_int), value :: level
       real(c_double), value :: time, possk
     end function
     integer(c_int) function amrgpu_advance_level(level, dtlevnew, cfl_level) &
          bind(C, name="amrgpu_advance_level")
       import :: c_int, c_double
       integer(c_int), value :: level
       real(c_double) :: dtlevnew, cfl_level
     end function
     integer(c_int) function amrgpu_get_registers(level, grid, svdflx_qc1d, lenbc) &
          bind(C, name="amrgpu_get_registers")
       import :: c_int, c_double
       integer(c_int), value :: level, grid
       real(c_double) :: svdflx_qc1d(*)
       integer(c_int) :: lenbc
     end function
     integer(c_int) function amrgpu_set_listbc(level, grid, n, list5) &
          bind(C, name="amrgpu_set_listbc")
       import :: c_int
       integer(c_int), value :: level, grid, n
       integer(c_int), intent(in) :: list5(*)
     end function
     integer(c_int) function amrgpu_set_gauges(n, x, y, t_start, t_end, min_dt) &
          bind(C, name="amrgpu_set_gauges")
       import :: c_int, c_double
       integer(c_int), value :: n
       real(c_double), intent(in) :: x(*), y(*), t_start(*), t_end(*), min_dt(*)
     end function
     integer(c_int) function amrgpu_set_gauge_sources(level, grid) &
          bind(C, name="amrgpu_set_gauge_sources")
       import :: c_int
       integer(c_int), intent(in) :: level(*), grid(*)
     end function
     integer(c_int) function amrgpu_get_gauge_records(maxrec, nrec, gauge, level, vals) &
          bind(C, name="amrgpu_get_gauge_records")
       import :: c_int, c_double
       integer(c_int), value :: maxrec
       integer(c_int) :: nrec, gauge(*), level(*)
       real(c_double) :: vals(*)
     end function
     integer(c_int) function amrgpu_update_level(level) bind(C, name="amrgpu_update_level")
       import :: c_int
       integer(c_int), value :: level
     end function
     integer(c_int) function amrgpu_conck(level, totmass) bind(C, name="amrgpu_conck")
       import :: c_int, c_double
       integer(c_int), value :: level
       real(c_double) :: totmass
     end function
     integer(c_int) function amrgpu_flag_level(level, flag_gradient, tolsp, flag_richardson, &
          tol, iorder, start_time, nregions, regions, mbuff, flags_base, offsets, nflagged) &
          bind(C, name="amrgpu_flag_level")
       import :: c_int, c_double, c_int64_t
       integer(c_int), value :: level, flag_gradient, flag_richardson, iorder, nregions, mbuff
       real(c_double), value :: tolsp, tol, start_time
       real(c_double), intent(in) :: regions(*)
       real(c_double) :: flags_base(*)
       integer(c_int64_t), intent(in) :: offsets(*)
       integer(c_int64_t) :: nflagged
     end function
  end interface

contains

  ! nonzero status -> message and stop, the reference's error convention
  subroutine amrgpu_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    integer :: i, k
    character(len=512) :: msg
    if (rc .eq. 0) return
    k = amrgpu_last_error(buf, 512)
    msg = ' '
    do i = 1, 512
       if (buf(i) .eq. c_null_char) exit
       msg(i:i) = buf(i)
    end do
    write(*,*) '*** ', where, ': ', trim(msg)
    stop
  end subroutine amrgpu_check

end module amrgpu_iface

! The state the device holds: whether the library is initialised, and per
! level the layout it was handed (number of grids, then per grid its box and
! its store1 / store2 heap locations).
module amrgpu_coherence
  use iso_c_binding
  use amr_module, only: maxlv
  implicit none
  logical :: dev_ready = .false.
  logical :: gauges_ready = .false.
  type level_layout
     integer, allocatable :: v(:)
  end type level_layout
  type(level_layout) :: held(maxlv)
  ! per level the listbc (prepc.f) the device was handed: per grid its
  ! entry count then its entries with the kid as an index on the finer level
  type(level_layout) :: lists(maxlv)
contains
  ! The layout of one level as the reference's node table describes it.
  subroutine amrgpu_level_layout(level, sig, n)
    use amr_module, only: numgrids, listOfGrids, listStart, node, ndilo, ndihi, ndjlo, &
         ndjhi, store1, store2
    integer, intent(in) :: level
    integer, intent(out) :: n
    integer, allocatable, intent(out) :: sig(:)
    integer :: j, mptr
    n = numgrids(level)
    allocate(sig(1 + 6 * n))
    sig(1) = n
    do j = 1, n
       mptr = listOfGrids(listStart(level) + j - 1)
       sig(6 * j - 4) = node(ndilo, mptr)
       sig(6 * j - 3) = node(ndihi, mptr)
       sig(6 * j - 2) = node(ndjlo, mptr)
       sig(6 * j - 1) = node(ndjhi, mptr)
       sig(6 * j) = node(store1, mptr)
       sig(6 * j + 1) = node(store2, mptr)
    end do
  end subroutine amrgpu_level_layout
  ! The listbc of every grid of `level` (node(cfluxptr), listsp(level)
  ! entries of five ints, zero-terminated, prepc.f:47-154), flattened: per
  ! grid the entry count, then (i, j, side, kid index on level+1, slot).
  subroutine amrgpu_level_lists(level, v)
    use amr_module, only: numgrids, listOfGrids, listStart, node, cfluxptr, listsp, alloc
    integer, intent(in) :: level
    integer, allocatable, intent(out) :: v(:)
    integer, allocatable :: tmp(:), one(:)
    integer :: j, mptr, n, k, pos, cap, nd, loc
    cap = 1 + numgrids(level) * (1 + 5 * max(listsp(level), 0))
    allocate(tmp(cap))
    allocate(one(5 * max(listsp(level), 1)))
    pos = 0
    do j = 1, numgrids(level)
       mptr = listOfGrids(listStart(level) + j - 1)
       n = 0
       if (node(cfluxptr, mptr) .ne. 0 .and. listsp(level) .gt. 0) then
          ! listbc(5, listsp) lives in alloc as integers (prepc.f:55-59;
          ! fluxsv.f and upbnd.f take it as an integer array): its entries up
          ! to the first zero cell index (upbnd.f:39-40)
          loc = node(cfluxptr, mptr)
          nd = (5 * listsp(level) + 1) / 2
          one = transfer(alloc(loc:loc + nd - 1), one, 5 * listsp(level))
          do k = 1, listsp(level)
             if (one(5 * k - 4) .eq. 0) exit
             n = k
          end do
       end if
       pos = pos + 1
       tmp(pos) = n
       do k = 1, n
          tmp(pos + 1:pos + 5) = one(5 * k - 4:5 * k)
          tmp(pos + 4) = amrgpu_kid_index(level + 1, one(5 * k - 1))
          pos = pos + 5
       end do
    end do
    allocate(v(pos))
    v(1:pos) = tmp(1:pos)
    deallocate(tmp, one)
  end subroutine amrgpu_level_lists

  ! the 0-based index of grid mkid in level l's list
  integer function amrgpu_kid_index(l, mkid)
    use amr_module, only: numgrids, listOfGrids, listStart
    integer, intent(in) :: l, mkid
    integer :: i
    amrgpu_kid_index = -1
    do i = 1, numgrids(l)
         if (listOfGrids(listStart(l) + i - 1) .eq. mkid) then
            amrgpu_kid_index = i - 1
            return
         end if
    end do
  end function amrgpu_kid_index

  ! Hand a level's lists (as amrgpu_level_lists flattens them) to the device.
  subroutine amrgpu_send_lists(level, v)
    use amr_module, only: numgrids
    use amrgpu_iface
    integer, intent(in) :: level
    integer, intent(in) :: v(:)
    integer :: j, n, pos
    pos = 1
    do j = 1, numgrids(level)
       n = v(pos)
       call amrgpu_check(amrgpu_set_listbc(level, j - 1, n, v(pos + 1:pos + 5 * n)), 'set_listbc')
       pos = pos + 1 + 5 * n
    end do
  end subroutine amrgpu_send_lists
end module amrgpu_coherence

! Initialise the library on first use; afterwards hand over every level
! whose layout differs from the one the device holds: its grids
! (amrgpu_set_level, set_corners), both time levels from alloc
! (store2 only below mxnest, where the reference keeps one) and its time.
subroutine amrgpu_sync_hierarchy(nvar, naux)
  use amr_module
  use amrgpu_iface
  use amrgpu_coherence
  implicit none
  integer, intent(in) :: nvar, naux
  type(amrgpu_config) :: cfg
  integer :: level, j, mptr, n, mw
  integer, allocatable :: sig(:)
  integer(c_int), allocatable :: ilo(:), ihi(:), jlo(:), jhi(:)
  integer(c_int64_t), allocatable :: off1(:), off2(:), offa(:)
  real(c_double), allocatable :: cx(:), cy(:)
  integer(c_int) :: isz(maxlv), jsz(maxlv), rx(maxlv), ry(maxlv)
  if (.not. dev_ready) then
     cfg%nvar = nvar
     cfg%mwaves = mwaves
     cfg%naux = naux
     cfg%nghost = nghost
     cfg%method = method(1:7)
     cfg%mthlim = 0
     do mw = 1, min(mwaves, 8)
        cfg%mthlim(mw) = mthlim(mw)
     end do
     cfg%mcapa = mcapa
     ! amr.data's aux_type strings (amr2.f90:346-352) as the library's codes
     cfg%auxtype = 0
     do mw = 1, min(naux, 8)
        select case (trim(auxtype(mw)))
        case ('xleft')
           cfg%auxtype(mw) = 1
        case ('yleft')
           cfg%auxtype(mw) = 2
        case ('capacity')
           cfg%auxtype(mw) = 3
        end select
     end do
     cfg%use_fwaves = 0
     if (use_fwaves) cfg%use_fwaves = 1
     cfg%cfl = cfl
     cfg%cflv1 = cflv1
     cfg%mthbc = mthbc(1:4)
     cfg%xlower = xlower
     cfg%xupper = xupper
     cfg%ylower = ylower
     cfg%yupper = yupper
     ! the application's solver and its /cparam/ constants
     call amrgpu_app_physics(cfg%solver, cfg%ubar, cfg%vbar, cfg%rho, cfg%bulk, cfg%gamma, &
          cfg%efix, cfg%precision)
     call amrgpu_check(amrgpu_init(cfg), 'init')
     dev_ready = .true.
  end if
  do level = 1, mxnest
     isz(level) = iregsz(level)
     jsz(level) = jregsz(level)
     ! the finest level has no ratio to a finer one (amr2.f90 reads
     ! intratx(1:mxnest-1) only)
     rx(level) = 1
     ry(level) = 1
     if (level .lt. mxnest) then
        rx(level) = intratx(level)
        ry(level) = intraty(level)
     end if
  end do
  ! a no-op when the geometry is the one the device has
  call amrgpu_check(amrgpu_set_geometry(mxnest, isz, jsz, rx, ry, hxposs, hyposs), 'set_geometry')
  do level = 1, mxnest
     call amrgpu_level_layout(level, sig, n)
     if (allocated(held(level)%v)) then
        if (size(held(level)%v) == size(sig)) then
           if (all(held(level)%v == sig)) cycle
        end if
     end if
     allocate(ilo(max(n,1)), ihi(max(n,1)), jlo(max(n,1)), jhi(max(n,1)))
     allocate(off1(max(n,1)), off2(max(n,1)), offa(max(n,1)), cx(max(n,1)), cy(max(n,1)))
     do j = 1, n
        mptr = listOfGrids(listStart(level) + j - 1)
        ilo(j) = node(ndilo, mptr)
        ihi(j) = node(ndihi, mptr)
        jlo(j) = node(ndjlo, mptr)
        jhi(j) = node(ndjhi, mptr)
        cx(j) = rnode(cornxlo, mptr)
        cy(j) = rnode(cornylo, mptr)
        off1(j) = node(store1, mptr) - 1          ! 0-based offsets into alloc
        off2(j) = node(store2, mptr) - 1
        if (naux .gt. 0) offa(j) = node(storeaux, mptr) - 1
     end do
     call amrgpu_check(amrgpu_set_level(level, n, ilo, ihi, jlo, jhi, hxposs(level), &
          hyposs(level)), 'set_level')
     if (n .gt. 0) then
        call amrgpu_check(amrgpu_set_corners(level, cx, cy), 'set_corners')
        call amrgpu_check(amrgpu_put_level(level, 0, alloc, off1), 'put_level new')
        if (level .lt. mxnest) call amrgpu_check(amrgpu_put_level(level, 1, alloc, off2), &
             'put_level old')
        if (naux .gt. 0) call amrgpu_check(amrgpu_put_level_aux(level, alloc, offa), &
             'put_level_aux')
        call amrgpu_check(amrgpu_set_time(level, rnode(timemult, listOfGrids(listStart(level))), &
             possk(level)), 'set_time')
     end if
     deallocate(ilo, ihi, jlo, jhi, off1, off2, offa, cx, cy)
     call move_alloc(sig, held(level)%v)
  end do
  ! the coarse side of the registers as prepc.f built it (its kid order is
  ! regrid's, before arrangeGrids sorts the grids): hand over every level
  ! whose lists differ from what the device holds
  do level = 1, mxnest - 1
     call amrgpu_level_lists(level, sig)
     if (allocated(lists(level)%v)) then
        if (size(lists(level)%v) == size(sig)) then
           if (all(lists(level)%v == sig)) cycle
        end if
     end if
     call amrgpu_send_lists(level, sig)
     call move_alloc(sig, lists(level)%v)
  end do
end subroutine amrgpu_sync_hierarchy


! The gauges of gauges.data (set_gauges has read them) to the device.
subroutine amrgpu_install_gauges()
  use gauges_module, only: num_gauges, gauges
  use amrgpu_iface
  use amrgpu_coherence, only: gauges_ready
  implicit none
  real(c_double), allocatable :: x(:), y(:), t0(:), t1(:), dtm(:)
  integer :: i
  ! amr2.f90 sets num_gauges to 0 while setgrd builds the first grids:
  ! install them at the first step that samples them
  if (gauges_ready .or. num_gauges .eq. 0) return
  gauges_ready = .true.
  ! the device samples q only (gauges_module.f90:409-417 samples aux
  ! components too when gauges.data asks for them)
  do i = 1, num_gauges
     if (allocated(gauges(i)%aux_out_vars)) then
        if (any(gauges(i)%aux_out_vars)) then
           print *, 'amrgpu: gauge output of aux components is not sampled on the device'
           stop 1
        end if
     end if
  end do
  allocate(x(num_gauges), y(num_gauges), t0(num_gauges), t1(num_gauges), dtm(num_gauges))
  do i = 1, num_gauges
     x(i) = gauges(i)%x
     y(i) = gauges(i)%y
     t0(i) = gauges(i)%t_start
     t1(i) = gauges(i)%t_end
     dtm(i) = gauges(i)%min_time_increment
  end do
  call amrgpu_check(amrgpu_set_gauges(num_gauges, x, y, t0, t1, dtm), 'set_gauges')
  deallocate(x, y, t0, t1, dtm)
end subroutine amrgpu_install_gauges

! Each gauge's best source grid (setbestsrc's mbestsrc, tick.f:203) as a
! level and an index in the order the device's levels were handed over.
subroutine amrgpu_gauge_sources()
  use amr_module
  use gauges_module, only: num_gauges, mbestsrc
  use amrgpu_iface
  implicit none
  integer(c_int), allocatable :: lev(:), grd(:)
  integer :: i, j, mptr, l
  if (num_gauges .eq. 0) return
  allocate(lev(num_gauges), grd(num_gauges))
  do i = 1, num_gauges
     lev(i) = 0
     grd(i) = -1
     mptr = mbestsrc(i)
     if (mptr .le. 0) cycle
     l = node(nestlevel, mptr)
     do j = 1, numgrids(l)
        if (listOfGrids(listStart(l) + j - 1) .eq. mptr) then
           lev(i) = l
           grd(i) = j - 1
           exit
        end if
     end do
  end do
  call amrgpu_check(amrgpu_set_gauge_sources(lev, grd), 'set_gauge_sources')
  deallocate(lev, grd)
end subroutine amrgpu_gauge_sources

! The records the device sampled during a level step, appended to the
! gauges' buffers as update_gauges appends its own (gauges_module.f90:
! 414-437), flushed to the gauge files when full.
subroutine amrgpu_collect_gauges(nvar)
  use gauges_module, only: num_gauges, gauges, MAX_BUFFER, print_gauges_and_reset_nextLoc
  use amrgpu_iface
  implicit none
  integer, intent(in) :: nvar
  integer(c_int) :: nrec
  integer(c_int), allocatable :: gid(:), glev(:)
  real(c_double), allocatable :: vals(:)
  integer :: r, i, n, k, index
  if (num_gauges .eq. 0) return
  allocate(gid(num_gauges), glev(num_gauges), vals(num_gauges * (nvar + 1)))
  call amrgpu_check(amrgpu_get_gauge_records(num_gauges, nrec, gid, glev, vals), 'gauge records')
  do r = 1, nrec
     i = gid(r) + 1
     index = gauges(i)%buffer_index
     gauges(i)%level(index) = glev(r)
     gauges(i)%data(1, index) = vals((r - 1) * (nvar + 1) + 1)
     k = 0
     do n = 1, nvar
        if (gauges(i)%q_out_vars(n)) then
           k = k + 1
           gauges(i)%data(1 + k, index) = vals((r - 1) * (nvar + 1) + 1 + n)
        end if
     end do
     gauges(i)%buffer_index = index + 1
     if (gauges(i)%buffer_index > MAX_BUFFER) call print_gauges_and_reset_nextLoc(i)
     gauges(i)%last_time = vals((r - 1) * (nvar + 1) + 1)
  end do
  deallocate(gid, glev, vals)
end subroutine amrgpu_collect_gauges

! Write the device's values of one level back into alloc: the new time
! (store1, ghost cells included) and, with old, the old time (store2).
subroutine amrgpu_pull_level(level, old)
  use amr_module
  use amrgpu_iface
  implicit none
  integer, intent(in) :: level
  logical, intent(in) :: old
  integer :: j, n
  integer(c_int64_t), allocatable :: off(:)
  n = numgrids(level)
  if (n .eq. 0) return
  allocate(off(n))
  do j = 1, n
     off(j) = node(store1, listOfGrids(listStart(level) + j - 1)) - 1
  end do
  call amrgpu_check(amrgpu_get_level(level, 0, alloc, off), 'get_level new')
  if (old) then
     do j = 1, n
        off(j) = node(store2, listOfGrids(listStart(level) + j - 1)) - 1
     end do
     call amrgpu_check(amrgpu_get_level(level, 1, alloc, off), 'get_level old')
  end if
  deallocate(off)
end subroutine amrgpu_pull_level

! Write the coarse-fine registers of the grids of fine level `level` back
! into their node(ffluxptr) blocks (svdflx then qc1d, prepf.f:33-48), so
! that a checkpoint (check.f writes all of alloc) holds what the
! reference's would.
subroutine amrgpu_pull_registers(level)
  use amr_module
  use amrgpu_iface
  implicit none
  integer, intent(in) :: level
  integer :: j, mptr
  integer(c_int) :: lenbc
  if (level .lt. 2 .or. level .gt. mxnest) return
  do j = 1, numgrids(level)
     mptr = listOfGrids(listStart(level) + j - 1)
     if (node(ffluxptr, mptr) .eq. 0) cycle
     call amrgpu_check(amrgpu_get_registers(level, j - 1, alloc(node(ffluxptr, mptr)), lenbc), &
          'get_registers')
  end do
end subroutine amrgpu_pull_registers

! advanc.f:10 -- same name and arguments; tick.f:240 calls it unchanged.
subroutine advanc(level, nvar, dtlevnew, vtime, naux)
  use amr_module
  use amrgpu_iface
  implicit none
  integer, intent(in) :: level, nvar, naux
  real(kind=8), intent(out) :: dtlevnew
  logical, intent(in) :: vtime
  real(c_double) :: dtn, cfl_dev, tlev
  integer :: j, mptr
  integer :: clock_start, clock_finish, clock_rate
  real(kind=8) :: cpu_start, cpu_finish
  call system_clock(clock_start, clock_rate)
  call cpu_time(cpu_start)
  call amrgpu_sync_hierarchy(nvar, naux)
  if (numgrids(level) .eq. 0) then
     dtlevnew = rinfinity
     cfl_level = 0.d0
     return
  end if
  tlev = rnode(timemult, listOfGrids(listStart(level)))
  call amrgpu_check(amrgpu_set_time(level, tlev, possk(level)), 'set_time')
  ! advanc.f:249-253: the device samples the gauges of this level's grids
  ! before their step (amrgpu_advance_level); their records join the
  ! gauges' buffers below
  call amrgpu_install_gauges()
  call amrgpu_gauge_sources()
  call amrgpu_check(amrgpu_advance_level(level, dtn, cfl_dev), 'advanc')
  call amrgpu_collect_gauges(nvar)
  call amrgpu_pull_level(level, level .lt. mxnest)
  ! this level's registers (qad, fluxad) and the next finer level's
  ! (saveqc, fluxsv)
  call amrgpu_pull_registers(level)
  if (level .lt. mxnest) then
     if (numgrids(level + 1) .gt. 0) call amrgpu_pull_registers(level + 1)
  end if
  dtlevnew = dtn
  cfl_level = cfl_dev                              ! advanc.f:99,138
  cflmax = dmax1(cflmax, cfl_level)
  do j = 1, numgrids(level)
     mptr = listOfGrids(listStart(level) + j - 1)
     rnode(timemult, mptr) = rnode(timemult, mptr) + possk(level)   ! advanc.f:289
     rvol = rvol + (node(ndihi,mptr) - node(ndilo,mptr) + 1) * &    ! advanc.f:222-225
          (node(ndjhi,mptr) - node(ndjlo,mptr) + 1)
     rvoll(level) = rvoll(level) + (node(ndihi,mptr) - node(ndilo,mptr) + 1) * &
          (node(ndjhi,mptr) - node(ndjlo,mptr) + 1)
  end do
  call system_clock(clock_finish, clock_rate)
  call cpu_time(cpu_finish)
  tvoll(level) = tvoll(level) + clock_finish - clock_start         ! advanc.f:131-135
  tvollCPU(level) = tvollCPU(level) + cpu_finish - cpu_start
  timeStepgrid = timeStepgrid + clock_finish - clock_start
  timeStepgridCPU = timeStepgridCPU + cpu_finish - cpu_start
end subroutine advanc

! update.f:16 -- same name and arguments (tick.f:295, setgrd.f:127).
subroutine update(level, nvar, naux)
  use amr_module, only: mxnest
  use amrgpu_iface
  implicit none
  integer, intent(in) :: level, nvar, naux
  call amrgpu_sync_hierarchy(nvar, naux)
  call amrgpu_check(amrgpu_update_level(level), 'update')
  call amrgpu_pull_level(level, .false.)
end subroutine update

