! amrgpu_shim.f90 -- Fortran side of the drop-in boundary: ISO_C_BINDING
! interfaces of include/amrgpu.h and replacements for the reference's
! advanc (advanc.f:10), update (update.f:16) and conck (conck.f:6), plus the
! layout hook a driver calls after regrid and the download it calls before
! valout / check / regrid flagging.  Compiled against the reference's
! amr_module; linked with libamrgpu.so in place of advanc.f, update.f and
! conck.f (Makefile.amr_2d EXCLUDE_SOURCES).
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
     integer(c_int) function amrgpu_get_level(level, which, alloc_base, offsets) &
          bind(C, name="amrgpu_get_level")
       import :: c_int, c_double, c_int64_t
       integer(c_int), value :: level, which
       real(c_double) :: alloc_base(*)
       integer(c_int64_t), intent(in) :: offsets(*)
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

! After regrid (regrid.f:72-118) and at start-up (setgrd): hand every level's
! grids and both time levels to the device.
subroutine amrgpu_upload_hierarchy(nvar)
  use amr_module
  use amrgpu_iface
  implicit none
  integer, intent(in) :: nvar
  integer :: level, j, mptr, n
  integer(c_int), allocatable :: ilo(:), ihi(:), jlo(:), jhi(:)
  integer(c_int64_t), allocatable :: off1(:), off2(:)
  real(c_double), allocatable :: cx(:), cy(:)
  integer(c_int) :: isz(maxlv), jsz(maxlv), rx(maxlv), ry(maxlv)
  do level = 1, mxnest
     isz(level) = iregsz(level)
     jsz(level) = jregsz(level)
     rx(level) = intratx(level)
     ry(level) = intraty(level)
  end do
  call amrgpu_check(amrgpu_set_geometry(mxnest, isz, jsz, rx, ry, hxposs, hyposs), 'set_geometry')
  do level = 1, mxnest
     n = numgrids(level)
     allocate(ilo(max(n,1)), ihi(max(n,1)), jlo(max(n,1)), jhi(max(n,1)))
     allocate(off1(max(n,1)), off2(max(n,1)), cx(max(n,1)), cy(max(n,1)))
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
     end do
     call amrgpu_check(amrgpu_set_level(level, n, ilo, ihi, jlo, jhi, hxposs(level), &
          hyposs(level)), 'set_level')
     if (n .gt. 0) then
        call amrgpu_check(amrgpu_set_corners(level, cx, cy), 'set_corners')
        call amrgpu_check(amrgpu_put_level(level, 0, alloc, off1), 'put_level new')
        if (level .lt. mxnest) call amrgpu_check(amrgpu_put_level(level, 1, alloc, off2), &
             'put_level old')
        call amrgpu_check(amrgpu_set_time(level, rnode(timemult, listOfGrids(listStart(level))), &
             possk(level)), 'set_time')
     end if
     deallocate(ilo, ihi, jlo, jhi, off1, off2, cx, cy)
  end do
end subroutine amrgpu_upload_hierarchy

! Before valout, check, conck on the host or regrid flagging: bring the new
! (and old) values of every level back into alloc.
subroutine amrgpu_download_hierarchy(nvar)
  use amr_module
  use amrgpu_iface
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
     call amrgpu_check(amrgpu_get_level(level, 0, alloc, off1), 'get_level new')
     if (level .lt. mxnest) call amrgpu_check(amrgpu_get_level(level, 1, alloc, off2), &
          'get_level old')
     deallocate(off1, off2)
  end do
end subroutine amrgpu_download_hierarchy

! advanc.f:10 -- same name and arguments; tick.f:240 calls it unchanged.
subroutine advanc(level, nvar, dtlevnew, vtime, naux)
  use amr_module
  use amrgpu_iface
  implicit none
  integer, intent(in) :: level, nvar, naux
  real(kind=8), intent(out) :: dtlevnew
  logical, intent(in) :: vtime
  real(c_double) :: dtn, cfl_dev
  integer :: j, mptr
  call amrgpu_check(amrgpu_advance_level(level, dtn, cfl_dev), 'advanc')
  dtlevnew = dtn
  cfl_level = cfl_dev                              ! advanc.f:138
  cflmax = dmax1(cflmax, cfl_level)
  do j = 1, numgrids(level)
     mptr = listOfGrids(listStart(level) + j - 1)
     rnode(timemult, mptr) = rnode(timemult, mptr) + possk(level)   ! advanc.f:289
     rvol = rvol + (node(ndihi,mptr) - node(ndilo,mptr) + 1) * &    ! advanc.f:222-225
          (node(ndjhi,mptr) - node(ndjlo,mptr) + 1)
     rvoll(level) = rvoll(level) + (node(ndihi,mptr) - node(ndilo,mptr) + 1) * &
          (node(ndjhi,mptr) - node(ndjlo,mptr) + 1)
  end do
end subroutine advanc

! update.f:16 -- same name and arguments (tick.f:295, setgrd.f:127).
subroutine update(level, nvar, naux)
  use amrgpu_iface
  implicit none
  integer, intent(in) :: level, nvar, naux
  call amrgpu_check(amrgpu_update_level(level), 'update')
end subroutine update
