! ref3_harness.f90 -- C-callable harness around the reference's own 3-D
! Fortran (TEST INFRASTRUCTURE ONLY; built into oracle/_ref, never shipped).
!
! stepgrid, step3, flux3 and inlinelimiter are compiled unmodified from the
! reference's src/3d by oracle/Makefile (target ref3).  rpn3/rpt3/rptt3 are
! the physics plugin the application links (flux3.f); here they forward to
! the oracle's solvers.

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

end module ref3_harness
