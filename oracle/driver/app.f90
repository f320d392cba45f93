! app.f90 -- the application side of a run of the reference's own driver
! (amr2, tick, regrid, ...; TEST INFRASTRUCTURE ONLY: built into oracle/_ref).
!
! An AMRClaw application supplies setprob, qinit and the Riemann solvers
! rpn2/rpt2 (src/2d/Makefile.amr_2d: they are not in the library).  Here:
!   setprob  reads setprob.data (written by amrclaw_amd.rundata): the initial
!            state, the solver and its /cparam/ constants, and the
!            precision a device run uses;
!   qinit    the initial states of the run configurations;
!   rpn2/rpt2 forward to the oracle's solvers (oracle/orc_riemann.c), whose
!            constants setprob installs;
!   amrgpu_app_physics hands the same solver and constants to the library's
!            Fortran shim (amrclaw_amd/fortran/amrgpu_shim.f90).
module drv_app
  implicit none
  integer :: problem = 1, solver = 1, efix = 1, precision = 0
  real(kind=8) :: ubar = 0.5d0, vbar = 1.d0, rho = 1.d0, bulk = 4.d0, gamma = 1.4d0
end module drv_app

subroutine setprob
  use drv_app
  implicit none
  interface
     subroutine orc_driver_physics(solver, ubar, vbar, rho, bulk, gamma, efix) &
          bind(C, name="orc_driver_physics")
       use iso_c_binding
       integer(c_int) :: solver, efix
       real(c_double) :: ubar, vbar, rho, bulk, gamma
     end subroutine orc_driver_physics
  end interface
  call opendatafile(7, 'setprob.data')
  read(7, *) problem
  read(7, *) solver
  read(7, *) ubar
  read(7, *) vbar
  read(7, *) rho
  read(7, *) bulk
  read(7, *) gamma
  read(7, *) efix
  read(7, *) precision
  close(7)
  call orc_driver_physics(solver, ubar, vbar, rho, bulk, gamma, efix)
end subroutine setprob

! Initial states at the cell centres:
!   1  the square of tests/advection_2d_square/qinit.f:20-26 (q = 1 for
!      0.1 < x, y < 0.6, else 0.1);
!   2  advection of a disk: q = 1 for r < 0.2 about (0.5, 0.5), else 0;
!   3  acoustics pulse: p = 1 + cos(pi r / 0.35) for r < 0.35, else 1,
!      u = v = 0;
!   4  Euler four-quadrant problem split at (0.8, 0.8): (rho, u, v, p) =
!      NE (1.5, 0, 0, 1.5), NW (0.5323, 1.206, 0, 0.3),
!      SW (0.138, 1.206, 1.206, 0.029), SE (0.5323, 0, 1.206, 0.3).
subroutine qinit(meqn, mbc, mx, my, xlower, ylower, dx, dy, q, maux, aux)
  use drv_app
  implicit none
  integer, intent(in) :: meqn, mbc, mx, my, maux
  real(kind=8), intent(in) :: xlower, ylower, dx, dy
  real(kind=8), intent(inout) :: q(meqn, 1-mbc:mx+mbc, 1-mbc:my+mbc)
  real(kind=8), intent(inout) :: aux(*)
  real(kind=8), parameter :: pi = 3.14159265358979323846d0
  real(kind=8) :: x, y, r, rh, u, v, p
  integer :: i, j
  do j = 1, my
     y = ylower + (j - 0.5d0) * dy
     do i = 1, mx
        x = xlower + (i - 0.5d0) * dx
        select case (problem)
        case (1)
           if (x > 0.1d0 .and. x < 0.6d0 .and. y > 0.1d0 .and. y < 0.6d0) then
              q(1, i, j) = 1.d0
           else
              q(1, i, j) = 0.1d0
           end if
        case (2)
           r = sqrt((x - 0.5d0)**2 + (y - 0.5d0)**2)
           q(1, i, j) = merge(1.d0, 0.d0, r < 0.2d0)
        case (3)
           r = sqrt(x * x + y * y)
           q(1, i, j) = merge(1.d0 + cos(pi * r / 0.35d0), 1.d0, r < 0.35d0)
           q(2, i, j) = 0.d0
           q(3, i, j) = 0.d0
        case (4)
           if (x >= 0.8d0 .and. y >= 0.8d0) then
              rh = 1.5d0; u = 0.d0; v = 0.d0; p = 1.5d0
           else if (y >= 0.8d0) then
              rh = 0.5323d0; u = 1.206d0; v = 0.d0; p = 0.3d0
           else if (x < 0.8d0) then
              rh = 0.138d0; u = 1.206d0; v = 1.206d0; p = 0.029d0
           else
              rh = 0.5323d0; u = 0.d0; v = 1.206d0; p = 0.3d0
           end if
           q(1, i, j) = rh
           q(2, i, j) = rh * u
           q(3, i, j) = rh * v
           q(4, i, j) = p / (gamma - 1.d0) + 0.5d0 * rh * (u * u + v * v)
        end select
     end do
  end do
end subroutine qinit

subroutine rpn2(ixy, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, auxl, auxr, wave, s, amdq, apdq)
  use iso_c_binding
  implicit none
  integer :: ixy, maxm, meqn, mwaves, maux, mbc, mx
  real(kind=8) :: ql(*), qr(*), auxl(*), auxr(*), wave(*), s(*), amdq(*), apdq(*)
  interface
     subroutine orc_rpn2_byref(ixy, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, auxl, auxr, &
          wave, s, amdq, apdq) bind(C, name="orc_rpn2_byref")
       import :: c_int, c_double
       integer(c_int) :: ixy, maxm, meqn, mwaves, maux, mbc, mx
       real(c_double) :: ql(*), qr(*), auxl(*), auxr(*), wave(*), s(*), amdq(*), apdq(*)
     end subroutine orc_rpn2_byref
  end interface
  call orc_rpn2_byref(ixy, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, auxl, auxr, wave, s, &
       amdq, apdq)
end subroutine rpn2

subroutine rpt2(ixy, imp, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, aux1, aux2, aux3, asdq, &
     bmasdq, bpasdq)
  use iso_c_binding
  implicit none
  integer :: ixy, imp, maxm, meqn, mwaves, maux, mbc, mx
  real(kind=8) :: ql(*), qr(*), aux1(*), aux2(*), aux3(*), asdq(*), bmasdq(*), bpasdq(*)
  interface
     subroutine orc_rpt2_byref(ixy, imp, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, aux1, aux2, &
          aux3, asdq, bmasdq, bpasdq) bind(C, name="orc_rpt2_byref")
       import :: c_int, c_double
       integer(c_int) :: ixy, imp, maxm, meqn, mwaves, maux, mbc, mx
       real(c_double) :: ql(*), qr(*), aux1(*), aux2(*), aux3(*), asdq(*), bmasdq(*), bpasdq(*)
     end subroutine orc_rpt2_byref
  end interface
  call orc_rpt2_byref(ixy, imp, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, aux1, aux2, aux3, &
       asdq, bmasdq, bpasdq)
end subroutine rpt2

! The solver and constants the device runs with (the shim calls this once).
subroutine amrgpu_app_physics(solver_, ubar_, vbar_, rho_, bulk_, gamma_, efix_, precision_)
  use iso_c_binding
  use drv_app
  implicit none
  integer(c_int), intent(out) :: solver_, efix_, precision_
  real(c_double), intent(out) :: ubar_, vbar_, rho_, bulk_, gamma_
  solver_ = solver
  ubar_ = ubar
  vbar_ = vbar
  rho_ = rho
  bulk_ = bulk
  gamma_ = gamma
  efix_ = efix
  precision_ = precision
end subroutine amrgpu_app_physics
