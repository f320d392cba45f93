! annulus_app.f90 -- the Riemann solvers and device physics of the
! reference's tests/advection_2d_annulus (TEST INFRASTRUCTURE ONLY: built
! into oracle/_ref by `make -C oracle annulus`).
!
! The application's setprob, qinit, setaux, mapc2p and stream are the
! reference's own files, compiled where they lie.  Its rpn2 / rpt2 are
! clawpack/riemann's rpn2_vc_advection.f / rpt2_vc_advection.f, which the
! reference does not ship: here they forward to the oracle's
! variable-coefficient advection solver (oracle/orc_riemann.c), installed
! on the first call.  amrgpu_app_physics hands the library's shim the same
! solver; the precision of a device run is AMRGPU_PRECISION (0 FP64,
! 1 FP32) from the environment, FP64 when unset.
module annulus_app
  implicit none
  logical :: installed = .false.
contains
  subroutine install()
    use iso_c_binding
    interface
       subroutine orc_driver_physics(solver, ubar, vbar, rho, bulk, gamma, efix) &
            bind(C, name="orc_driver_physics")
         import :: c_int, c_double
         integer(c_int) :: solver, efix
         real(c_double) :: ubar, vbar, rho, bulk, gamma
       end subroutine orc_driver_physics
    end interface
    integer(c_int) :: solver = 5, efix = 0       ! the oracle's code of the solver
    real(c_double) :: zero = 0.d0, one = 1.d0
    if (installed) return
    call orc_driver_physics(solver, zero, zero, one, one, one, efix)
    installed = .true.
  end subroutine install
end module annulus_app

subroutine rpn2(ixy, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, auxl, auxr, wave, s, amdq, apdq)
  use iso_c_binding
  use annulus_app
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
  call install()
  call orc_rpn2_byref(ixy, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, auxl, auxr, wave, s, &
       amdq, apdq)
end subroutine rpn2

subroutine rpt2(ixy, imp, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, aux1, aux2, aux3, asdq, &
     bmasdq, bpasdq)
  use iso_c_binding
  use annulus_app
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
  call install()
  call orc_rpt2_byref(ixy, imp, maxm, meqn, mwaves, maux, mbc, mx, ql, qr, aux1, aux2, aux3, &
       asdq, bmasdq, bpasdq)
end subroutine rpt2

! The solver the device runs (AMRGPU_VCADV) and the precision of the run.
subroutine amrgpu_app_physics(solver_, ubar_, vbar_, rho_, bulk_, gamma_, efix_, precision_)
  use iso_c_binding
  implicit none
  integer(c_int), intent(out) :: solver_, efix_, precision_
  real(c_double), intent(out) :: ubar_, vbar_, rho_, bulk_, gamma_
  character(len=8) :: env
  integer :: n, st
  solver_ = 4
  ubar_ = 0.d0
  vbar_ = 0.d0
  rho_ = 1.d0
  bulk_ = 1.d0
  gamma_ = 1.4d0
  efix_ = 0
  precision_ = 0
  call get_environment_variable('AMRGPU_PRECISION', env, n, st)
  if (st == 0 .and. n > 0) read(env(1:n), *) precision_
end subroutine amrgpu_app_physics
