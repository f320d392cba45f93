! amrgpu3_shim.f90 -- ISO_C_BINDING boundary of the 3-D patch step.
!
! Drop-in for src/3d/stepgrid.f:4 (same name and arguments; par_advanc,
! src/3d/advanc.f:263, calls it unchanged): the step of one grid runs on the
! GPU through amrgpu3_stepgrid (include/amrgpu3.h), which returns the new
! solution, the six flux arrays fluxsv/fluxad read (advanc.f:284,294) and
! dtnew.  amrgpu3_init is called once after the run parameters are read
! (amr3.f90), with the configuration filled from method, mthlim, cfl and the
! solver's constants.
module amrgpu3_iface
  use iso_c_binding
  implicit none

  type, bind(C) :: amrgpu3_config
     integer(c_int) :: solver, nvar, mwaves, naux, nghost
     integer(c_int) :: method(7), mthlim(8), mcapa, use_fwaves
     real(c_double) :: cfl, cflv1
     real(c_double) :: ubar, vbar, wbar, gamma
     integer(c_int) :: efix, precision
  end type amrgpu3_config

  interface
     integer(c_int) function amrgpu3_init(cfg) bind(C, name="amrgpu3_init")
       import :: c_int, amrgpu3_config
       type(amrgpu3_config), intent(in) :: cfg
     end function
     integer(c_int) function amrgpu3_last_error(buf, len) bind(C, name="amrgpu3_last_error")
       import :: c_int, c_char
       character(kind=c_char) :: buf(*)
       integer(c_int), value :: len
     end function
     integer(c_int) function amrgpu3_stepgrid(mx, my, mz, q_in, q_out, dx, dy, dz, dt, &
          fm, fp, gm, gp, hm, hp, cflgrid, dtnew) bind(C, name="amrgpu3_stepgrid")
       import :: c_int, c_double
       integer(c_int), value :: mx, my, mz
       real(c_double), intent(in) :: q_in(*)
       real(c_double) :: q_out(*)
       real(c_double), value :: dx, dy, dz, dt
       real(c_double) :: fm(*), fp(*), gm(*), gp(*), hm(*), hp(*)
       real(c_double) :: cflgrid, dtnew
     end function
  end interface

contains

  subroutine amrgpu3_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    integer :: i, k
    character(len=512) :: msg
    if (rc .eq. 0) return
    k = amrgpu3_last_error(buf, 512)
    msg = ' '
    do i = 1, 512
       if (buf(i) .eq. c_null_char) exit
       msg(i:i) = buf(i)
    end do
    write(*, '(a, a, a)') trim(where), ': ', trim(msg)
    stop
  end subroutine amrgpu3_check

end module amrgpu3_iface

! src/3d/stepgrid.f:4 -- same name and arguments.  q is overwritten with the
! new solution (the ghost cells keep their values), fm..hp receive the
! fluxes of the faces of interior cells, dtnew = dt cfl / cflgrid.
subroutine stepgrid(q, fm, fp, gm, gp, hm, hp, mitot, mjtot, mktot, mbc, dt, dtnew, dx, dy, &
                    dz, nvar, xlow, ylow, zlow, time, mptr, maux, aux)
  use amr_module
  use amrgpu3_iface
  implicit none
  integer, intent(in) :: mitot, mjtot, mktot, mbc, nvar, mptr, maux
  double precision, intent(inout) :: q(nvar, mitot, mjtot, mktot)
  double precision, intent(inout) :: fm(nvar, mitot, mjtot, mktot), fp(nvar, mitot, mjtot, mktot)
  double precision, intent(inout) :: gm(nvar, mitot, mjtot, mktot), gp(nvar, mitot, mjtot, mktot)
  double precision, intent(inout) :: hm(nvar, mitot, mjtot, mktot), hp(nvar, mitot, mjtot, mktot)
  double precision, intent(in) :: aux(maux, mitot, mjtot, mktot)
  double precision, intent(in) :: dt, dx, dy, dz, xlow, ylow, zlow, time
  double precision, intent(out) :: dtnew
  real(c_double) :: cflgrid, dtn
  call amrgpu3_check(amrgpu3_stepgrid(mitot - 2 * mbc, mjtot - 2 * mbc, mktot - 2 * mbc, &
       q, q, dx, dy, dz, dt, fm, fp, gm, gp, hm, hp, cflgrid, dtn), 'stepgrid')
  dtnew = dtn
  ! stepgrid.f:215-220: report a Courant number above cflv1
  if (cflgrid .gt. cflv1) then
     write(*, 810) cflgrid
     write(outunit, 810) cflgrid
  endif
810 format('*** WARNING *** Courant number  =', d12.4, '  is larger than cflv(1) ')
end subroutine stepgrid
