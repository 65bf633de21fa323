! ISO_C_BINDING interface to libsmac_amd.so (include/smac_amd.h) for Fortran callers.
!
! The derived types mirror the C structs field for field (bind(C)); arrays are passed as the
! reference's own Fortran arrays (j fastest), so a caller hands over q(jmax,kmax,3),
! btc(jmax,kmax,3,3), ... unchanged.
module smac_amd
  use iso_c_binding
  implicit none
  private
  public :: smac_params, smac_bc_region, smac_iter_stats
  public :: smac_lrelax, smac_lusgs, smac_create, smac_destroy, smac_step, smac_run
  public :: smac_set_state, smac_get_state, smac_last_error_f
  public :: SMAC_OK, SMAC_FP64, SMAC_ORDER_REFERENCE

  integer(c_int), parameter :: SMAC_OK = 0, SMAC_FP64 = 0, SMAC_ORDER_REFERENCE = 0

  type, bind(C) :: smac_params
    real(c_double) :: alpha, beta, reynum, dtau, dt, underr, cdis, pin, pout
    real(c_double) :: chord, xmom, ymom
    integer(c_int) :: impsch, ivis, iturb, iflxo, njsp, nksp, nbcimp, ntjsp, ntksp
    integer(c_int) :: precision, line_order
  end type

  type, bind(C) :: smac_bc_region
    integer(c_int) :: ibcval, jbeg, jend, kbeg, kend
  end type

  type, bind(C) :: smac_iter_stats
    integer(c_int) :: nt, jres, kres, nonfinite
    real(c_double) :: resmax, resmax0, divmax, turres, rhsnorm
    real(c_double) :: cl, cd, cm, clp, cdp, fxp, fyp, fxs, fys, fmom
  end type

  interface
    integer(c_int) function smac_lrelax(p, jmax, kmax, npair, wake, q, dq, s, btc, bjm, &
                                        bjp, bkm, bkp) bind(C, name="smac_lrelax")
      import :: c_int, c_double, smac_params
      type(smac_params), intent(in) :: p
      integer(c_int), value :: jmax, kmax, npair
      integer(c_int), intent(in) :: wake(*)
      real(c_double), intent(in) :: q(*), btc(*), bjm(*), bjp(*), bkm(*), bkp(*)
      real(c_double), intent(inout) :: dq(*), s(*)
    end function

    integer(c_int) function smac_lusgs(p, jmax, kmax, neqs, npair, wake, q, dq, rhs, btc, &
                                       bjm, bjp, bkm, bkp, scr1) bind(C, name="smac_lusgs")
      import :: c_int, c_double, smac_params
      type(smac_params), intent(in) :: p
      integer(c_int), value :: jmax, kmax, neqs, npair
      integer(c_int), intent(in) :: wake(*)
      real(c_double), intent(in) :: q(*), btc(*), bjm(*), bjp(*), bkm(*), bkp(*)
      real(c_double), intent(inout) :: dq(*), rhs(*), scr1(*)
    end function

    integer(c_int) function smac_create(p, jmax, kmax, x, y, rtxy, dj, nreg, regs, h) &
        bind(C, name="smac_create")
      import :: c_int, c_double, c_ptr, smac_params, smac_bc_region
      type(smac_params), intent(in) :: p
      integer(c_int), value :: jmax, kmax, nreg
      real(c_double), intent(in) :: x(*), y(*), rtxy(*), dj(*)
      type(smac_bc_region), intent(in) :: regs(*)
      type(c_ptr), intent(out) :: h
    end function

    integer(c_int) function smac_destroy(h) bind(C, name="smac_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function

    integer(c_int) function smac_step(h, st) bind(C, name="smac_step")
      import :: c_int, c_ptr, smac_iter_stats
      type(c_ptr), value :: h
      type(smac_iter_stats), intent(out) :: st
    end function

    integer(c_int) function smac_run(h, niter, drop, st, done) bind(C, name="smac_run")
      import :: c_int, c_double, c_ptr, smac_iter_stats
      type(c_ptr), value :: h
      integer(c_int), value :: niter
      real(c_double), value :: drop
      type(smac_iter_stats), intent(out) :: st
      integer(c_int), intent(out) :: done
    end function

    integer(c_int) function smac_set_state(h, q, anut, nt, resmax0, clift) &
        bind(C, name="smac_set_state")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(in) :: q(*), anut(*)
      integer(c_int), value :: nt
      real(c_double), value :: resmax0, clift
    end function

    integer(c_int) function smac_get_state(h, q, anut, vnut) bind(C, name="smac_get_state")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      real(c_double), intent(out) :: q(*), anut(*), vnut(*)
    end function

    integer(c_int) function smac_get_stats(h, st) bind(C, name="smac_get_stats")
      import :: c_int, c_ptr, smac_iter_stats
      type(c_ptr), value :: h
      type(smac_iter_stats), intent(out) :: st
    end function

    integer(c_int) function smac_set_option(h, name, value) bind(C, name="smac_set_option")
      import :: c_int, c_char, c_ptr
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: name(*)   ! NUL-terminated
      integer(c_int), value :: value
    end function

    integer(c_int) function smac_batch_create(cases, n, b) bind(C, name="smac_batch_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: cases(*)
      integer(c_int), value :: n
      type(c_ptr), intent(out) :: b
    end function

    integer(c_int) function smac_batch_run(b, npass) bind(C, name="smac_batch_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: npass
    end function

    integer(c_int) function smac_batch_run_until(b, max_pass, drop, passes) &
        bind(C, name="smac_batch_run_until")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: max_pass
      real(c_double), value :: drop
      integer(c_int) :: passes(*)
    end function

    integer(c_int) function smac_batch_timings(b, out) bind(C, name="smac_batch_timings")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double), intent(out) :: out(3)
    end function

    integer(c_int) function smac_batch_destroy(b) bind(C, name="smac_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: b
    end function

    type(c_ptr) function smac_last_error_c() bind(C, name="smac_last_error")
      import :: c_ptr
    end function
  end interface

contains

  ! smac_last_error as a Fortran string (for stop/print at the call site)
  function smac_last_error_f() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: chars(:)
    integer :: i
    msg = ' '
    call c_f_pointer(smac_last_error_c(), chars, [256])
    do i = 1, 256
      if (chars(i) == c_null_char) exit
      msg(i:i) = chars(i)
    end do
  end function
end module smac_amd
