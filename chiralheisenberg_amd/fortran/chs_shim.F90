! chs_shim.F90 -- drop-in replacements for the fermion routines of
! rhmc/cheisdwf.f, with the reference's own subroutine names, argument lists
! and COMMON blocks, forwarding to the C-ABI of include/chs.h (libchs, HIP on
! the MI355X) through iso_c_binding.
!
! A program built from rhmc/cheisdwf.f links this file instead of its own
! dslash, dslashd, qmrherm, congrad, derivs, force and hamilton
! (rhmc/cheisdwf.f:2634, 3071, 599, 1064, 747, 470, 522).  Compile it with the
! same default-kind flags as the caller (e.g. -fdefault-real-8) and the
! caller's lattice sizes:
!     flang -cpp -DKSIZE=12 -DKSIZET=12 -DKTHIRD=8 -c chs_shim.F90
! The reference's sizes are compile-time parameters (rhmc/cheisdwf.f:31); the
! context is created on the first call with those sizes and /para/ am3, ibound.
! Every call checks the status the C-ABI returns and stops with the library's
! message on failure (the reference has no status to return).

#ifndef KSIZE
#define KSIZE 12
#endif
#ifndef KSIZET
#define KSIZET 12
#endif
#ifndef KTHIRD
#define KTHIRD 8
#endif

module chs_binding
  use iso_c_binding
  implicit none
  integer, parameter :: ksize = KSIZE, ksizet = KSIZET, kthird = KTHIRD
  integer, parameter :: kvol = ksize*ksize*ksizet
  integer, parameter :: kferm = 4*2*kthird*kvol
  type(c_ptr), save :: ctx = c_null_ptr

  interface
    function chs_ctx_create(L, Lt, Ls, ibound, am3, device) bind(C, name="chs_ctx_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: L, Lt, Ls, ibound, device
      real(c_double), value :: am3
      type(c_ptr) :: chs_ctx_create
    end function
    function chs_create_error() bind(C, name="chs_create_error")
      import :: c_ptr
      type(c_ptr) :: chs_create_error
    end function
    function chs_last_error(c) bind(C, name="chs_last_error")
      import :: c_ptr
      type(c_ptr), value :: c
      type(c_ptr) :: chs_last_error
    end function
    integer(c_int) function c_dslash(c, phi, r, chi, am, hpin, imass) bind(C, name="chs_dslash")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: c
      complex(c_double_complex) :: phi(*), r(*)
      real(c_double) :: chi(*)
      real(c_double), value :: am, hpin
      integer(c_int), value :: imass
    end function
    integer(c_int) function c_dslashd(c, phi, r, chi, am, hpin, imass) bind(C, name="chs_dslashd")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: c
      complex(c_double_complex) :: phi(*), r(*)
      real(c_double) :: chi(*)
      real(c_double), value :: am, hpin
      integer(c_int), value :: imass
    end function
    integer(c_int) function c_qmrherm(c, phi, res, itercg, am, hpin, imass, anum, aden, ndiag, iflag, chi, x, &
                                      dsdchi) bind(C, name="chs_qmrherm")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: c
      complex(c_double_complex) :: phi(*), x(*)
      real(c_double), value :: res, am, hpin
      integer(c_int) :: itercg
      integer(c_int), value :: imass, ndiag, iflag
      real(c_double) :: anum(*), aden(*), chi(*), dsdchi(*)
    end function
    integer(c_int) function c_congrad(c, phi, res, itercg, am, hpin, imass, chi, x) bind(C, name="chs_congrad")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: c
      complex(c_double_complex) :: phi(*), x(*)
      real(c_double), value :: res, am, hpin
      integer(c_int) :: itercg
      integer(c_int), value :: imass
      real(c_double) :: chi(*)
    end function
    integer(c_int) function c_force(c, phi, res1, am, hpin, imass, chi, beta, anum, aden, bnum, bden, ndiag, &
                                    dsdchi, x, its) bind(C, name="chs_force")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: c
      complex(c_double_complex) :: phi(*), x(*)
      real(c_double), value :: res1, am, hpin, beta
      integer(c_int), value :: imass, ndiag
      real(c_double) :: chi(*), anum(*), aden(*), bnum(*), bden(*), dsdchi(*)
      integer(c_int) :: its(2)
    end function
    integer(c_int) function c_hamilton(c, phi, h, hg, hp, s, res2, isweep, am, hpin, imass, chi, pp, beta, &
                                       bnum, bden, ndiag, x, itercg) bind(C, name="chs_hamilton")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: c
      complex(c_double_complex) :: phi(*), x(*)
      real(c_double) :: h, hg, hp, s
      real(c_double), value :: res2, am, hpin, beta
      integer(c_int), value :: isweep, imass, ndiag
      real(c_double) :: chi(*), pp(*), bnum(*), bden(*)
      integer(c_int) :: itercg
    end function
    integer(c_int) function c_derivs(c, r, x2, anum, imass, dsdchi) bind(C, name="chs_derivs")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: c
      complex(c_double_complex) :: r(*), x2(*)
      real(c_double), value :: anum
      integer(c_int), value :: imass
      real(c_double) :: dsdchi(*)
    end function
    integer(c_int) function c_mirror_register(c, a) bind(C, name="chs_mirror_register")
      import :: c_int, c_double_complex, c_ptr
      type(c_ptr), value :: c
      complex(c_double_complex) :: a(*)
    end function
    integer(c_int) function c_mirror_touch(c, a) bind(C, name="chs_mirror_touch")
      import :: c_int, c_double_complex, c_ptr
      type(c_ptr), value :: c
      complex(c_double_complex) :: a(*)
    end function
    integer(c_int) function c_set_force_precision(c, prec) bind(C, name="chs_set_force_precision")
      import :: c_int, c_ptr
      type(c_ptr), value :: c
      integer(c_int), value :: prec
    end function
    integer(c_int) function c_set_shift_freeze(c, f) bind(C, name="chs_set_shift_freeze")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: c
      real(c_double), value :: f
    end function
  end interface

contains

  subroutine chs_stop(msg_ptr, where)
    type(c_ptr), intent(in) :: msg_ptr
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: cs(:)
    character(len=512) :: msg
    integer :: i
    msg = ' '
    if (c_associated(msg_ptr)) then
      call c_f_pointer(msg_ptr, cs, [512])
      do i = 1, 512
        if (cs(i) == c_null_char) exit
        msg(i:i) = cs(i)
      end do
    end if
    write(6, '(a, ": ", a)') where, trim(msg)
    stop 1
  end subroutine

  subroutine chs_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) call chs_stop(chs_last_error(ctx), where)
  end subroutine

  ! a multi-shift solve that stopped at its iteration cap: the reference
  ! carries on silently (rhmc/cheisdwf.f:639); here it carries on with a
  ! message, and stops on any other error
  subroutine chs_nonconv(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc == 3) then
      write(6, '(a, ": iteration cap reached, result kept")') where
    else
      call chs_check(rc, where)
    end if
  end subroutine

  ! create the context on first use, with /para/ am3, ibound of the caller
  subroutine chs_ensure(am3, ibound)
    real(c_double), intent(in) :: am3
    integer, intent(in) :: ibound
    if (.not. c_associated(ctx)) then
      ctx = chs_ctx_create(int(ksize, c_int), int(ksizet, c_int), int(kthird, c_int), int(ibound, c_int), am3, 0_c_int)
      if (.not. c_associated(ctx)) call chs_stop(chs_create_error(), 'chs_ctx_create')
    end if
  end subroutine

end module chs_binding

! ---------------------------------------------------------------------------
! The reference's routines.  Implicit typing and COMMON blocks as in
! rhmc/cheisdwf.f; reals are converted to C doubles at the boundary so the
! shim works whatever the default real kind of the caller.

subroutine dslash(Phi, R, chi, am, hpin, imass)
  use chs_binding
  common /para/ beta, am3, ibound
  complex*16 Phi(kferm), R(kferm)
  real*8 chi(kvol, 3)
  call chs_ensure(real(am3, c_double), ibound)
  call chs_check(c_dslash(ctx, Phi, R, chi, real(am, c_double), real(hpin, c_double), int(imass, c_int)), 'dslash')
end subroutine

subroutine dslashd(Phi, R, chi, am, hpin, imass)
  use chs_binding
  common /para/ beta, am3, ibound
  complex*16 Phi(kferm), R(kferm)
  real*8 chi(kvol, 3)
  call chs_ensure(real(am3, c_double), ibound)
  call chs_check(c_dslashd(ctx, Phi, R, chi, real(am, c_double), real(hpin, c_double), int(imass, c_int)), 'dslashd')
end subroutine

subroutine qmrherm(Phi, res, itercg, am, hpin, imass, anum, aden, ndiag, iflag)
  use chs_binding
  common /trial/ chi(kvol, 3), pp(kvol, 3)
  common /sforce/ dSdchi(kvol, 3)
  common /para/ bbb, am3, ibound
  common /vector/ x(kferm)
  complex*16 Phi(kferm), x
  real*8 chi, anum(0:ndiag), aden(ndiag)
  real(c_double) :: f(kvol, 3)
  integer(c_int) :: it, rc
  call chs_ensure(real(am3, c_double), ibound)
  f = real(dSdchi, c_double)
  rc = c_qmrherm(ctx, Phi, real(res, c_double), it, real(am, c_double), real(hpin, c_double), &
                 int(imass, c_int), anum, aden, int(ndiag, c_int), int(iflag, c_int), chi, x, f)
  call chs_nonconv(rc, 'qmrherm')
  if (iflag == 1) dSdchi = f
  itercg = it
end subroutine

subroutine congrad(Phi, res, itercg, am, hpin, imass)
  use chs_binding
  common /trial/ chi(kvol, 3), pp(kvol, 3)
  common /para/ bbb, am3, ibound
  common /vector/ x(kferm)
  complex*16 Phi(kferm), x
  real*8 chi
  integer(c_int) :: it, rc
  call chs_ensure(real(am3, c_double), ibound)
  rc = c_congrad(ctx, Phi, real(res, c_double), it, real(am, c_double), real(hpin, c_double), int(imass, c_int), chi, x)
  ! like the reference, running into the iteration cap only writes its
  ! message, to the terminal and to the program's output file (unit 7,
  ! rhmc/cheisdwf.f:1151-1153)
  if (rc == 3) then
    write(6, 1000)
    write(7, 1000)
1000 format(' # iterations of congrad exceeds niterc')
  else
    call chs_check(rc, 'congrad')
  end if
  itercg = it
end subroutine

subroutine derivs(R, X2, anum, imass)
  use chs_binding
  common /sforce/ dSdchi(kvol, 3)
  complex*16 R(kferm), X2(kferm)
  real*8 anum
  real(c_double) :: f(kvol, 3)
  if (.not. c_associated(ctx)) call chs_stop(c_null_ptr, 'derivs called before any other routine')
  f = real(dSdchi, c_double)
  call chs_check(c_derivs(ctx, R, X2, anum, int(imass, c_int), f), 'derivs')
  dSdchi = f
end subroutine

subroutine force(Phi, res1, am, hpin, imass)
  use chs_binding
  integer, parameter :: ndiagg = 12
  common /trial/ chi(kvol, 3), pp(kvol, 3)
  common /sforce/ dSdchi(kvol, 3)
  common /para/ beta, am3, ibound
  common /param/ ancg, ancgh, ancgf, ancgpf
  common /vector/ X1(kferm)
  common /remez2g/ anum(0:ndiagg), aden(ndiagg), bnum(0:ndiagg), bden(ndiagg)
  complex*16 Phi(kferm), X1
  real*8 chi, anum, aden, bnum, bden
  real(c_double) :: f(kvol, 3)
  integer(c_int) :: its(2)
  call chs_ensure(real(am3, c_double), ibound)
  call chs_nonconv(c_force(ctx, Phi, real(res1, c_double), real(am, c_double), real(hpin, c_double), &
                           int(imass, c_int), chi, real(beta, c_double), anum, aden, bnum, bden, int(ndiagg, c_int), &
                           f, X1, its), 'force')
  dSdchi = f
  ancg = ancg + float(its(1))
  ancgf = ancgf + float(its(2))
end subroutine

subroutine hamilton(Phi, h, hg, hp, s, res2, isweep, am, hpin, imass)
  use chs_binding
  integer, parameter :: ndiag = 25
  common /trial/ chi(kvol, 3), pp(kvol, 3)
  common /remez2/ anum(0:ndiag), aden(ndiag), bnum(0:ndiag), bden(ndiag)
  common /para/ beta, am3, ibound
  common /param/ ancg, ancgh, ancgf, ancgpf
  common /vector/ X1(kferm)
  complex*16 Phi(kferm), X1
  real*8 chi, h, hg, hp, s
  real*8 anum, aden, bnum, bden
  real(c_double) :: p(kvol, 3), hh, hhg, hhp, ss
  integer(c_int) :: it
  call chs_ensure(real(am3, c_double), ibound)
  p = real(pp, c_double)
  call chs_nonconv(c_hamilton(ctx, Phi, hh, hhg, hhp, ss, real(res2, c_double), int(isweep, c_int), &
                              real(am, c_double), real(hpin, c_double), int(imass, c_int), chi, p, &
                              real(beta, c_double), bnum, bden, int(ndiag, c_int), X1, it), 'hamilton')
  h = hh
  hg = hhg
  hp = hhp
  s = ss
  ancgh = ancgh + float(it)
end subroutine

! ---------------------------------------------------------------------------
! Device mirror of a read-only spinor the caller passes to many calls (Phi of
! the trajectory): after `call chs_mirror(Phi)` once, the library keeps Phi on
! the device and uploads it again only after `call chs_touch(Phi)`, which the
! caller adds after each write to Phi (the pseudofermion heatbath,
! rhmc/cheisdwf.f:238-245).  No counterpart in the reference.

subroutine chs_mirror(Phi)
  use chs_binding
  common /para/ bbb, am3, ibound
  complex*16 Phi(kferm)
  call chs_ensure(real(am3, c_double), ibound)
  call chs_check(c_mirror_register(ctx, Phi), 'chs_mirror')
end subroutine

subroutine chs_touch(Phi)
  use chs_binding
  complex*16 Phi(kferm)
  if (.not. c_associated(ctx)) call chs_stop(c_null_ptr, 'chs_touch called before chs_mirror')
  call chs_check(c_mirror_touch(ctx, Phi), 'chs_touch')
end subroutine

! ---------------------------------------------------------------------------
! Solver options of the shim's context (no counterpart in the reference):
! iprec = 0 keeps the force's multi-shift solves in FP64 (the default), 1
! stores their spinors in FP32 (chs_set_force_precision); freeze is the
! shift-freezing factor f of every multi-shift solve, 0 (the default)
! updating every shift as the reference does (chs_set_shift_freeze).
! Call it after /para/ is set and before or between the other routines.

subroutine chs_options(iprec, freeze)
  use chs_binding
  common /para/ bbb, am3, ibound
  integer iprec
  double precision freeze
  call chs_ensure(real(am3, c_double), ibound)
  if (iprec /= 0 .and. iprec /= 1) call chs_stop(c_null_ptr, 'chs_options: iprec must be 0 or 1')
  call chs_check(c_set_force_precision(ctx, int(iprec, c_int)), 'chs_options')
  call chs_check(c_set_shift_freeze(ctx, real(freeze, c_double)), 'chs_options')
end subroutine
