! Fortran drop-in for the reference box FGT entry point.
!
! `subroutine boxfgt(...)` keeps the reference's 28-argument Fortran-77
! interface (bfgt/boxfgt.f:5-8, declarations :91-109), so the heat,
! unsteady-Stokes and Navier-Stokes drivers link against this object instead
! of bfgt/boxfgt.o with no source change.  It forwards to the HIP library
! through the C ABI `bfgt_hip_boxfgt` (include/boxfgt_hip.h) via
! iso_c_binding -- every argument by reference as in the reference, so that
! ifnewtree=1 can rewrite itree, centers and nlevels -- and stops with a
! message where the reference would silently misbehave (a tree outgrowing
! nboxes, more than 200 plane waves, ...).
!
! `boxfgt_set_precision(iprec)` selects the arithmetic of later calls:
! 0 = all-FP64 (default), 1 = split-FP32 with FP64 weights and accumulation,
! 2 = three-slice bf16 leaf transforms (FP32 accumulation), near field split-FP32.
! The library keeps the plan of a call for later calls on the same tree with
! the same parameters; `boxfgt_release_plans()` frees the kept plans.
module boxfgt_hip_mod
  use iso_c_binding
  implicit none
  private
  public :: bfgt_hip_boxfgt_c, bfgt_last_error_c, bfgt_hip_release_plans_c, boxfgt_precision, c_string

  integer, save :: boxfgt_precision = 0

  interface
    integer(c_int) function bfgt_hip_boxfgt_c(nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, &
        nboxes, nlevels, ltree, itree, iptr, centers, boxsize, fvals, ifpgh, pot, grad, hess, &
        ifnewtree, ntarg, targs, ifpghtarg, pote, grade, hesse, timeinfo, precision) &
        bind(C, name="bfgt_hip_boxfgt")
      import :: c_int, c_double
      integer(c_int), value :: nd, ndim, ipoly, iperiod, norder, npbox, nboxes, ltree
      integer(c_int) :: nlevels
      real(c_double), value :: delta, eps
      integer(c_int) :: itree(*), iptr(*)
      real(c_double) :: centers(*), boxsize(*), fvals(*)
      integer(c_int), value :: ifpgh
      real(c_double) :: pot(*), grad(*), hess(*)
      integer(c_int), value :: ifnewtree, ntarg
      real(c_double) :: targs(*)
      integer(c_int), value :: ifpghtarg
      real(c_double) :: pote(*), grade(*), hesse(*), timeinfo(*)
      integer(c_int), value :: precision
    end function

    type(c_ptr) function bfgt_last_error_c() bind(C, name="bfgt_last_error")
      import :: c_ptr
    end function

    integer(c_int) function bfgt_hip_release_plans_c() bind(C, name="bfgt_hip_release_plans")
      import :: c_int
    end function
  end interface

contains

  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=256) :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: i
    s = ' '
    if (.not. c_associated(p)) return
    call c_f_pointer(p, chars, [256])
    do i = 1, 256
      if (chars(i) == c_null_char) exit
      s(i:i) = chars(i)
    end do
  end function

end module boxfgt_hip_mod


subroutine boxfgt_set_precision(iprec)
  use boxfgt_hip_mod, only: boxfgt_precision
  implicit none
  integer, intent(in) :: iprec
  boxfgt_precision = iprec
end subroutine


subroutine boxfgt_release_plans()
  use iso_c_binding
  use boxfgt_hip_mod, only: bfgt_hip_release_plans_c
  implicit none
  integer(c_int) :: ier
  ier = bfgt_hip_release_plans_c()
end subroutine


subroutine boxfgt(nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, &
    nboxes, nlevels, ltree, itree, iptr, centers, boxsize, fvals, &
    ifpgh, pot, grad, hess, ifnewtree, ntarg, targs, &
    ifpghtarg, pote, grade, hesse, timeinfo)
  use iso_c_binding
  use boxfgt_hip_mod
  implicit none
  integer :: nd, ndim, ipoly, iperiod, norder, npbox, nboxes, nlevels, ltree
  real(8) :: delta, eps
  integer :: itree(ltree), iptr(8)
  real(8) :: centers(ndim, nboxes), boxsize(0:nlevels)
  real(8) :: fvals(nd, npbox, nboxes)
  integer :: ifpgh, ifnewtree, ntarg, ifpghtarg
  real(8) :: pot(nd, npbox, nboxes), grad(*), hess(*)
  real(8) :: targs(ndim, *), pote(*), grade(*), hesse(*), timeinfo(*)
  integer(c_int) :: ier

  ier = bfgt_hip_boxfgt_c(nd, ndim, delta, eps, ipoly, iperiod, norder, npbox, nboxes, nlevels, &
      ltree, itree, iptr, centers, boxsize, fvals, ifpgh, pot, grad, hess, ifnewtree, ntarg, &
      targs, ifpghtarg, pote, grade, hesse, timeinfo, boxfgt_precision)
  if (ier /= 0) then
    write(*, '(a,i0,a,a)') 'boxfgt (HIP): status ', ier, ': ', trim(c_string(bfgt_last_error_c()))
    error stop 'boxfgt (HIP) failed'
  end if
end subroutine
