! C-callable shim around the reference scaling routines (test infrastructure
! only).
!
! Links against src/scaling.f90 and src/matrix_util.f90 compiled from
! /root/reference as they lie (see oracle/Makefile); this file is ours and
! only marshals arguments so the tests can compare the product's device
! scalings with the reference's equilib_scale_sym (src/scaling.f90:403-416,
! inf_norm_equilib_sym :480-521), auction_scale_sym (:269-309) and
! hungarian_scale_sym (:134-170).
! Matrices are the lower triangle in CSC with 1-based ptr/row, as the
! reference takes them.
module ref_scaling_shim
  use iso_c_binding
  use spral_scaling, only : equilib_scale_sym, equilib_options, equilib_inform, &
       auction_scale_sym, auction_options, auction_inform, hungarian_scale_sym, &
       hungarian_options, hungarian_inform
  implicit none
contains
  subroutine ref_equilib_scale_sym(n, ptr, row, val, scaling, max_iterations, &
       tol, iterations, flag) bind(C, name="ref_equilib_scale_sym")
    integer(c_int), value :: n, max_iterations
    integer(c_int64_t), intent(in) :: ptr(n+1)
    integer(c_int), intent(in) :: row(*)
    real(c_double), intent(in) :: val(*)
    real(c_double), intent(out) :: scaling(n)
    real(c_float), value :: tol
    integer(c_int), intent(out) :: iterations, flag
    type(equilib_options) :: options
    type(equilib_inform) :: inform
    options%max_iterations = max_iterations
    options%tol = tol
    call equilib_scale_sym(n, ptr, row, val, scaling, options, inform)
    iterations = inform%iterations
    flag = inform%flag
  end subroutine ref_equilib_scale_sym

  subroutine ref_auction_scale_sym(n, ptr, row, val, scaling, match, &
       matched, iterations, flag) bind(C, name="ref_auction_scale_sym")
    integer(c_int), value :: n
    integer(c_int64_t), intent(in) :: ptr(n+1)
    integer(c_int), intent(in) :: row(*)
    real(c_double), intent(in) :: val(*)
    real(c_double), intent(out) :: scaling(n)
    integer(c_int), intent(out) :: match(n)
    integer(c_int), intent(out) :: matched, iterations, flag
    type(auction_options) :: options
    type(auction_inform) :: inform
    call auction_scale_sym(n, ptr, row, val, scaling, options, inform, match)
    matched = inform%matched
    iterations = inform%iterations
    flag = inform%flag
  end subroutine ref_auction_scale_sym

  subroutine ref_hungarian_scale_sym(n, ptr, row, val, scaling, match, &
       scale_if_singular, matched, flag) bind(C, name="ref_hungarian_scale_sym")
    integer(c_int), value :: n, scale_if_singular
    integer(c_int64_t), intent(in) :: ptr(n+1)
    integer(c_int), intent(in) :: row(*)
    real(c_double), intent(in) :: val(*)
    real(c_double), intent(out) :: scaling(n)
    integer(c_int), intent(out) :: match(n)
    integer(c_int), intent(out) :: matched, flag
    type(hungarian_options) :: options
    type(hungarian_inform) :: inform
    options%scale_if_singular = (scale_if_singular /= 0)
    call hungarian_scale_sym(n, ptr, row, val, scaling, options, inform, match)
    matched = inform%matched
    flag = inform%flag
  end subroutine ref_hungarian_scale_sym
end module ref_scaling_shim
