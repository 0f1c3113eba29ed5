! C-callable shim around the reference Rutherford-Boeing routines (test
! infrastructure only).
!
! Links against src/rutherford_boeing.f90, src/matrix_util.f90 and
! src/random.f90 compiled from /root/reference as they lie (see
! oracle/Makefile); this file is ours and only marshals arguments, so that
! the tests can compare the product's RB writer and reader with rb_write
! (src/rutherford_boeing.f90:635-732) and rb_read (:274-575).  Two calls
! for a read: ref_rb_read keeps the arrays and returns the sizes,
! ref_rb_fetch copies them out and releases them.
module ref_rb_shim
  use iso_c_binding
  use spral_rutherford_boeing
  implicit none
  integer(c_int64_t), allocatable, save :: g_ptr(:)
  integer(c_int), allocatable, save :: g_row(:)
  real(c_double), allocatable, save :: g_val(:)
contains
  function cstr(s, n) result(f)
    character(kind=c_char), intent(in) :: s(*)
    integer(c_int), intent(in) :: n
    character(len=n) :: f
    integer :: i
    do i = 1, n
      f(i:i) = s(i)
    end do
  end function cstr

  subroutine ref_rb_write(fname, flen, matrix_type, m, n, ptr, row, val, has_val, &
       info) bind(C, name="ref_rb_write")
    character(kind=c_char), intent(in) :: fname(*)
    integer(c_int), value :: flen, matrix_type, m, n, has_val
    integer(c_int64_t), intent(in) :: ptr(n+1)
    integer(c_int), intent(in) :: row(*)
    real(c_double), intent(in) :: val(*)
    integer(c_int), intent(out) :: info
    type(rb_write_options) :: options
    integer(c_int64_t) :: nz
    nz = ptr(n+1) - 1
    if (has_val /= 0) then
      call rb_write(cstr(fname, flen), matrix_type, m, n, ptr, row(1:nz), options, info, &
           val=val(1:nz))
    else
      call rb_write(cstr(fname, flen), matrix_type, m, n, ptr, row(1:nz), options, info)
    end if
  end subroutine ref_rb_write

  subroutine ref_rb_read(fname, flen, lwr_upr_full, values, add_diagonal, m, n, nz, &
       has_val, matrix_type, info) bind(C, name="ref_rb_read")
    character(kind=c_char), intent(in) :: fname(*)
    integer(c_int), value :: flen, lwr_upr_full, values, add_diagonal
    integer(c_int), intent(out) :: m, n, has_val, matrix_type, info
    integer(c_int64_t), intent(out) :: nz
    type(rb_read_options) :: options
    integer(c_int64_t), allocatable :: ptr(:)
    integer(c_int), allocatable :: row(:)
    real(c_double), allocatable :: val(:)
    options%lwr_upr_full = lwr_upr_full
    options%values = values
    options%add_diagonal = (add_diagonal /= 0)
    if (allocated(g_ptr)) deallocate(g_ptr)
    if (allocated(g_row)) deallocate(g_row)
    if (allocated(g_val)) deallocate(g_val)
    call rb_read(cstr(fname, flen), m, n, ptr, row, val, options, info, &
         matrix_type=matrix_type)
    nz = 0
    has_val = 0
    if (info < 0) return
    nz = ptr(n+1) - 1
    allocate(g_ptr(n+1), g_row(nz))
    g_ptr(:) = ptr(1:n+1)
    g_row(:) = row(1:nz)
    if (allocated(val)) then
      has_val = 1
      allocate(g_val(nz))
      g_val(:) = val(1:nz)
    end if
  end subroutine ref_rb_read

  subroutine ref_rb_fetch(ptr, row, val) bind(C, name="ref_rb_fetch")
    integer(c_int64_t), intent(out) :: ptr(*)
    integer(c_int), intent(out) :: row(*)
    real(c_double), intent(out) :: val(*)
    ptr(1:size(g_ptr)) = g_ptr(:)
    row(1:size(g_row)) = g_row(:)
    if (allocated(g_val)) val(1:size(g_val)) = g_val(:)
  end subroutine ref_rb_fetch
end module ref_rb_shim
