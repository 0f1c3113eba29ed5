! C-callable shim around the reference analyse (test infrastructure only).
!
! Links against src/core_analyse.f90 compiled from /root/reference as it lies
! (see oracle/Makefile); this file is ours and only marshals arguments so the
! tests can compare the product's analyse with the reference's
! basic_analyse (src/core_analyse.f90:38-151) integer-for-integer.
! Two calls: ref_analyse_run computes and keeps the result, returning sizes;
! ref_analyse_fetch copies the arrays out and releases them.
module ref_analyse_shim
  use iso_c_binding
  use spral_core_analyse, only : basic_analyse
  implicit none
  integer, allocatable, save :: g_sptr(:), g_sparent(:), g_rlist(:), g_perm(:)
  integer(c_int64_t), allocatable, save :: g_rptr(:)
  integer, save :: g_nnodes = 0
contains
  subroutine ref_analyse_run(n, ptr, row, perm, nemin, nnodes, rlist_len, &
       nfact, nflops, info) bind(C, name="ref_analyse_run")
    integer(c_int), value :: n, nemin
    integer(c_int64_t), intent(in) :: ptr(n+1)
    integer(c_int), intent(in) :: row(*)
    integer(c_int), intent(in) :: perm(n)
    integer(c_int), intent(out) :: nnodes, info
    integer(c_int64_t), intent(out) :: rlist_len, nfact, nflops
    integer :: st
    if (allocated(g_perm)) deallocate(g_perm)
    allocate(g_perm(n))
    g_perm(:) = perm(:)
    call basic_analyse(n, ptr, row, g_perm, g_nnodes, g_sptr, g_sparent, &
         g_rptr, g_rlist, nemin, info, st, nfact, nflops)
    nnodes = g_nnodes
    rlist_len = 0
    if (info >= 0) rlist_len = g_rptr(g_nnodes+1) - 1
  end subroutine ref_analyse_run

  subroutine ref_analyse_fetch(n, perm, sptr, sparent, rptr, rlist) &
       bind(C, name="ref_analyse_fetch")
    integer(c_int), value :: n
    integer(c_int), intent(out) :: perm(n), sptr(*), sparent(*), rlist(*)
    integer(c_int64_t), intent(out) :: rptr(*)
    integer(c_int64_t) :: nr
    perm(1:n) = g_perm(1:n)
    sptr(1:g_nnodes+1) = g_sptr(1:g_nnodes+1)
    sparent(1:g_nnodes) = g_sparent(1:g_nnodes)
    rptr(1:g_nnodes+1) = g_rptr(1:g_nnodes+1)
    nr = g_rptr(g_nnodes+1) - 1
    rlist(1:nr) = g_rlist(1:nr)
    deallocate(g_perm, g_sptr, g_sparent, g_rptr, g_rlist)
  end subroutine ref_analyse_fetch
end module ref_analyse_shim
