! Test driver for the Fortran binding module spral_ssids_gpu_subtree: a
! Fortran host that reaches the HIP library only through the reference's
! abstract types symbolic_subtree_base / numeric_subtree_base
! (src/ssids/subtree.f90:26-119), as src/ssids/fkeep.F90 does.  For each
! matrix it prints one line
!   case n num_neg num_two rank bwd enquire_neg alter_same x1..x5(doc only)
! which the GPU test parses.
program hip_subtree_driver
  use, intrinsic :: iso_c_binding
  use spral_ssids_contrib, only : contrib_type
  use spral_ssids_datatypes
  use spral_ssids_inform, only : ssids_inform
  use spral_ssids_subtree, only : symbolic_subtree_base, numeric_subtree_base
  use spral_ssids_gpu_subtree
  implicit none

  interface
     type(C_PTR) function c_analyse(n, ptr, row, order, nemin, flag) &
          bind(C, name="spral_ssids_hip_analyse")
       import :: C_PTR, C_INT, C_INT64_T
       integer(C_INT), value :: n, nemin
       integer(C_INT64_T), dimension(*), intent(in) :: ptr
       integer(C_INT), dimension(*), intent(in) :: row
       integer(C_INT), dimension(*), intent(inout) :: order
       integer(C_INT), intent(out) :: flag
     end function c_analyse
     subroutine c_analyse_info(ak, nnodes, rlist_len, nlist_len, num_factor, num_flops, &
          maxfront, maxsupernode) bind(C, name="spral_ssids_hip_analyse_info")
       import :: C_PTR, C_INT, C_INT64_T
       type(C_PTR), value :: ak
       integer(C_INT), intent(out) :: nnodes, maxfront, maxsupernode
       integer(C_INT64_T), intent(out) :: rlist_len, nlist_len, num_factor, num_flops
     end subroutine c_analyse_info
     subroutine c_analyse_arrays(ak, sptr, sparent, rptr, rlist, nptr, nlist) &
          bind(C, name="spral_ssids_hip_analyse_arrays")
       import :: C_PTR, C_INT, C_INT64_T
       type(C_PTR), value :: ak
       integer(C_INT), dimension(*), intent(out) :: sptr, sparent, rlist
       integer(C_INT64_T), dimension(*), intent(out) :: rptr, nptr, nlist
     end subroutine c_analyse_arrays
     subroutine c_analyse_free(ak) bind(C, name="spral_ssids_hip_analyse_free")
       import :: C_PTR
       type(C_PTR), value :: ak
     end subroutine c_analyse_free
     integer(C_INT) function c_order_md(n, ptr, row, order) &
          bind(C, name="spral_ssids_hip_order_min_degree")
       import :: C_INT, C_INT64_T
       integer(C_INT), value :: n
       integer(C_INT64_T), dimension(*), intent(in) :: ptr
       integer(C_INT), dimension(*), intent(in) :: row
       integer(C_INT), dimension(*), intent(out) :: order
     end function c_order_md
  end interface

  integer(C_INT64_T), allocatable :: ptr(:)
  integer(C_INT), allocatable :: row(:)
  real(C_DOUBLE), allocatable :: val(:)
  integer :: n

  ! the 5x5 example of examples/Fortran/ssids.f90:30-42 (solution 1..5)
  n = 5
  allocate(ptr(6), row(9), val(9))
  ptr = [1, 3, 6, 8, 9, 10]
  row = [1, 2, 2, 3, 5, 3, 4, 4, 5]
  val = [2.0d0, 1.0d0, 4.0d0, 1.0d0, 1.0d0, 3.0d0, 2.0d0, -1.0d0, 2.0d0]
  call run_case("doc5", n, ptr, row, val)
  deallocate(ptr, row, val)

  ! 7-point Laplacian on a 7x6x5 box, diagonal 4 (indefinite)
  call laplace(7, 6, 5, 4.0d0, n, ptr, row, val)
  call run_case("lap765", n, ptr, row, val)

contains

  subroutine laplace(nx, ny, nz, diag, n, ptr, row, val)
    integer, intent(in) :: nx, ny, nz
    real(C_DOUBLE), intent(in) :: diag
    integer, intent(out) :: n
    integer(C_INT64_T), allocatable, intent(out) :: ptr(:)
    integer(C_INT), allocatable, intent(out) :: row(:)
    real(C_DOUBLE), allocatable, intent(out) :: val(:)
    integer :: i, j, k, c, nz_
    n = nx * ny * nz
    allocate(ptr(n + 1), row(4 * n), val(4 * n))
    nz_ = 0
    ptr(1) = 1
    do k = 0, nz - 1
       do j = 0, ny - 1
          do i = 0, nx - 1
             c = 1 + i + nx * (j + ny * k)
             nz_ = nz_ + 1; row(nz_) = c; val(nz_) = diag
             if (i + 1 < nx) then
                nz_ = nz_ + 1; row(nz_) = c + 1; val(nz_) = -1.0d0
             end if
             if (j + 1 < ny) then
                nz_ = nz_ + 1; row(nz_) = c + nx; val(nz_) = -1.0d0
             end if
             if (k + 1 < nz) then
                nz_ = nz_ + 1; row(nz_) = c + nx * ny; val(nz_) = -1.0d0
             end if
             ptr(c + 1) = nz_ + 1
          end do
       end do
    end do
  end subroutine laplace

  ! y = A x for the symmetric matrix held as its lower triangle
  subroutine symv(n, ptr, row, val, x, y)
    integer, intent(in) :: n
    integer(C_INT64_T), intent(in) :: ptr(:)
    integer(C_INT), intent(in) :: row(:)
    real(C_DOUBLE), intent(in) :: val(:), x(:)
    real(C_DOUBLE), intent(out) :: y(:)
    integer :: j, r
    integer(C_INT64_T) :: e
    y(1:n) = 0.0d0
    do j = 1, n
       do e = ptr(j), ptr(j + 1) - 1
          r = row(e)
          y(r) = y(r) + val(e) * x(j)
          if (r /= j) y(j) = y(j) + val(e) * x(r)
       end do
    end do
  end subroutine symv

  subroutine run_case(name, n, ptr, row, val)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    integer(C_INT64_T), intent(in) :: ptr(:)
    integer(C_INT), intent(in) :: row(:)
    real(C_DOUBLE), intent(in), target :: val(:)
    type(C_PTR) :: ak
    integer(C_INT) :: flag, nnodes, maxfront, maxsup, rc
    integer(C_INT64_T) :: rlen, nlen, nfact, nflops
    integer(C_INT), allocatable :: order(:), sptr(:), sparent(:), rlist(:)
    integer(C_INT64_T), allocatable :: rptr(:), nptr(:), nlist(:,:)
    class(symbolic_subtree_base), pointer :: sym
    class(numeric_subtree_base), pointer :: num
    type(contrib_type), allocatable :: child_contrib(:)
    type(ssids_options) :: options
    type(ssids_inform) :: inform
    real(C_DOUBLE), allocatable :: b(:), x(:), xe(:), r(:), x0(:), d(:,:)
    integer, allocatable :: piv(:)
    integer :: i, neg, k
    real(C_DOUBLE) :: anorm, bwd, det, tr
    logical :: same

    allocate(order(n))
    rc = c_order_md(int(n, C_INT), ptr, row, order)
    ak = c_analyse(int(n, C_INT), ptr, row, order, 32_C_INT, flag)
    if (.not. c_associated(ak)) stop 1
    call c_analyse_info(ak, nnodes, rlen, nlen, nfact, nflops, maxfront, maxsup)
    allocate(sptr(nnodes + 1), sparent(nnodes), rptr(nnodes + 1), rlist(rlen), &
         nptr(nnodes + 1), nlist(2, nlen))
    call c_analyse_arrays(ak, sptr, sparent, rptr, rlist, nptr, nlist)
    call c_analyse_free(ak)

    sym => construct_gpu_symbolic_subtree(0, n, 1, nnodes + 1, sptr, sparent, rptr, rlist, &
         nptr, nlist, options)
    if (.not. associated(sym)) stop 2
    allocate(child_contrib(0))
    inform%matrix_rank = n
    num => sym%factor(.false., val, child_contrib, options, inform)
    if (.not. associated(num)) stop 3

    ! b = A x_true with x_true = 1..n; x arrives in elimination order
    allocate(b(n), x(n), xe(n), r(n), x0(n), d(2, n), piv(n))
    do i = 1, n
       x0(i) = real(i, C_DOUBLE)
    end do
    call symv(n, ptr, row, val, x0, b)
    do i = 1, n
       xe(order(i)) = b(i)
    end do
    ! fkeep%inner_solve: fwd, then diag_bwd (src/ssids/fkeep.F90:269-297)
    call num%solve_fwd(1, xe, n, inform)
    call num%solve_diag_bwd(1, xe, n, inform)
    if (inform%flag < 0) stop 4
    do i = 1, n
       x(i) = xe(order(i))
    end do
    call symv(n, ptr, row, val, x, r)
    anorm = 0.0d0
    do i = 1, n
       r(i) = b(i) - r(i)
    end do
    call symv(n, ptr, row, abs(val), [(1.0d0, i = 1, n)], x0)
    anorm = maxval(x0(1:n))
    bwd = maxval(abs(r)) / (anorm * maxval(abs(x)) + maxval(abs(b)))

    ! enquire D^{-1} and count negative eigenvalues of its blocks; alter
    ! with the same D^{-1} leaves the solve unchanged bit for bit
    neg = -1
    same = .false.
    select type (num)
    type is (gpu_numeric_subtree)
       call num%enquire_indef(piv_order=piv, d=d)
       neg = 0
       k = 1
       do while (k <= n)
          if (d(2, k) /= 0.0d0 .and. k < n) then
             det = d(1, k) * d(1, k + 1) - d(2, k)**2
             tr = d(1, k) + d(1, k + 1)
             if (det < 0.0d0) then
                neg = neg + 1
             else if (tr < 0.0d0) then
                neg = neg + 2
             end if
             k = k + 2
          else
             if (d(1, k) < 0.0d0) neg = neg + 1
             k = k + 1
          end if
       end do
       call num%alter(d)
       do i = 1, n
          x0(order(i)) = b(i)
       end do
       call num%solve_fwd(1, x0, n, inform)
       call num%solve_diag_bwd(1, x0, n, inform)
       same = all(x0 == xe)
    end select

    if (name == "doc5") then
       write (*, "(a,1x,a,5(1x,i0),1x,es10.3,1x,i0,1x,l1,5(1x,f20.15))") "CASE", name, n, &
            inform%num_neg, inform%num_two, inform%matrix_rank, nnodes, bwd, neg, same, x(1:5)
    else
       write (*, "(a,1x,a,5(1x,i0),1x,es10.3,1x,i0,1x,l1)") "CASE", name, n, &
            inform%num_neg, inform%num_two, inform%matrix_rank, nnodes, bwd, neg, same
    end if

    call num%cleanup()
    deallocate(num)
    call sym%cleanup()
    deallocate(sym)
  end subroutine run_case

end program hip_subtree_driver
