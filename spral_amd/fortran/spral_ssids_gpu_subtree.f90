!> Binding module that replaces src/ssids/gpu/subtree.f90 (module
!> spral_ssids_gpu_subtree) of the reference: the same public names
!> (gpu_symbolic_subtree, construct_gpu_symbolic_subtree, gpu_numeric_subtree,
!> gpu_free_contrib, as used by src/ssids/anal.F90:19 and
!> src/ssids/contrib_free.f90:13), the deferred procedures of
!> symbolic_subtree_base / numeric_subtree_base (src/ssids/subtree.f90:26-119),
!> each one call through iso_c_binding into libspral_ssids_hip.so
!> (include/spral_ssids_hip.h).  All numeric work runs in the HIP library.
module spral_ssids_gpu_subtree
  use, intrinsic :: iso_c_binding
  use spral_ssids_contrib,   only : contrib_type
  use spral_ssids_datatypes
  use spral_ssids_inform,    only : ssids_inform
  use spral_ssids_subtree,   only : symbolic_subtree_base, numeric_subtree_base
  implicit none

  private
  public :: gpu_symbolic_subtree, construct_gpu_symbolic_subtree
  public :: gpu_numeric_subtree, gpu_free_contrib
  public :: hip_options, hip_stats

  !> struct spral_ssids_hip_options
  type, bind(C) :: hip_options
     integer(C_INT)    :: print_level, action
     real(C_DOUBLE)    :: small, u, multiplier
     integer(C_INT)    :: pivot_method, failed_pivot_method, schur_precision
     integer(C_INT)    :: refine_max
     real(C_DOUBLE)    :: refine_tol
     integer(C_INT)    :: timing, level_schedule
  end type hip_options

  !> struct spral_ssids_hip_stats
  type, bind(C) :: hip_stats
     integer(C_INT)     :: flag, num_delay
     integer(C_INT64_T) :: num_factor, num_flops
     integer(C_INT)     :: num_neg, num_two, num_zero, maxfront, maxsupernode
     integer(C_INT)     :: not_first_pass, not_second_pass, hip_error, num_levels
     real(C_DOUBLE)     :: factor_ms, kernel_ms, kernel_flops, kernel_bytes
     integer(C_INT)     :: kernel_launches
     integer(C_INT64_T) :: pool_peak_bytes
     integer(C_INT)     :: static_plan
  end type hip_stats

  interface
     subroutine hip_default_options(opt) bind(C, name="spral_ssids_hip_default_options")
       import :: hip_options
       type(hip_options), intent(out) :: opt
     end subroutine hip_default_options
     type(C_PTR) function hip_create_symbolic(device, n, sa, en, sptr, sparent, rptr, &
          rlist, nptr, nlist, opt, flag) bind(C, name="spral_ssids_hip_create_symbolic")
       import :: C_PTR, C_INT, C_INT64_T, hip_options
       integer(C_INT), value :: device, n, sa, en
       integer(C_INT), dimension(*), intent(in) :: sptr, sparent, rlist
       integer(C_INT64_T), dimension(*), intent(in) :: rptr, nptr, nlist
       type(hip_options), intent(in) :: opt
       integer(C_INT), intent(out) :: flag
     end function hip_create_symbolic
     subroutine hip_destroy_symbolic(s) bind(C, name="spral_ssids_hip_destroy_symbolic")
       import :: C_PTR
       type(C_PTR), value :: s
     end subroutine hip_destroy_symbolic
     type(C_PTR) function hip_factor(posdef, s, aval, naval, scaling, opt, st) &
          bind(C, name="spral_ssids_hip_factor")
       import :: C_PTR, C_BOOL, C_INT64_T, hip_options, hip_stats
       logical(C_BOOL), value :: posdef
       type(C_PTR), value :: s, aval, scaling
       integer(C_INT64_T), value :: naval
       type(hip_options), intent(in) :: opt
       type(hip_stats), intent(out) :: st
     end function hip_factor
     subroutine hip_destroy_numeric(f) bind(C, name="spral_ssids_hip_destroy_numeric")
       import :: C_PTR
       type(C_PTR), value :: f
     end subroutine hip_destroy_numeric
     integer(C_INT) function hip_solve_fwd(f, nrhs, x, ldx) &
          bind(C, name="spral_ssids_hip_solve_fwd")
       import :: C_PTR, C_INT, C_DOUBLE
       type(C_PTR), value :: f
       integer(C_INT), value :: nrhs, ldx
       real(C_DOUBLE), dimension(*), intent(inout) :: x
     end function hip_solve_fwd
     integer(C_INT) function hip_solve_diag(f, nrhs, x, ldx) &
          bind(C, name="spral_ssids_hip_solve_diag")
       import :: C_PTR, C_INT, C_DOUBLE
       type(C_PTR), value :: f
       integer(C_INT), value :: nrhs, ldx
       real(C_DOUBLE), dimension(*), intent(inout) :: x
     end function hip_solve_diag
     integer(C_INT) function hip_solve_diag_bwd(f, nrhs, x, ldx) &
          bind(C, name="spral_ssids_hip_solve_diag_bwd")
       import :: C_PTR, C_INT, C_DOUBLE
       type(C_PTR), value :: f
       integer(C_INT), value :: nrhs, ldx
       real(C_DOUBLE), dimension(*), intent(inout) :: x
     end function hip_solve_diag_bwd
     integer(C_INT) function hip_solve_bwd(f, nrhs, x, ldx) &
          bind(C, name="spral_ssids_hip_solve_bwd")
       import :: C_PTR, C_INT, C_DOUBLE
       type(C_PTR), value :: f
       integer(C_INT), value :: nrhs, ldx
       real(C_DOUBLE), dimension(*), intent(inout) :: x
     end function hip_solve_bwd
     integer(C_INT) function hip_enquire_indef(f, piv_order, d) &
          bind(C, name="spral_ssids_hip_enquire_indef")
       import :: C_PTR, C_INT
       type(C_PTR), value :: f, piv_order, d
     end function hip_enquire_indef
     integer(C_INT) function hip_alter(f, d) bind(C, name="spral_ssids_hip_alter")
       import :: C_PTR, C_INT, C_DOUBLE
       type(C_PTR), value :: f
       real(C_DOUBLE), dimension(2,*), intent(in) :: d
     end function hip_alter
     integer(C_INT) function hip_enquire_posdef(f, d) &
          bind(C, name="spral_ssids_hip_enquire_posdef")
       import :: C_PTR, C_INT
       type(C_PTR), value :: f, d
     end function hip_enquire_posdef
     integer(C_INT) function hip_get_contrib(f, n, val, ldval, rlist, ndelay, delay_perm, &
          delay_val, lddelay) bind(C, name="spral_ssids_hip_get_contrib")
       import :: C_PTR, C_INT
       type(C_PTR), value :: f
       integer(C_INT), intent(out) :: n, ldval, ndelay, lddelay
       type(C_PTR), intent(out) :: val, rlist, delay_perm, delay_val
     end function hip_get_contrib
     subroutine hip_free_contrib(f) bind(C, name="spral_ssids_hip_free_contrib")
       import :: C_PTR
       type(C_PTR), value :: f
     end subroutine hip_free_contrib
  end interface

  !> The analysed subtree, held by the library (row lists, child maps and
  !> the A-entry map uploaded to HBM once).
  type, extends(symbolic_subtree_base) :: gpu_symbolic_subtree
     type(C_PTR) :: handle = C_NULL_PTR
     integer :: n = 0
     integer(long) :: max_a_idx = 0    ! entries of aval the subtree reads
   contains
     procedure :: factor
     procedure :: cleanup => symbolic_cleanup
  end type gpu_symbolic_subtree

  !> The factors, resident in HBM and owned by the library's handle.
  type, extends(numeric_subtree_base) :: gpu_numeric_subtree
     type(C_PTR) :: handle = C_NULL_PTR
     logical :: posdef = .false.
     integer :: n = 0
   contains
     procedure :: get_contrib
     procedure :: solve_fwd
     procedure :: solve_diag
     procedure :: solve_diag_bwd
     procedure :: solve_bwd
     procedure :: enquire_posdef
     procedure :: enquire_indef
     procedure :: alter
     procedure :: cleanup => numeric_cleanup
  end type gpu_numeric_subtree

contains

  !> Options the numeric path honours, from ssids_options (as
  !> cpu_copy_options_in, src/ssids/cpu/cpu_iface.f90:57-69); the rest keep
  !> the library's defaults.
  subroutine to_c_options(foptions, copt)
    type(ssids_options), intent(in) :: foptions
    type(hip_options), intent(out) :: copt
    call hip_default_options(copt)
    copt%print_level = foptions%print_level
    copt%action = merge(1_C_INT, 0_C_INT, foptions%action)
    copt%small = foptions%small
    copt%u = foptions%u
    copt%multiplier = foptions%multiplier
    copt%pivot_method = min(3, max(1, foptions%pivot_method))
    copt%failed_pivot_method = min(2, max(1, foptions%failed_pivot_method))
  end subroutine to_c_options

  !> Fold the subtree's statistics into inform, as cpu_copy_stats_out
  !> (src/ssids/cpu/cpu_iface.f90:74-94).
  subroutine merge_stats(st, inform)
    type(hip_stats), intent(in) :: st
    type(ssids_inform), intent(inout) :: inform
    if (st%flag < 0) then
       inform%flag = min(inform%flag, int(st%flag))
    else
       inform%flag = max(inform%flag, int(st%flag))
    end if
    inform%num_delay = inform%num_delay + st%num_delay
    inform%num_factor = inform%num_factor + st%num_factor
    inform%num_flops = inform%num_flops + st%num_flops
    inform%num_neg = inform%num_neg + st%num_neg
    inform%num_two = inform%num_two + st%num_two
    inform%maxfront = max(inform%maxfront, int(st%maxfront))
    inform%maxsupernode = max(inform%maxsupernode, int(st%maxsupernode))
    inform%not_first_pass = inform%not_first_pass + st%not_first_pass
    inform%not_second_pass = inform%not_second_pass + st%not_second_pass
    inform%matrix_rank = inform%matrix_rank - st%num_zero
    if (st%hip_error /= 0) inform%cuda_error = st%hip_error
  end subroutine merge_stats

  !> A negative library return code as an ssids_inform flag.
  subroutine set_solve_flag(rc, inform)
    integer(C_INT), intent(in) :: rc
    type(ssids_inform), intent(inout) :: inform
    if (rc < 0) inform%flag = min(inform%flag, int(rc))
  end subroutine set_solve_flag

  !> Replaces construct_gpu_symbolic_subtree (src/ssids/gpu/subtree.f90:76-160).
  function construct_gpu_symbolic_subtree(device, n, sa, en, sptr, sparent, rptr, rlist, &
       nptr, nlist, options) result(this)
    class(gpu_symbolic_subtree), pointer :: this
    integer, intent(in) :: device, n, sa, en
    integer, dimension(*), target, intent(in) :: sptr, sparent, rlist
    integer(long), dimension(*), target, intent(in) :: rptr, nptr
    integer(long), dimension(2,*), target, intent(in) :: nlist
    class(ssids_options), intent(in) :: options
    type(gpu_symbolic_subtree), pointer :: sym
    type(hip_options) :: copt
    integer(C_INT) :: flag
    integer(long) :: k
    integer :: st
    nullify(this)
    allocate(sym, stat=st)
    if (st /= 0) return
    select type (options)
    type is (ssids_options)
       call to_c_options(options, copt)
    class default
       call hip_default_options(copt)
    end select
    sym%n = n
    sym%max_a_idx = 0
    do k = nptr(sa), nptr(en) - 1
       sym%max_a_idx = max(sym%max_a_idx, nlist(1, k))
    end do
    sym%handle = hip_create_symbolic(int(device, C_INT), int(n, C_INT), int(sa, C_INT), &
         int(en, C_INT), sptr, sparent, rptr, rlist, nptr, nlist, copt, flag)
    if (.not. c_associated(sym%handle)) then
       deallocate(sym)
       return
    end if
    this => sym
  end function construct_gpu_symbolic_subtree

  !> Replaces gpu_symbolic_subtree%factor (src/ssids/gpu/subtree.f90:304-460).
  !> Contribution blocks from child subtrees held elsewhere are not taken:
  !> the whole tree is one part on the GPU (options%gpu_only).
  function factor(this, posdef, aval, child_contrib, options, inform, scaling)
    class(numeric_subtree_base), pointer :: factor
    class(gpu_symbolic_subtree), target, intent(inout) :: this
    logical, intent(in) :: posdef
    real(wp), dimension(*), target, intent(in) :: aval
    type(contrib_type), dimension(:), target, intent(inout) :: child_contrib
    type(ssids_options), intent(in) :: options
    type(ssids_inform), intent(inout) :: inform
    real(wp), dimension(*), target, optional, intent(in) :: scaling
    type(gpu_numeric_subtree), pointer :: num
    type(hip_options) :: copt
    type(hip_stats) :: st
    type(C_PTR) :: sc
    integer :: stat
    nullify(factor)
    if (size(child_contrib) > 0) then
       inform%flag = SSIDS_ERROR_UNIMPLEMENTED
       return
    end if
    call to_c_options(options, copt)
    sc = C_NULL_PTR
    if (present(scaling)) sc = c_loc(scaling)
    allocate(num, stat=stat)
    if (stat /= 0) then
       inform%flag = SSIDS_ERROR_ALLOCATION
       inform%stat = stat
       return
    end if
    num%posdef = posdef
    num%n = this%n
    num%handle = hip_factor(logical(posdef, C_BOOL), this%handle, c_loc(aval), &
         int(this%max_a_idx, C_INT64_T), sc, copt, st)
    call merge_stats(st, inform)
    if (.not. c_associated(num%handle)) then
       deallocate(num)
       return
    end if
    factor => num
  end function factor

  !> Replaces the symbolic cleanup (src/ssids/gpu/subtree.f90:236-302).
  subroutine symbolic_cleanup(this)
    class(gpu_symbolic_subtree), intent(inout) :: this
    if (c_associated(this%handle)) call hip_destroy_symbolic(this%handle)
    this%handle = C_NULL_PTR
  end subroutine symbolic_cleanup

  !> Replaces get_contrib (src/ssids/gpu/subtree.f90:522-538): the subtree
  !> root's block (n x n, ldval) and delayed columns, in host memory owned
  !> by the numeric handle until gpu_free_contrib.
  function get_contrib(this)
    type(contrib_type) :: get_contrib
    class(gpu_numeric_subtree), intent(in) :: this
    integer(C_INT) :: n, ldval, ndelay, lddelay, flag
    type(C_PTR) :: val, rlist, delay_perm, delay_val
    flag = hip_get_contrib(this%handle, n, val, ldval, rlist, ndelay, delay_perm, &
         delay_val, lddelay)
    ! contrib_type has no error channel and the parent subtree spins on
    ! %ready: a device failure while exporting the block stops the run
    if (flag /= 0) then
       write (0, "(a,i0)") "spral_ssids_hip_get_contrib failed, flag ", flag
       error stop 1
    end if
    get_contrib%n = n
    get_contrib%ldval = ldval
    get_contrib%ndelay = ndelay
    get_contrib%lddelay = lddelay
    nullify(get_contrib%val, get_contrib%rlist, get_contrib%delay_perm, &
         get_contrib%delay_val)
    if (n > 0) then
       call c_f_pointer(val, get_contrib%val, [int(ldval) * int(n)])
       call c_f_pointer(rlist, get_contrib%rlist, [int(n)])
    end if
    if (ndelay > 0) then
       call c_f_pointer(delay_perm, get_contrib%delay_perm, [int(ndelay)])
       call c_f_pointer(delay_val, get_contrib%delay_val, [int(lddelay) * int(ndelay)])
    end if
    get_contrib%owner = 1  ! gpu
    get_contrib%posdef = logical(this%posdef, C_BOOL)
    get_contrib%owner_ptr = this%handle
    get_contrib%ready = .true.
  end function get_contrib

  !> Replaces gpu_free_contrib: the memory belongs to the numeric handle.
  subroutine gpu_free_contrib(contrib)
    type(contrib_type), intent(inout) :: contrib
    if (c_associated(contrib%owner_ptr)) call hip_free_contrib(contrib%owner_ptr)
    nullify(contrib%val, contrib%rlist, contrib%delay_perm, contrib%delay_val)
    contrib%owner_ptr = C_NULL_PTR
  end subroutine gpu_free_contrib

  !> Replaces gpu_numeric_subtree%solve_fwd (src/ssids/gpu/subtree.f90:555-600);
  !> x(ldx,nrhs) arrives in elimination order (src/ssids/fkeep.F90:252-266),
  !> every right-hand side solved in one pass.
  subroutine solve_fwd(this, nrhs, x, ldx, inform)
    class(gpu_numeric_subtree), intent(inout) :: this
    integer, intent(in) :: nrhs
    real(wp), dimension(*), intent(inout) :: x
    integer, intent(in) :: ldx
    type(ssids_inform), intent(inout) :: inform
    call set_solve_flag(hip_solve_fwd(this%handle, int(nrhs, C_INT), x, int(ldx, C_INT)), &
         inform)
  end subroutine solve_fwd

  subroutine solve_diag(this, nrhs, x, ldx, inform)
    class(gpu_numeric_subtree), intent(inout) :: this
    integer, intent(in) :: nrhs
    real(wp), dimension(*), intent(inout) :: x
    integer, intent(in) :: ldx
    type(ssids_inform), intent(inout) :: inform
    call set_solve_flag(hip_solve_diag(this%handle, int(nrhs, C_INT), x, int(ldx, C_INT)), &
         inform)
  end subroutine solve_diag

  subroutine solve_diag_bwd(this, nrhs, x, ldx, inform)
    class(gpu_numeric_subtree), intent(inout) :: this
    integer, intent(in) :: nrhs
    real(wp), dimension(*), intent(inout) :: x
    integer, intent(in) :: ldx
    type(ssids_inform), intent(inout) :: inform
    call set_solve_flag(hip_solve_diag_bwd(this%handle, int(nrhs, C_INT), x, &
         int(ldx, C_INT)), inform)
  end subroutine solve_diag_bwd

  subroutine solve_bwd(this, nrhs, x, ldx, inform)
    class(gpu_numeric_subtree), intent(inout) :: this
    integer, intent(in) :: nrhs
    real(wp), dimension(*), intent(inout) :: x
    integer, intent(in) :: ldx
    type(ssids_inform), intent(inout) :: inform
    call set_solve_flag(hip_solve_bwd(this%handle, int(nrhs, C_INT), x, int(ldx, C_INT)), &
         inform)
  end subroutine solve_bwd

  !> Replaces enquire_posdef (src/ssids/gpu/subtree.f90:702-738): the
  !> diagonal of the Cholesky factor in pivot order.
  subroutine enquire_posdef(this, d)
    class(gpu_numeric_subtree), target, intent(in) :: this
    real(wp), dimension(*), target, intent(out) :: d
    integer(C_INT) :: rc
    rc = hip_enquire_posdef(this%handle, c_loc(d))
  end subroutine enquire_posdef

  !> Replaces enquire_indef (src/ssids/gpu/subtree.f90:740-844): piv_order
  !> 0-based (the caller adds 1, src/ssids/ssids.f90:1343-1352), d(1:2,:) the
  !> blocks of D^{-1} in pivot order.
  subroutine enquire_indef(this, piv_order, d)
    class(gpu_numeric_subtree), target, intent(in) :: this
    integer, dimension(*), target, optional, intent(out) :: piv_order
    real(wp), dimension(2,*), target, optional, intent(out) :: d
    type(C_PTR) :: pp, dp
    integer(C_INT) :: rc
    pp = C_NULL_PTR
    dp = C_NULL_PTR
    if (present(piv_order)) pp = c_loc(piv_order)
    if (present(d)) dp = c_loc(d)
    rc = hip_enquire_indef(this%handle, pp, dp)
  end subroutine enquire_indef

  !> Replaces alter (src/ssids/gpu/subtree.f90:909-981).
  subroutine alter(this, d)
    class(gpu_numeric_subtree), target, intent(inout) :: this
    real(wp), dimension(2,*), intent(in) :: d
    integer(C_INT) :: rc
    rc = hip_alter(this%handle, d)
  end subroutine alter

  !> Replaces the numeric cleanup (src/ssids/gpu/subtree.f90:462-520).
  subroutine numeric_cleanup(this)
    class(gpu_numeric_subtree), intent(inout) :: this
    if (c_associated(this%handle)) call hip_destroy_numeric(this%handle)
    this%handle = C_NULL_PTR
  end subroutine numeric_cleanup

end module spral_ssids_gpu_subtree
