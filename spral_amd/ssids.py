"""Host-side mirror of the SSIDS API (src/ssids/ssids.f90) over the HIP C-ABI.

Names, argument meaning and flag values follow the reference:
  ssids_analyse        src/ssids/ssids.f90:148-389   (analyse_double)
  ssids_factor         src/ssids/ssids.f90:767-1109  (ssids_factor_ptr64_double)
  ssids_solve          src/ssids/ssids.f90:1140-1250 (ssids_solve_mult_double)
  ssids_enquire_indef  src/ssids/ssids.f90:1300-1356
  ssids_alter          src/ssids/ssids.f90:1362-1399
  ssids_free           src/ssids/ssids.f90:1403-1434
The numeric work runs on the GPU through libspral_ssids_hip.so; there is no
CPU fallback.  Matrices are lower-triangle CSC with 1-based ptr/row as in the
reference.  Orderings: options.ordering = 0 takes the user's order (as the
reference); 1 (the reference's METIS) uses this package's exact minimum degree
up to n = AMD_THRESHOLD and its approximate minimum degree above, since METIS is
not available; 2 is the reference's matching-based order (maximum-product
matching, matched pairs ordered together; needs val) with approximate minimum
degree on the compressed graph; 3 is geometric nested dissection for an
nx*ny*nz grid (pass grid=(nx,ny,nz)); 4 is the approximate minimum degree at
any size.  Scalings: options.scaling 1 is the matching (Hungarian) scaling, 2
the auction scaling, 3 the scaling of ordering 2, >= 4 norm equilibration on
the device.
"""
import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np

from . import _lib

# flags (src/ssids/datatypes.f90:25-59)
SSIDS_SUCCESS = 0
SSIDS_ERROR_CALL_SEQUENCE = -1
SSIDS_ERROR_A_N_OOR = -2
SSIDS_ERROR_A_PTR = -3
SSIDS_ERROR_SINGULAR = -5
SSIDS_ERROR_NOT_POS_DEF = -6
SSIDS_ERROR_ORDER = -8
SSIDS_ERROR_VAL = -9
SSIDS_ERROR_X_SIZE = -10
SSIDS_ERROR_JOB_OOR = -11
SSIDS_ERROR_NOT_LLT = -13
SSIDS_ERROR_NOT_LDLT = -14
SSIDS_ERROR_NO_SAVED_SCALING = -15
SSIDS_ERROR_ALLOCATION = -50
SSIDS_ERROR_CUDA_UNKNOWN = -51
SSIDS_ERROR_UNIMPLEMENTED = -98
SSIDS_WARNING_ANAL_SINGULAR = 6
SSIDS_WARNING_FACT_SINGULAR = 7
SSIDS_WARNING_MATCH_ORD_NO_SCALE = 8

SSIDS_SOLVE_JOB_ALL = 0
SSIDS_SOLVE_JOB_FWD = 1
SSIDS_SOLVE_JOB_DIAG = 2
SSIDS_SOLVE_JOB_BWD = 3
SSIDS_SOLVE_JOB_DIAG_BWD = 4


@dataclass
class SSIDSOptions:
    """Subset of ssids_options (src/ssids/datatypes.f90:188-284) the path uses."""
    print_level: int = 0
    ordering: int = 1
    nemin: int = 32
    action: bool = True
    pivot_method: int = 2
    small: float = 1e-20
    # default-real literals in the reference (src/ssids/datatypes.f90:262,268):
    # single-precision 0.01 and 1.1 widened to double
    u: float = float(np.float32(0.01))
    multiplier: float = float(np.float32(1.1))
    failed_pivot_method: int = 1
    scaling: int = 0
    use_gpu: bool = True
    schur_precision: int = 0
    device: int = 0
    # tree levels of the factorization (no reference counterpart): 0 the
    # cheapest by the cost model, 1 height from the leaves, 2 greedy, 3
    # layered; ssids_autotune picks one by timing them on the device
    level_schedule: int = 0

    def to_c(self):
        o = _lib.default_options()
        o.print_level = self.print_level
        o.action = 1 if self.action else 0
        o.small = self.small
        o.u = self.u
        o.multiplier = self.multiplier
        o.pivot_method = min(3, max(1, self.pivot_method))
        o.failed_pivot_method = min(2, max(1, self.failed_pivot_method))
        o.schur_precision = self.schur_precision
        o.level_schedule = self.level_schedule
        return o


@dataclass
class SSIDSInform:
    """Subset of ssids_inform (src/ssids/inform.f90:17-50)."""
    flag: int = 0
    matrix_rank: int = 0
    maxfront: int = 0
    maxsupernode: int = 0
    num_delay: int = 0
    num_factor: int = 0
    num_flops: int = 0
    num_neg: int = 0
    num_sup: int = 0
    num_two: int = 0
    not_first_pass: int = 0
    not_second_pass: int = 0
    cuda_error: int = 0
    factor_ms: float = 0.0
    num_levels: int = 0
    pool_peak_bytes: int = 0  # peak device pool use of the factorization
    static_plan: int = 0      # 1: the factorization ran on the static plan


@dataclass
class AKeep:
    n: int = 0
    ptr: np.ndarray = None
    row: np.ndarray = None
    order: np.ndarray = None   # final pivot order (1-based; 0 unused)
    invp: np.ndarray = None    # invp[i] = variable at pivot position i (0-based)
    nnodes: int = 0
    sptr: np.ndarray = None
    sparent: np.ndarray = None
    rptr: np.ndarray = None
    rlist: np.ndarray = None
    nptr: np.ndarray = None
    nlist: np.ndarray = None
    num_factor: int = 0
    num_flops: int = 0
    symbolic: int = None       # C handle of the device symbolic subtree
    inform: SSIDSInform = field(default_factory=SSIDSInform)
    nz: int = 0
    scaling: np.ndarray = None  # matching scaling of ordering = 2 (user order)
    level_schedule: int = 0    # schedule of the symbolic handle (ssids_autotune)
    schedule_ms: dict = field(default_factory=dict)  # timed schedules: ms each


@dataclass
class FKeep:
    numeric: int = None
    pos_def: bool = False
    symbolic: int = None       # the analyse handle the factors belong to
    scaling: np.ndarray = None
    inform: SSIDSInform = field(default_factory=SSIDSInform)


def _as_i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# METIS stand-ins (options.ordering = 1): exact minimum degree up to this
# order, approximate minimum degree above it (or always, with ordering = 4)
AMD_THRESHOLD = 20000


def ssids_analyse(check, n, ptr, row, options=None, order=None, grid=None, val=None):
    """Analyse the sparsity pattern; returns (akeep, inform).  val is needed
    for the matching-based order (options.ordering = 2), as in the
    reference."""
    options = options or SSIDSOptions()
    L = _lib.lib()
    inform = SSIDSInform()
    ak = AKeep(n=n)
    if n < 0:
        inform.flag = SSIDS_ERROR_A_N_OOR
        ak.inform = inform
        return ak, inform
    ptr = _as_i64(ptr)
    row = _as_i32(row)
    ak.ptr, ak.row, ak.nz = ptr, row, int(ptr[-1] - 1)
    if options.ordering == 2 and val is None:
        inform.flag = SSIDS_ERROR_VAL
        ak.inform = inform
        return ak, inform
    if options.ordering == 0:
        if order is None:
            inform.flag = SSIDS_ERROR_ORDER
            ak.inform = inform
            return ak, inform
        ordv = _as_i32(np.abs(order)).copy()
    elif options.ordering == 2:
        # matching-based order (src/ssids/ssids.f90:318-340): the matching's
        # pairs take consecutive positions; its scaling is kept for
        # options.scaling = 3 in the factorization
        ordv = np.zeros(n, np.int32)
        ak.scaling = np.zeros(n)
        mf = L.spral_match_order(n, ptr, row, np.ascontiguousarray(val, np.float64), ordv,
                                 ak.scaling)
        if mf < 0:
            inform.flag = SSIDS_ERROR_ALLOCATION
            ak.inform = inform
            return ak, inform
        if mf == 1:
            inform.flag = SSIDS_WARNING_ANAL_SINGULAR
    elif options.ordering == 3:
        nx, ny, nz = grid
        ordv = np.zeros(n, np.int32)
        L.spral_ssids_hip_order_grid_nd(nx, ny, nz, ordv)
    elif options.ordering == 4 or n > AMD_THRESHOLD:
        # approximate minimum degree: the exact one's explicit graph grows
        # with the fill
        ordv = np.zeros(n, np.int32)
        L.spral_ssids_hip_order_amd(n, ptr, row, ordv)
    else:
        ordv = np.zeros(n, np.int32)
        L.spral_ssids_hip_order_min_degree(n, ptr, row, ordv)
    flag = C.c_int(0)
    h = L.spral_ssids_hip_analyse(n, ptr, row, ordv, options.nemin, C.byref(flag))
    if not h:
        inform.flag = flag.value
        ak.inform = inform
        return ak, inform
    try:
        nn, rl, nl = C.c_int(), C.c_int64(), C.c_int64()
        nf, fl, mf, ms = C.c_int64(), C.c_int64(), C.c_int(), C.c_int()
        L.spral_ssids_hip_analyse_info(h, C.byref(nn), C.byref(rl), C.byref(nl), C.byref(nf),
                                       C.byref(fl), C.byref(mf), C.byref(ms))
        ak.nnodes = nn.value
        ak.sptr = np.zeros(nn.value + 1, np.int32)
        ak.sparent = np.zeros(nn.value, np.int32)
        ak.rptr = np.zeros(nn.value + 1, np.int64)
        ak.rlist = np.zeros(rl.value, np.int32)
        ak.nptr = np.zeros(nn.value + 1, np.int64)
        ak.nlist = np.zeros(2 * nl.value, np.int64)
        L.spral_ssids_hip_analyse_arrays(h, ak.sptr, ak.sparent, ak.rptr, ak.rlist, ak.nptr,
                                         ak.nlist)
        ak.num_factor, ak.num_flops = nf.value, fl.value
        inform.maxfront, inform.maxsupernode = mf.value, ms.value
    finally:
        L.spral_ssids_hip_analyse_free(h)
    ak.order = ordv
    invp = np.zeros(n, np.int64)
    used = ordv > 0
    invp[ordv[used] - 1] = np.nonzero(used)[0]
    ak.invp = invp
    inform.flag = max(inform.flag, flag.value)
    inform.num_factor, inform.num_flops = ak.num_factor, ak.num_flops
    inform.num_sup = ak.nnodes
    inform.matrix_rank = int(ak.sptr[-1] - 1) if ak.nnodes else 0
    # device symbolic subtree covering the whole tree
    if options.use_gpu and ak.nnodes > 0:
        f = C.c_int(0)
        opts = options.to_c()
        sym = L.spral_ssids_hip_create_symbolic(options.device, n, 1, ak.nnodes + 1, ak.sptr,
                                                ak.sparent, ak.rptr, ak.rlist, ak.nptr,
                                                ak.nlist, C.byref(opts), C.byref(f))
        if not sym:
            inform.flag = f.value
        ak.symbolic = sym
        ak.level_schedule = options.level_schedule
    ak.inform = inform
    return ak, inform


def _inform_from_stats(st, akeep):
    inf = SSIDSInform()
    inf.flag = st.flag
    inf.num_delay = st.num_delay
    inf.num_factor = st.num_factor
    inf.num_flops = st.num_flops
    inf.num_neg = st.num_neg
    inf.num_two = st.num_two
    inf.maxfront = st.maxfront
    inf.maxsupernode = st.maxsupernode
    inf.not_first_pass = st.not_first_pass
    inf.not_second_pass = st.not_second_pass
    inf.cuda_error = st.hip_error
    inf.matrix_rank = akeep.inform.matrix_rank - st.num_zero
    inf.num_sup = akeep.nnodes
    inf.factor_ms = st.factor_ms
    inf.num_levels = st.num_levels
    inf.pool_peak_bytes = st.pool_peak_bytes
    inf.static_plan = st.static_plan
    return inf


def equilib_scale_sym(n, ptr, row, val, max_iterations=10, tol=1e-8):
    """Symmetric infinity-norm equilibration of the lower CSC (1-based ptr,
    row) on the device, as equilib_scale_sym (src/scaling.f90:403-416).
    Returns (scaling, iterations); tol is used at single precision, as the
    reference's equilib_options%tol."""
    L = _lib.lib()
    ptr = _as_i64(ptr)
    row = _as_i32(row)
    val = np.ascontiguousarray(val, dtype=np.float64)
    s = np.zeros(max(n, 1))
    its = C.c_int(0)
    rc = L.spral_ssids_hip_equilib_scale_sym(n, ptr.ctypes.data, row.ctypes.data,
                                             val.ctypes.data, s.ctypes.data,
                                             int(max_iterations), float(tol), C.byref(its))
    if rc != 0:
        raise RuntimeError(f"spral_ssids_hip_equilib_scale_sym failed ({rc})")
    return s[:n], its.value


def _match_order_warning(inf, akeep, options):
    """Every factorization of a matching-based analyse that does not use the
    matching's scaling (options.scaling != 3) warns, refactors included
    (src/ssids/ssids.f90:915-918); a factorization warning of its own takes
    precedence."""
    if inf.flag == SSIDS_SUCCESS and akeep.scaling is not None and options.scaling != 3:
        inf.flag = SSIDS_WARNING_MATCH_ORD_NO_SCALE


def ssids_factor(posdef, val, akeep, options=None, scale=None, fkeep=None):
    """Numeric factorization on the GPU; returns (fkeep, inform)."""
    options = options or SSIDSOptions()
    L = _lib.lib()
    if akeep.inform.flag < 0 or akeep.symbolic is None:
        inf = SSIDSInform(flag=SSIDS_ERROR_CALL_SEQUENCE)
        return FKeep(inform=inf), inf
    val = np.ascontiguousarray(val, dtype=np.float64)
    st = _lib.Stats()
    opts = options.to_c()
    sc = None
    scp = None
    if options.scaling >= 4:
        # norm equilibration (MC77-like), as src/ssids/ssids.f90:1003-1028:
        # computed on the device from the values, returned through `scale`
        s = equilib_scale_sym(akeep.n, akeep.ptr, akeep.row, val)[0]
        if scale is not None:
            scale[:] = s
        scale = s
    elif options.scaling == 1:
        # maximum-product matching scaling (src/ssids/ssids.f90:922-950),
        # scale_if_singular = options.action
        s = np.zeros(akeep.n)
        nm = C.c_int(0)
        hp, hr = _as_i64(akeep.ptr), _as_i32(akeep.row)
        hf = L.spral_hungarian_scale_sym(akeep.n, hp.ctypes.data, hr.ctypes.data,
                                         val.ctypes.data, s.ctypes.data, None,
                                         1 if options.action else 0, C.byref(nm))
        if hf == -2:
            inf = SSIDSInform(flag=SSIDS_ERROR_SINGULAR)
            return FKeep(inform=inf), inf
        if hf == -1:
            inf = SSIDSInform(flag=SSIDS_ERROR_ALLOCATION)
            return FKeep(inform=inf), inf
        if scale is not None:
            scale[:] = s
        scale = s
    elif options.scaling == 3:
        # the scaling of the matching-based order (src/ssids/ssids.f90:979-987)
        if getattr(akeep, "scaling", None) is None:
            inf = SSIDSInform(flag=SSIDS_ERROR_NO_SAVED_SCALING)
            return FKeep(inform=inf), inf
        if scale is not None:
            scale[:] = akeep.scaling
        scale = akeep.scaling
    elif options.scaling == 2:
        # auction scaling (src/ssids/ssids.f90:952-977) with the default
        # auction_options (src/scaling.f90:33-38)
        s = np.zeros(akeep.n)
        hp, hr = _as_i64(akeep.ptr), _as_i32(akeep.row)
        mu = np.array([10, 100, 100], np.int32)
        mp = np.array([0.90, 0.0, 0.0], np.float32)
        nm, it, un = C.c_int(0), C.c_int(0), C.c_int(0)
        af = L.spral_auction_scale_sym(akeep.n, hp.ctypes.data, hr.ctypes.data,
                                       val.ctypes.data, s.ctypes.data, None, 30000,
                                       mu.ctypes.data, mp.ctypes.data,
                                       float(np.float32(0.01)), C.byref(nm), C.byref(it),
                                       C.byref(un))
        if af < 0:
            inf = SSIDSInform(flag=SSIDS_ERROR_ALLOCATION)
            return FKeep(inform=inf), inf
        if scale is not None:
            scale[:] = s
        scale = s
    if scale is not None:
        # scaling is applied in elimination order inside the subtree
        sc = np.ascontiguousarray(np.asarray(scale, np.float64)[akeep.invp])
        scp = sc.ctypes.data_as(C.c_void_p)
    if (fkeep is not None and fkeep.numeric and fkeep.symbolic == akeep.symbolic
            and fkeep.pos_def == bool(posdef)):
        # a re-factorization of the same analyse reuses the handle: its device
        # pool, its static plan and the delays seen so far (the reference's
        # fkeep is intent(inout) in ssids_factor)
        from . import devmem
        d_val = devmem.DeviceArray.from_host(val)
        d_sc = devmem.DeviceArray.from_host(sc) if sc is not None else None
        L.spral_ssids_hip_refactor_device(fkeep.numeric, d_val.ptr,
                                          d_sc.ptr if d_sc is not None else None,
                                          C.byref(opts), C.byref(st))
        inf = _inform_from_stats(st, akeep)
        _match_order_warning(inf, akeep, options)
        fkeep.scaling, fkeep.inform = sc, inf
        if inf.flag < 0:
            L.spral_ssids_hip_destroy_numeric(fkeep.numeric)
            fkeep.numeric = None
        elif inf.flag == SSIDS_WARNING_FACT_SINGULAR and not options.action:
            inf.flag = SSIDS_ERROR_SINGULAR
        return fkeep, inf
    if fkeep is not None and fkeep.numeric:
        L.spral_ssids_hip_destroy_numeric(fkeep.numeric)
        fkeep.numeric = None
    num = L.spral_ssids_hip_factor(bool(posdef), akeep.symbolic, val, val.size, scp,
                                   C.byref(opts), C.byref(st))
    inf = _inform_from_stats(st, akeep)
    if num:
        _match_order_warning(inf, akeep, options)
    fk = FKeep(numeric=num, pos_def=bool(posdef), scaling=sc, inform=inf,
               symbolic=akeep.symbolic)
    if num and inf.flag == SSIDS_WARNING_FACT_SINGULAR and not options.action:
        inf.flag = SSIDS_ERROR_SINGULAR
    return fk, inf


def ssids_solve(x, akeep, fkeep, options=None, job=None):
    """Solve in place; x is (n,) or (n, nrhs).  Returns inform."""
    options = options or SSIDSOptions()
    L = _lib.lib()
    inf = SSIDSInform()
    if fkeep is None or not fkeep.numeric or fkeep.inform.flag < 0:
        inf.flag = SSIDS_ERROR_CALL_SEQUENCE
        return inf
    inf.flag = max(SSIDS_SUCCESS, fkeep.inform.flag)
    job = SSIDS_SOLVE_JOB_ALL if job is None else job
    if job < 0 or job > 4:
        inf.flag = SSIDS_ERROR_JOB_OOR
        return inf
    if fkeep.pos_def and job in (SSIDS_SOLVE_JOB_DIAG, SSIDS_SOLVE_JOB_DIAG_BWD):
        inf.flag = SSIDS_ERROR_JOB_OOR  # src/ssids/ssids.f90:1237-1240
        return inf
    xa = np.asarray(x)
    one = xa.ndim == 1
    X = xa.reshape(akeep.n, -1)
    n = akeep.n
    x2 = np.asfortranarray(X[akeep.invp, :], dtype=np.float64)
    if fkeep.scaling is not None and job in (0, 1):
        x2 *= fkeep.scaling[:, None]
    x2c = np.ascontiguousarray(x2.T)  # (nrhs, n) row-major == (n, nrhs) col-major
    fn = {0: None, 1: L.spral_ssids_hip_solve_fwd, 2: L.spral_ssids_hip_solve_diag,
          3: L.spral_ssids_hip_solve_bwd, 4: L.spral_ssids_hip_solve_diag_bwd}[job]
    nrhs = x2c.shape[0]
    if job == 0:
        # fkeep%inner_solve: solve_fwd then solve_diag_bwd (src/ssids/fkeep.F90:269-297)
        flag = _solve_all(L, fkeep.numeric, nrhs, x2c, n)
    else:
        flag = fn(fkeep.numeric, nrhs, x2c.reshape(-1), n)
    if flag < 0:
        inf.flag = flag
        return inf
    res = x2c.T
    if fkeep.scaling is not None and job in (0, 3, 4):
        res = res * fkeep.scaling[:, None]
    out = np.empty_like(res)
    out[akeep.invp, :] = res
    if one:
        xa[:] = out[:, 0]
    else:
        X[:, :] = out
    return inf


REFINE_WARNING_NOT_CONVERGED = 60  # a column above tol after maxit corrections
REFINE_ERROR_NONFINITE = -60       # a NaN or infinity in b, x or the residual


def ssids_solve_refine(x, akeep, fkeep, options=None, maxit=5, tol=1e-14):
    """Solve with FP64 iterative refinement on the device (no counterpart in
    the reference).  x (n,) or (n, nrhs) holds b on entry and the solution on
    exit.  Returns (inform, bwd per RHS, refinement steps); inform.flag is
    REFINE_WARNING_NOT_CONVERGED when some column stays above tol and
    REFINE_ERROR_NONFINITE when a column is not finite."""
    L = _lib.lib()
    inf = SSIDSInform()
    if fkeep is None or not fkeep.numeric or fkeep.inform.flag < 0:
        inf.flag = SSIDS_ERROR_CALL_SEQUENCE
        return inf, None, 0
    inf.flag = max(SSIDS_SUCCESS, fkeep.inform.flag)
    xa = np.asarray(x)
    one = xa.ndim == 1
    X = xa.reshape(akeep.n, -1)
    n, nrhs = akeep.n, X.shape[1]
    b2 = np.ascontiguousarray(np.asarray(X[akeep.invp, :], np.float64).T)
    x2 = np.zeros_like(b2)
    bwd = np.zeros(nrhs)
    it = C.c_int(0)
    flag = L.spral_ssids_hip_solve_refine(fkeep.numeric, nrhs, b2.reshape(-1), x2.reshape(-1),
                                          n, maxit, tol, bwd, C.byref(it))
    if flag < 0:
        # SPRAL_SSIDS_HIP_ERROR_NONFINITE (-60): a NaN or infinity in b, x or
        # the residual; bwd of that column is NaN
        inf.flag = flag
        return inf, (bwd if flag == REFINE_ERROR_NONFINITE else None), it.value
    if flag == REFINE_WARNING_NOT_CONVERGED:
        inf.flag = max(inf.flag, flag)
    out = np.empty((n, nrhs))
    out[akeep.invp, :] = x2.T
    if one:
        xa[:] = out[:, 0]
    else:
        X[:, :] = out
    return inf, bwd, it.value


def _solve_all(L, num, nrhs, x2c, n):
    # src/ssids/fkeep.F90:269-297: fwd then diag_bwd (for L L^T the subtree's
    # diag_bwd is its bwd)
    flag = L.spral_ssids_hip_solve_fwd(num, nrhs, x2c.reshape(-1), n)
    if flag < 0:
        return flag
    return L.spral_ssids_hip_solve_diag_bwd(num, nrhs, x2c.reshape(-1), n)


def ssids_enquire_indef(akeep, fkeep, options=None):
    """Returns (piv_order (1-based, negative for 2x2 members), d (2,n), inform)
    as ssids_enquire_indef (src/ssids/ssids.f90:1300-1356)."""
    L = _lib.lib()
    inf = SSIDSInform()
    if fkeep is None or not fkeep.numeric:
        inf.flag = SSIDS_ERROR_CALL_SEQUENCE
        return None, None, inf
    if fkeep.pos_def:
        inf.flag = SSIDS_ERROR_NOT_LDLT
        return None, None, inf
    n = akeep.n
    po = np.zeros(n, np.int32)
    d = np.zeros(2 * n, np.float64)
    flag = L.spral_ssids_hip_enquire_indef(fkeep.numeric, po.ctypes.data_as(C.c_void_p),
                                           d.ctypes.data_as(C.c_void_p))
    if flag < 0:
        inf.flag = flag
        return None, None, inf
    piv = np.zeros(n, np.int32)
    piv[akeep.invp] = po  # subtree result is by elimination position
    piv = np.where(piv >= 0, piv + 1, piv - 1)
    return piv, d.reshape(n, 2).T.copy(), inf


def ssids_enquire_posdef(akeep, fkeep, options=None):
    """Returns (d, inform): the diagonal of the Cholesky factor in pivot order,
    as ssids_enquire_posdef (src/ssids/ssids.f90:1256-1291)."""
    L = _lib.lib()
    inf = SSIDSInform()
    if fkeep is None or not fkeep.numeric:
        inf.flag = SSIDS_ERROR_CALL_SEQUENCE
        return None, inf
    if not fkeep.pos_def:
        inf.flag = SSIDS_ERROR_NOT_LLT
        return None, inf
    d = np.zeros(akeep.n, np.float64)
    flag = L.spral_ssids_hip_enquire_posdef(fkeep.numeric, d)
    if flag < 0:
        inf.flag = flag
        return None, inf
    return d, inf


def ssids_alter(d, akeep, fkeep, options=None):
    """Overwrite D^{-1} (d is (2, n) in pivot order)."""
    L = _lib.lib()
    inf = SSIDSInform()
    if fkeep is None or not fkeep.numeric:
        inf.flag = SSIDS_ERROR_CALL_SEQUENCE
        return inf
    if fkeep.pos_def:
        inf.flag = SSIDS_ERROR_NOT_LDLT
        return inf
    dd = np.ascontiguousarray(np.asarray(d, np.float64).T.reshape(-1))
    inf.flag = L.spral_ssids_hip_alter(fkeep.numeric, dd)
    return inf


def ssids_free(akeep=None, fkeep=None):
    L = _lib.lib()
    if fkeep is not None and fkeep.numeric:
        L.spral_ssids_hip_destroy_numeric(fkeep.numeric)
        fkeep.numeric = None
    if akeep is not None and akeep.symbolic:
        L.spral_ssids_hip_destroy_symbolic(akeep.symbolic)
        akeep.symbolic = None
    return 0


def ssids_autotune(akeep, val, options=None, schedules=(1, 2, 3), reps=2, posdef=False,
                   scale=None):
    """Schedule autotuner (no reference counterpart).  Builds a device
    symbolic handle per tree-level schedule (options.level_schedule: 1 height
    from the leaves, 2 greedy by urgency, 3 layered heavy fronts), factors val
    on each (one factorization, then reps refactors on the same handle, as a
    resident pattern is refactored), and keeps in akeep the handle whose best
    refactor took the least device time, replacing the analyse's handle.
    Every schedule factors each front from the same inputs in the same order,
    so the factors do not depend on the choice.  Call after ssids_analyse and
    before the first ssids_factor of akeep; scale is passed to each
    ssids_factor as given.  Returns {schedule: ms}; akeep records the choice
    (level_schedule) and the times (schedule_ms)."""
    options = options or SSIDSOptions()
    if akeep.inform.flag < 0 or akeep.symbolic is None:
        raise ValueError("ssids_autotune needs a successful ssids_analyse on the GPU")
    if not schedules or reps < 1:
        raise ValueError("ssids_autotune needs schedules and reps >= 1")
    L = _lib.lib()
    original = akeep.symbolic
    handles, times = {}, {}
    try:
        for s in schedules:
            o = replace(options, level_schedule=int(s))
            f = C.c_int(0)
            h = L.spral_ssids_hip_create_symbolic(
                o.device, akeep.n, 1, akeep.nnodes + 1, _as_i32(akeep.sptr),
                _as_i32(akeep.sparent), _as_i64(akeep.rptr), _as_i32(akeep.rlist),
                _as_i64(akeep.nptr), _as_i64(akeep.nlist), C.byref(o.to_c()), C.byref(f))
            if not h:
                raise RuntimeError(f"symbolic handle for schedule {s}: flag {f.value}")
            handles[int(s)] = h
            akeep.symbolic = h
            fk, best = None, float("inf")
            try:
                for r in range(reps + 1):
                    fk, inf = ssids_factor(posdef, val, akeep, o, fkeep=fk,
                                           scale=None if scale is None else scale.copy())
                    if inf.flag < 0:
                        raise RuntimeError(f"factorization under schedule {s}: flag {inf.flag}")
                    if r > 0:
                        best = min(best, float(inf.factor_ms))
            finally:
                if fk is not None:
                    ssids_free(fkeep=fk)
            times[int(s)] = best
    except Exception:
        for h in handles.values():
            L.spral_ssids_hip_destroy_symbolic(h)
        akeep.symbolic = original
        raise
    pick = min(times, key=lambda s: (times[s], s))
    for s, h in handles.items():
        if s != pick:
            L.spral_ssids_hip_destroy_symbolic(h)
    L.spral_ssids_hip_destroy_symbolic(original)
    akeep.symbolic = handles[pick]
    akeep.level_schedule = pick
    akeep.schedule_ms = times
    return times


# ---------------------------------------------------------------------------
# Analyse residency: an analysed pattern saved and reloaded, and a cache of
# analyses keyed by pattern for a resident service.  Arrays only (numpy .npz,
# nothing pickled); loading rebuilds the device symbolic handle.
# ---------------------------------------------------------------------------
_AKEEP_ARRAYS = ("ptr", "row", "order", "invp", "sptr", "sparent", "rptr", "rlist", "nptr",
                 "nlist")


def ssids_save_akeep(akeep, path):
    """Writes the analyse result (the arrays of akeep and its inform sizes)."""
    extra = {} if akeep.scaling is None else {"scaling": akeep.scaling}
    np.savez(path, n=akeep.n, nnodes=akeep.nnodes, num_factor=akeep.num_factor,
             num_flops=akeep.num_flops, nz=akeep.nz, flag=akeep.inform.flag,
             maxfront=akeep.inform.maxfront, maxsupernode=akeep.inform.maxsupernode,
             matrix_rank=akeep.inform.matrix_rank, level_schedule=akeep.level_schedule,
             **{k: getattr(akeep, k) for k in _AKEEP_ARRAYS}, **extra)


def ssids_load_akeep(path, options=None):
    """Reads a saved analyse; with options.use_gpu the device symbolic handle
    is rebuilt from the arrays (no ordering or analyse is rerun), under the
    saved level schedule unless options.level_schedule names another."""
    options = options or SSIDSOptions()
    z = np.load(path, allow_pickle=False)
    ak = AKeep(n=int(z["n"]))
    if "level_schedule" in z.files:
        ak.level_schedule = int(z["level_schedule"])
    if options.level_schedule == 0 and ak.level_schedule != 0:
        options = replace(options, level_schedule=ak.level_schedule)
    ak.level_schedule = options.level_schedule
    for k in _AKEEP_ARRAYS:
        setattr(ak, k, z[k])
    ak.nnodes, ak.nz = int(z["nnodes"]), int(z["nz"])
    ak.num_factor, ak.num_flops = int(z["num_factor"]), int(z["num_flops"])
    if "scaling" in z.files:
        ak.scaling = z["scaling"]
    inf = SSIDSInform(flag=int(z["flag"]), maxfront=int(z["maxfront"]),
                      maxsupernode=int(z["maxsupernode"]), matrix_rank=int(z["matrix_rank"]),
                      num_factor=ak.num_factor, num_flops=ak.num_flops, num_sup=ak.nnodes)
    if options.use_gpu and ak.nnodes > 0:
        L = _lib.lib()
        f = C.c_int(0)
        opts = options.to_c()
        ak.symbolic = L.spral_ssids_hip_create_symbolic(
            options.device, ak.n, 1, ak.nnodes + 1, _as_i32(ak.sptr), _as_i32(ak.sparent),
            _as_i64(ak.rptr), _as_i32(ak.rlist), _as_i64(ak.nptr), _as_i64(ak.nlist),
            C.byref(opts), C.byref(f))
        if not ak.symbolic:
            inf.flag = f.value
    ak.inform = inf
    return ak, inf


def ssids_save_fkeep(fkeep, path):
    """Writes the resident factors of fkeep (no reference counterpart: the
    reference keeps factors in memory only).  Returns 0, or -1 on an I/O
    error."""
    if fkeep is None or not fkeep.numeric:
        return SSIDS_ERROR_CALL_SEQUENCE
    return _lib.lib().spral_ssids_hip_save_numeric(fkeep.numeric, str(path).encode())


def ssids_load_fkeep(akeep, path):
    """Rebuilds resident factors saved by ssids_save_fkeep on akeep (an
    analyse of the same pattern and ordering: its device symbolic handle must
    exist).  Returns (fkeep, inform); inform.flag < 0 when the file is
    unreadable (-1), belongs to another analyse (-2) or the device failed."""
    L = _lib.lib()
    if akeep is None or akeep.symbolic is None:
        inf = SSIDSInform(flag=SSIDS_ERROR_CALL_SEQUENCE)
        return FKeep(inform=inf), inf
    st = _lib.Stats()
    f = C.c_int(0)
    num = L.spral_ssids_hip_load_numeric(akeep.symbolic, str(path).encode(), C.byref(st),
                                         C.byref(f))
    if not num:
        inf = SSIDSInform(flag=f.value)
        return FKeep(inform=inf), inf
    inf = _inform_from_stats(st, akeep)
    sc = np.zeros(akeep.n)
    has = L.spral_ssids_hip_numeric_scaling(num, sc)
    en = _lib.lib().spral_ssids_hip_enquire_posdef(num, np.zeros(max(akeep.n, 1)))
    fk = FKeep(numeric=num, pos_def=(en == 0), scaling=(sc if has == 1 else None), inform=inf,
               symbolic=akeep.symbolic)
    return fk, inf


class AnalyseCache:
    """Analyses kept per sparsity pattern: a pattern seen before (same n,
    ptr, row and ordering options) returns its resident akeep without a new
    ordering, analyse or device upload."""

    def __init__(self):
        self._store = {}

    @staticmethod
    def key(n, ptr, row, options, grid=None):
        import hashlib
        h = hashlib.sha256()
        h.update(np.int64(n).tobytes())
        h.update(_as_i64(ptr).tobytes())
        h.update(_as_i32(row).tobytes())
        h.update(repr((options.ordering, options.nemin, options.device, grid)).encode())
        return h.hexdigest()

    def analyse(self, n, ptr, row, options=None, grid=None, val=None):
        """val: the values the matching-based order (options.ordering = 2)
        is computed from, for the first analyse of the pattern."""
        options = options or SSIDSOptions()
        k = self.key(n, ptr, row, options, grid)
        if k not in self._store:
            self._store[k] = ssids_analyse(False, n, ptr, row, options, grid=grid, val=val)
        return self._store[k]

    def __len__(self):
        return len(self._store)

    def free(self):
        for ak, _ in self._store.values():
            ssids_free(ak)
        self._store.clear()
