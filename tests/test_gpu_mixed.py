"""Mixed-precision factorization (schur_precision = 1).

The contribution blocks are held in FP32, and the children's Schur updates
run on FP32 MFMA. The panel, the pivot blocks, their threshold tests, the
solves and the refinement residual stay FP64. src/ssids has no FP32
anywhere, so there is no reference routine to match; the bar is the one
the accuracy target states. Checked against the FP64 path of this build
(itself parity-checked against the CPU oracle):
* inertia (num_neg) and rank identical to the FP64 factorization;
* unrefined scaled backward error at FP32 level, <= MIXED_RAW_TOL;
* after FP64 iterative refinement against the original matrix,
  <= 1e-14 within 5 steps.
"""
import ctypes as C

import numpy as np
import pytest

from spral_amd import _lib, devmem, matrices, ssids
from helpers import analyse, rhs

pytestmark = pytest.mark.gpu

BWD_TOL = 1e-14
MIXED_RAW_TOL = 1e-5  # FP32 rounding of the Schur updates, amplified by growth


def factor_pair(m, grid=None):
    ak, inf = analyse(m, grid=grid, use_gpu=True)
    assert inf.flag >= 0, inf
    kw = {} if grid is None else {"ordering": 3}
    o64 = ssids.SSIDSOptions(**kw)
    o32 = ssids.SSIDSOptions(schur_precision=1, **kw)
    f64, i64 = ssids.ssids_factor(False, m.val, ak, o64)
    f32, i32 = ssids.ssids_factor(False, m.val, ak, o32)
    assert i64.flag >= 0 and i32.flag >= 0, (i64, i32)
    return ak, (f64, i64, o64), (f32, i32, o32)


def check_mixed(m, ak, mixed, nrhs=2, seed=0, max_steps=5):
    f32, i32, o32 = mixed
    xt, b = rhs(m, nrhs, seed)
    x = np.array(b, copy=True)
    assert ssids.ssids_solve(x, ak, f32, o32).flag >= 0
    raw = matrices.bwd_error(m.full, x, b).max()
    xr = np.array(b, copy=True)
    inf, bwd, its = ssids.ssids_solve_refine(xr, ak, f32, o32, maxit=max_steps, tol=BWD_TOL)
    ref = matrices.bwd_error(m.full, xr, b).max()
    print(f"  mixed: raw bwd {raw:.2e} refined {ref:.2e} ({its} steps)", flush=True)
    assert inf.flag >= 0
    assert raw <= MIXED_RAW_TOL
    assert ref <= BWD_TOL
    return raw, its


def test_mixed_stencil_grid(lib):
    m = matrices.laplace3d(16, diag=5.7)
    ak, (f64, i64, _), mixed = factor_pair(m, grid=(16, 16, 16))
    assert mixed[1].num_neg == i64.num_neg and mixed[1].matrix_rank == i64.matrix_rank == m.n
    raw, its = check_mixed(m, ak, mixed)
    assert raw > 1e-12  # the FP32 updates are in effect
    ssids.ssids_free(None, f64)
    ssids.ssids_free(ak, mixed[0])


def test_mixed_kkt_config1(lib):
    m = matrices.kkt_laplace()
    ak, (f64, i64, _), mixed = factor_pair(m)
    assert mixed[1].num_neg == i64.num_neg == 1024
    assert mixed[1].matrix_rank == m.n
    check_mixed(m, ak, mixed, seed=3)
    ssids.ssids_free(None, f64)
    ssids.ssids_free(ak, mixed[0])


def test_mixed_kkt_stencil27(lib):
    """The mixed-precision KKT workload at 12^3 (n = 2,025): H is SPD and A
    has full row rank (checked on the CPU in test_matrices), so both
    factorizations give exactly m_con negative pivots. Many constraint
    columns fail APP and go through the TPP fallback."""
    mc = 297
    m = matrices.kkt_stencil27(12, mc)
    ak, (f64, i64, o64), mixed = factor_pair(m)
    print(f"  delays fp64={i64.num_delay} mixed={mixed[1].num_delay}", flush=True)
    assert i64.num_neg == mixed[1].num_neg == mc
    assert i64.matrix_rank == mixed[1].matrix_rank == m.n
    xt, b = rhs(m, 1, 7)
    x = np.array(b, copy=True)
    ssids.ssids_solve_refine(x, ak, f64, o64, tol=BWD_TOL)
    assert matrices.bwd_error(m.full, x, b).max() <= BWD_TOL
    check_mixed(m, ak, mixed, seed=7)
    ssids.ssids_free(None, f64)
    ssids.ssids_free(ak, mixed[0])


@pytest.mark.parametrize("seed", range(4))
def test_mixed_random_with_delays(lib, seed):
    m = matrices.random_indef(200 + 60 * seed, density=0.03, seed=40 + seed)
    ev = np.linalg.eigvalsh(m.full.toarray())
    gap = np.min(np.abs(ev)) / np.max(np.abs(ev))
    ak, (f64, i64, _), mixed = factor_pair(m)
    print(f"  seed {seed}: delays fp64={i64.num_delay} mixed={mixed[1].num_delay} "
          f"min|ev|/max|ev| {gap:.1e}", flush=True)
    # inertia is fixed by the matrix whenever no eigenvalue is within the
    # FP32 perturbation of zero
    if gap > 1e-5:
        assert mixed[1].num_neg == i64.num_neg == int((ev < 0).sum())
        assert mixed[1].matrix_rank == i64.matrix_rank
    if mixed[1].matrix_rank == m.n and gap > 1e-5:
        check_mixed(m, ak, mixed, seed=seed)
    ssids.ssids_free(None, f64)
    ssids.ssids_free(ak, mixed[0])


def test_fp16_not_built(lib):
    m = matrices.laplace3d(4, diag=6.0)
    ak, inf = analyse(m, use_gpu=True)
    fk, finf = ssids.ssids_factor(False, m.val, ak, ssids.SSIDSOptions(schur_precision=2))
    assert finf.flag == ssids.SSIDS_ERROR_UNIMPLEMENTED
    assert not fk.numeric


def test_precision_switch_on_one_handle(lib):
    """Refactors on one numeric handle: FP64, mixed, FP64 again. The plan is
    rebuilt for each precision, and the two FP64 solutions are bitwise equal."""
    L = _lib.lib()
    m = matrices.laplace3d(12, diag=5.7)
    ak, inf = analyse(m, grid=(12, 12, 12), use_gpu=True)
    d_val = devmem.DeviceArray.from_host(m.val)
    xt, b = rhs(m, 1, 5)
    b = b[:, 0]
    o64 = ssids.SSIDSOptions(ordering=3).to_c()
    o32 = ssids.SSIDSOptions(ordering=3, schur_precision=1).to_c()
    st = _lib.Stats()
    num = L.spral_ssids_hip_factor_device(False, ak.symbolic, d_val.ptr, m.val.size, None,
                                          C.byref(o64), C.byref(st))
    assert num and st.flag >= 0

    def solve():
        x = np.array(b[ak.invp], copy=True)
        assert ssids._solve_all(L, num, 1, x, m.n) == 0  # fwd, then diag + bwd
        out = np.empty_like(x)
        out[ak.invp] = x
        return out

    x1 = solve()
    for o, tol in ((o32, MIXED_RAW_TOL), (o64, None)):
        s = _lib.Stats()
        assert L.spral_ssids_hip_refactor_device(num, d_val.ptr, None, C.byref(o),
                                                 C.byref(s)) >= 0
        x = solve()
        if tol is None:
            assert np.array_equal(x, x1)
        else:
            assert matrices.bwd_error(m.full, x, b).max() <= tol
            assert not np.array_equal(x, x1)
    L.spral_ssids_hip_destroy_numeric(num)
    ssids.ssids_free(ak)


@pytest.mark.timeout(900)
def test_kkt_stencil27_largest_grid(lib):
    """The 27-point KKT workload at 56^3 (n = 205,800; 1.77e14 flops; a
    71,616-row root front), the largest grid whose factors fit one GPU (80^3
    needs 287 GB of factor entries with any ordering tried), in the matching-
    based order with its scaling, FP64: exactly m_con negative pivots, full
    rank, and a refined solution at the 1e-14 bar."""
    g = 56
    mc = int(round(g ** 3 * 88000 / 512000))
    m = matrices.kkt_stencil27(g, mc)
    opts = ssids.SSIDSOptions(ordering=2, scaling=3)
    ak, ainf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, opts, val=m.val)
    assert ainf.flag >= 0, ainf
    fk, inf = ssids.ssids_factor(False, m.val, ak, opts)
    print(f"  56^3: flag {inf.flag}, delays {inf.num_delay}, "
          f"pool peak {inf.pool_peak_bytes / 2 ** 30:.1f} GiB", flush=True)
    assert inf.flag >= 0, inf
    assert inf.num_neg == mc
    assert inf.matrix_rank == m.n
    xt, b = rhs(m, 1, 56)
    x = np.array(b, copy=True)
    ssids.ssids_solve_refine(x, ak, fk, opts, tol=BWD_TOL)
    assert matrices.bwd_error(m.full, x, b).max() <= BWD_TOL
    ssids.ssids_free(ak, fk)
