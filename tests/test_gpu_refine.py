"""FP64 iterative refinement entry points: device buffers of any leading
dimension, and non-finite input reported instead of taken as converged."""
import ctypes as C

import numpy as np
import pytest

from spral_amd import _lib, devmem, matrices, ssids
from helpers import analyse, rhs

pytestmark = pytest.mark.gpu

BWD_TOL = 1e-14


def _factor(m):
    ak, inf = analyse(m, use_gpu=True)
    assert inf.flag >= 0
    opts = ssids.SSIDSOptions()
    fk, finf = ssids.ssids_factor(False, m.val, ak, opts)
    assert finf.flag >= 0
    return ak, fk, opts


def _refine_device(num, b_cols, ld, n, maxit=5, tol=BWD_TOL):
    """Refine with b (n, k) placed in device buffers of leading dimension ld."""
    k = b_cols.shape[1]
    host = np.zeros((k, ld))
    host[:, :n] = b_cols.T
    d_b = devmem.DeviceArray.from_host(host.reshape(-1))
    d_x = devmem.DeviceArray(ld * k)
    bwd = np.zeros(k)
    it = C.c_int(0)
    f = _lib.lib().spral_ssids_hip_solve_refine_device(num, k, d_b.ptr, d_x.ptr, ld, maxit, tol,
                                                       bwd, C.byref(it), None)
    devmem.synchronize()
    x = d_x.download().reshape(k, ld)[:, :n].T.copy()
    return f, x, bwd


def test_refine_device_leading_dimension_then_more_columns(lib):
    """A call with ld = 2n and one column, then ld = n and two columns, on one
    handle (the per-column norms grow with the column count, not with
    ld * nrhs): each column equals the host entry's result bit for bit."""
    m = matrices.kkt_laplace(grid=6, m_con=40)
    ak, fk, opts = _factor(m)
    n = m.n
    xt, b = rhs(m, 2, 31)
    be = b[ak.invp, :]  # elimination order, as the subtree sees it
    f1, x1, bwd1 = _refine_device(fk.numeric, be[:, :1], 2 * n, n)
    f2, x2, bwd2 = _refine_device(fk.numeric, be, n, n)
    assert f1 == 0 and f2 == 0
    for k in range(2):
        xh = b[:, k].copy()
        inf, bwd, its = ssids.ssids_solve_refine(xh, ak, fk, opts)
        assert inf.flag == 0 and bwd[0] <= BWD_TOL
        assert np.array_equal(x2[:, k], xh[ak.invp])
    assert np.array_equal(x1[:, 0], x2[:, 0])
    ssids.ssids_free(ak, fk)


def test_refine_reports_nonfinite(lib):
    """A NaN in one right-hand side is reported (flag -60, that column's bwd
    NaN), never counted as converged; the other column is unaffected."""
    m = matrices.kkt_laplace(grid=6, m_con=40)
    ak, fk, opts = _factor(m)
    xt, b = rhs(m, 2, 32)
    b[5, 1] = np.nan
    x = b.copy()
    inf, bwd, its = ssids.ssids_solve_refine(x, ak, fk, opts)
    assert inf.flag == ssids.REFINE_ERROR_NONFINITE
    assert np.isnan(bwd[1]) and np.isfinite(bwd[0])
    b[5, 1] = np.inf
    x = b.copy()
    inf, bwd, its = ssids.ssids_solve_refine(x, ak, fk, opts)
    assert inf.flag == ssids.REFINE_ERROR_NONFINITE and np.isnan(bwd[1])
    ssids.ssids_free(ak, fk)


def test_refine_reports_not_converged(lib):
    """No correction allowed and an unreachable tolerance: the warning flag,
    with the backward error of the unrefined solve."""
    m = matrices.random_indef(200, density=0.03, seed=3)
    ak, fk, opts = _factor(m)
    xt, b = rhs(m, 1, 33)
    x = b.copy()
    inf, bwd, its = ssids.ssids_solve_refine(x, ak, fk, opts, maxit=0, tol=1e-30)
    assert inf.flag == ssids.REFINE_WARNING_NOT_CONVERGED and its == 0
    assert 0 < bwd[0] < 1e-12
    ssids.ssids_free(ak, fk)
