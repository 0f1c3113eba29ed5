"""The allocation-free solve entry point (spral_ssids_hip_solve_device_noalloc)
against spral_ssids_hip_solve_device, through the C-ABI: two calls in a row,
the second on a smaller system with fewer right-hand sides, each bitwise
equal to the allocating entry's result (the same kernels on the same
operands; only the workspace differs), and a request beyond the fixed
workspace rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

from spral_amd import _lib, devmem, matrices, ssids
from helpers import analyse, rhs

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 8 << 20
ERROR_SCRATCH_TOO_SMALL = -61


def _factor(m):
    ak, inf = analyse(m, use_gpu=True)
    assert inf.flag >= 0
    opts = ssids.SSIDSOptions()
    fk, finf = ssids.ssids_factor(False, m.val, ak, opts)
    assert finf.flag >= 0
    return ak, fk


def _solve_both(num, n, b_cols, job):
    """b_cols (n, k): solve with both entries on device copies of b."""
    L = _lib.lib()
    k = b_cols.shape[1]
    host = np.ascontiguousarray(b_cols.T).reshape(-1)
    d_a = devmem.DeviceArray.from_host(host)
    d_b = devmem.DeviceArray.from_host(host)
    need = L.spral_ssids_hip_solve_workspace_doubles(num, k)
    assert 0 < need <= SCRATCH_DOUBLES, need
    f_a = L.spral_ssids_hip_solve_device(num, job, k, d_a.ptr, n, None)
    f_b = L.spral_ssids_hip_solve_device_noalloc(num, job, k, d_b.ptr, n, None)
    devmem.synchronize()
    assert f_a == 0 and f_b == 0, (f_a, f_b)
    return d_a.download().reshape(k, n).T, d_b.download().reshape(k, n).T


def test_noalloc_solve_equals_allocating_entry(lib):
    big = matrices.kkt_laplace(grid=10, m_con=200)       # n = 1,200
    small = matrices.random_indef(300, density=0.02, seed=21)
    ak1, fk1 = _factor(big)
    ak2, fk2 = _factor(small)
    _, b1 = rhs(big, 4, 3)
    _, b2 = rhs(small, 1, 5)
    print(f"noalloc call 1: n={big.n} nrhs=4 job=0", flush=True)
    xa, xb = _solve_both(fk1.numeric, big.n, b1[ak1.invp, :], 0)
    assert np.array_equal(xa, xb)
    assert matrices.bwd_error(big.full, _unperm(xb, ak1), b1).max() <= 1e-13
    print(f"noalloc call 2: n={small.n} nrhs=1 job=0", flush=True)
    xa, xb = _solve_both(fk2.numeric, small.n, b2[ak2.invp, :], 0)
    assert np.array_equal(xa, xb)
    assert matrices.bwd_error(small.full, _unperm(xb, ak2), b2).max() <= 1e-13
    for job in (1, 2, 3, 4):
        print(f"noalloc job={job}", flush=True)
        xa, xb = _solve_both(fk2.numeric, small.n, b2[ak2.invp, :], job)
        assert np.array_equal(xa, xb)
    ssids.ssids_free(ak1, fk1)
    ssids.ssids_free(ak2, fk2)


def test_noalloc_rejects_oversized_request(lib):
    m = matrices.random_indef(300, density=0.02, seed=21)
    ak, fk = _factor(m)
    L = _lib.lib()
    per = L.spral_ssids_hip_solve_workspace_doubles(fk.numeric, 1)
    k = SCRATCH_DOUBLES // per + 1
    # rejected on the host before any launch, so the buffer is never touched
    f = L.spral_ssids_hip_solve_device_noalloc(fk.numeric, 0, k, None, m.n, None)
    assert f == ERROR_SCRATCH_TOO_SMALL
    ssids.ssids_free(ak, fk)


def _unperm(x_elim, ak):
    out = np.empty_like(x_elim)
    out[ak.invp, :] = x_elim
    return out
