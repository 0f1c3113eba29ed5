"""GPU equilibration scaling (options%scaling >= 4) against the oracle.

The device routine replaces equilib_scale_sym (src/scaling.f90:403-416;
inf_norm_equilib_sym :480-521).  Every step of the algorithm is exactly
rounded, so the device result must equal the oracle's (and the reference
routine's, where oracle/_ref is built) bit for bit, with the same iteration
count.  The factor path with scaling = 4 must keep the inertia of the
unscaled matrix and reach the 1e-14 scaled backward error.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import scaling as osc
from spral_amd import devmem, matrices, ssids
from helpers import analyse, rhs
from test_scaling import random_sym

pytestmark = pytest.mark.gpu

BWD_TOL = 1e-14


def stencil_random(g, seed):
    """7-point pattern on g^3 with N(0,1) values of widely varying size."""
    m = matrices.laplace3d(g, diag=6.0)
    rng = np.random.default_rng(seed)
    val = rng.standard_normal(m.val.size) * 10.0 ** rng.uniform(-4, 4, m.val.size)
    return m.n, m.ptr, m.row, val


@pytest.mark.parametrize("seed", range(24))
def test_device_matches_oracle_bitwise(lib, seed):
    n = seed + 1 if seed < 20 else 500 + 311 * seed
    ptr, row, val = random_sym(n, 7000 + seed, diag=(seed % 4 != 1))
    print(f"equilib seed={seed} n={n} nz={len(val)}", flush=True)
    for maxit in (1, 4, 10):
        s_d, it_d = ssids.equilib_scale_sym(n, ptr, row, val, max_iterations=maxit)
        s_o, it_o = osc.equilib_scale_sym(n, ptr, row, val, max_iterations=maxit)
        assert it_d == it_o, (maxit, it_d, it_o)
        assert np.array_equal(s_d, s_o), (maxit, np.max(np.abs(s_d - s_o)))
        if osc.ref_available():
            s_r, it_r, _ = osc.ref_equilib_scale_sym(n, ptr, row, val, max_iterations=maxit)
            assert it_d == it_r and np.array_equal(s_d, s_r)


def test_device_entry_and_large_stencil(lib):
    n, ptr, row, val = stencil_random(40, 11)  # n = 64,000
    print(f"equilib stencil n={n} nz={len(val)}", flush=True)
    s_h, it_h = ssids.equilib_scale_sym(n, ptr, row, val)
    s_o, it_o = osc.equilib_scale_sym(n, ptr, row, val)
    assert it_h == it_o and np.array_equal(s_h, s_o)
    # the same through the device-array entry
    d_ptr = devmem.DeviceArray.from_host(np.asarray(ptr, np.int64))
    d_row = devmem.DeviceArray.from_host(np.asarray(row, np.int32))
    d_val = devmem.DeviceArray.from_host(val)
    d_s = devmem.DeviceArray(n)
    its = C.c_int(-1)
    rc = lib.spral_ssids_hip_equilib_scale_sym_device(n, d_ptr.ptr, d_row.ptr, d_val.ptr,
                                                      d_s.ptr, 10, 1e-8, C.byref(its))
    assert rc == 0 and its.value == it_h
    assert np.array_equal(d_s.download(), s_h)


def test_invalid_patterns_rejected(lib):
    s = np.zeros(3)
    its = C.c_int(0)
    val = np.ones(3)
    bad = [(np.array([0, 1, 2, 3], np.int64), np.array([1, 2, 3], np.int32)),   # ptr[0] != 1
           (np.array([1, 3, 2, 4], np.int64), np.array([1, 2, 3], np.int32)),   # decreasing
           (np.array([1, 2, 3, 4], np.int64), np.array([1, 4, 3], np.int32))]   # row > n
    for ptr, row in bad:
        rc = lib.spral_ssids_hip_equilib_scale_sym(3, ptr.ctypes.data, row.ctypes.data,
                                                   val.ctypes.data, s.ctypes.data, 10, 1e-8,
                                                   C.byref(its))
        assert rc == -1


def badly_scaled_kkt(seed=5):
    """The config-1 KKT system under a random diagonal scaling D K D with
    D spread over 10^-3 .. 10^3: inertia is unchanged by the congruence."""
    m = matrices.kkt_laplace()
    rng = np.random.default_rng(seed)
    d = 10.0 ** rng.uniform(-3, 3, m.n)
    low = sp.tril(m.full, format="csc")
    low = sp.diags(d) @ low @ sp.diags(d)
    low = sp.csc_matrix(low)
    low.sort_indices()
    full = low + sp.tril(low, k=-1).T
    out = matrices.LowerCSC.__new__(matrices.LowerCSC)
    out.n = m.n
    out.ptr = (low.indptr + 1).astype(np.int64)
    out.row = (low.indices + 1).astype(np.int32)
    out.val = low.data.copy()
    out.full = sp.csc_matrix(full)
    return m, out


def test_factor_with_equilibration(lib):
    m0, m = badly_scaled_kkt()
    ak, inf = analyse(m, use_gpu=True)
    assert inf.flag >= 0
    opts = ssids.SSIDSOptions(scaling=4)
    scale = np.zeros(m.n)
    fk, finf = ssids.ssids_factor(False, m.val, ak, opts, scale=scale)
    assert finf.flag >= 0, finf
    # the scaling handed back is the equilibration of the values
    s_o, _ = osc.equilib_scale_sym(m.n, m.ptr, m.row, m.val)
    assert np.array_equal(scale, s_o)
    # inertia of D K D is that of K (config 1: 1024 negative eigenvalues)
    assert finf.num_neg == 1024 and finf.matrix_rank == m.n
    xt, b = rhs(m, 2, 21)
    x = np.array(b, copy=True)
    inf2, bwd, steps = ssids.ssids_solve_refine(x, ak, fk, opts)
    print(f"  scaled kkt: delays={finf.num_delay} two={finf.num_two} refine steps={steps} "
          f"bwd={np.max(bwd):.2e}", flush=True)
    assert inf2.flag >= 0
    assert matrices.bwd_error(m.full, x, b).max() <= BWD_TOL
    ssids.ssids_free(ak, fk)
