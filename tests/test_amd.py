"""Approximate minimum degree ordering (order_amd; Amestoy, Davis and Duff,
SIAM J. Matrix Anal. Appl. 17(4), 1996), the METIS stand-in for large
general patterns.

Any permutation gives a correct factorization, so the checks are: the output
is a permutation, it is deterministic, and its fill is that of a minimum
degree ordering. Closed forms (supernode amalgamation off, nemin = 1, so
num_factor counts the entries of L exactly): a path has no fill (2n - 1); a
star eliminates its leaves before the hub (2n - 1; the hub first would fill
the whole triangle); a dense block fills n(n+1)/2 in any order. On random
patterns, the config-1 KKT system and stencils the fill stays within 10% of
the exact minimum degree ordering's.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from spral_amd import _lib, matrices, ssids


def lower_csc(full):
    low = sp.csc_matrix(sp.tril(full))
    low.sort_indices()
    m = matrices.LowerCSC.__new__(matrices.LowerCSC)
    m.n = full.shape[0]
    m.ptr = (low.indptr + 1).astype(np.int64)
    m.row = (low.indices + 1).astype(np.int32)
    m.val = low.data.copy()
    m.full = sp.csc_matrix(full)
    return m


def amd_order(m):
    o = np.zeros(m.n, np.int32)
    assert _lib.lib().spral_ssids_hip_order_amd(m.n, m.ptr, m.row, o) == 0
    return o


def fill(m, ordering, nemin=32, order=None):
    o = ssids.SSIDSOptions(ordering=ordering, use_gpu=False, nemin=nemin)
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o, order=order)
    assert inf.flag >= 0
    return ak.num_factor


def is_permutation(o, n):
    return np.array_equal(np.sort(o), np.arange(1, n + 1))


@pytest.mark.parametrize("seed", range(12))
def test_permutation_and_determinism(lib, seed):
    n = [1, 2, 5, 17, 64, 150, 300, 40, 90, 250, 500, 333][seed]
    m = matrices.random_indef(n, density=min(0.5, 6.0 / n), seed=seed)
    o1, o2 = amd_order(m), amd_order(m)
    assert is_permutation(o1, n)
    assert np.array_equal(o1, o2)


def test_closed_forms(lib):
    n = 200
    path = sp.diags([np.full(n - 1, -1.0), np.full(n, 4.0), np.full(n - 1, -1.0)], [-1, 0, 1])
    assert fill(lower_csc(path), 4, nemin=1) == 2 * n - 1
    star = sp.lil_matrix((n, n))
    star.setdiag(4.0)
    star[0, 1:] = 1.0
    star[1:, 0] = 1.0
    o = amd_order(lower_csc(star.tocsc()))
    assert o[0] >= n - 1  # the hub goes last (or next to last: no fill either way)
    assert fill(lower_csc(star.tocsc()), 4, nemin=1) == 2 * n - 1
    k = 40
    dense = sp.csc_matrix(np.ones((k, k)) + k * np.eye(k))
    assert fill(lower_csc(dense), 4, nemin=1) == k * (k + 1) // 2


def test_empty_and_diagonal(lib):
    m = lower_csc(sp.identity(30, format="csc") * 2.0)
    o = amd_order(m)
    assert is_permutation(o, 30)
    assert fill(m, 4, nemin=1) == 30


@pytest.mark.parametrize("case", ["random", "kkt", "stencil"])
def test_fill_near_exact_minimum_degree(lib, case):
    if case == "random":
        m = matrices.random_indef(500, density=0.01, seed=3)
    elif case == "kkt":
        m = matrices.kkt_laplace()
    else:
        m = matrices.laplace3d(16, diag=6.0)
    f_amd, f_md = fill(m, 4, nemin=1), fill(m, 1, nemin=1)
    print(f"  {case}: amd {f_amd} exact md {f_md} ({f_amd / f_md:.3f})")
    assert f_amd <= 1.10 * f_md
    # and well below the natural order's
    f_nat = fill(m, 0, nemin=1, order=np.arange(1, m.n + 1, dtype=np.int32))
    assert f_amd < 0.8 * f_nat


def test_large_pattern_is_fast(lib):
    import time
    m = matrices.laplace3d(48, diag=6.0)  # n = 110,592
    t = time.perf_counter()
    o = amd_order(m)
    dt = time.perf_counter() - t
    assert is_permutation(o, m.n)
    assert dt < 10.0, dt
