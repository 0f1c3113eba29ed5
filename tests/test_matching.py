"""Maximum-product matching scaling (spral_hungarian_scale_sym, host code in
the library) against the reference's hungarian_scale_sym
(src/scaling.f90:134-170, compiled where it lies into
oracle/_ref/libref_scaling.so): the matching, its cardinality, the flag and
the scaling bit for bit; and, without the reference, the properties that
define the scaling: every matched entry of S A S has modulus 1 and no entry
exceeds 1 (to rounding), and a diagonal matrix gets s_i = |a_ii|^{-1/2}."""
import ctypes as C

import numpy as np
import pytest

from oracle import scaling as osc
from spral_amd import matrices
from test_scaling import random_sym


def hungarian(lib, n, ptr, row, val, scale_if_singular=False):
    ptr = np.ascontiguousarray(ptr, np.int64)
    row = np.ascontiguousarray(row, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    s = np.zeros(max(n, 1))
    mt = np.zeros(max(n, 1), np.int32)
    nm = C.c_int(0)
    fl = lib.spral_hungarian_scale_sym(n, ptr.ctypes.data, row.ctypes.data, val.ctypes.data,
                                       s.ctypes.data, mt.ctypes.data,
                                       1 if scale_if_singular else 0, C.byref(nm))
    return s[:n], mt[:n], nm.value, fl


def cases():
    for seed in range(40):
        n = seed + 1 if seed < 25 else 60 + 41 * seed
        yield f"random{seed}", n, random_sym(n, 9100 + seed, diag=(seed % 3 != 1))
    # KKT systems: zero (2,2) block, matching pairs constraints with H
    for g, mc in ((4, 20), (6, 70)):
        m = matrices.kkt_laplace(grid=g, m_con=mc, seed=77 + g)
        yield f"kkt{g}", m.n, (m.ptr, m.row, m.val)
    # integer values: many equal costs, so the tie rules decide the matching
    for seed in range(6):
        n = 30 + 10 * seed
        ptr, row, val = random_sym(n, 500 + seed, diag=(seed % 2 == 0))
        val = np.round(np.abs(val) * 0 + np.random.default_rng(seed).integers(1, 4, len(val)))
        yield f"ties{seed}", n, (ptr, row, val.astype(np.float64))


CASES = list(cases())


@pytest.mark.skipif(not osc.ref_available(), reason="reference scaling (oracle/_ref) not built")
@pytest.mark.parametrize("name,n,mat", CASES, ids=[c[0] for c in CASES])
def test_matches_reference_bitwise(lib, name, n, mat):
    ptr, row, val = mat
    for sis in (False, True):
        s, mt, nm, fl = hungarian(lib, n, ptr, row, val, sis)
        rs, rmt, rnm, rfl = osc.ref_hungarian_scale_sym(n, ptr, row, val, sis)
        assert (nm, fl) == (rnm, rfl), (name, nm, rnm, fl, rfl)
        assert np.array_equal(mt, rmt), name
        assert np.array_equal(s, rs), (name, np.max(np.abs(s - rs) / rs))


@pytest.mark.parametrize("name,n,mat", CASES[:30], ids=[c[0] for c in CASES[:30]])
def test_scaled_matrix_property(lib, name, n, mat):
    ptr, row, val = mat
    s, mt, nm, fl = hungarian(lib, n, ptr, row, val, True)
    if nm < n:
        assert fl == 1
        return
    assert fl == 0
    col = np.repeat(np.arange(n), np.diff(ptr))
    r = np.asarray(row) - 1
    sv = np.abs(s[col] * np.asarray(val) * s[r])
    # every entry <= 1 and every matched entry == 1, to rounding
    assert sv.max() <= 1 + 1e-12
    full = {}
    for c_, r_, v in zip(col, r, sv):
        full[(r_, c_)] = v
        full[(c_, r_)] = v
    for i in range(n):
        j = mt[i] - 1
        assert abs(full[(i, j)] - 1.0) <= 1e-12, (i, j, full[(i, j)])


def test_diagonal_closed_form(lib):
    n = 40
    d = np.random.default_rng(2).standard_normal(n) * 10.0 ** np.linspace(-6, 6, n)
    ptr = np.arange(1, n + 2, dtype=np.int64)
    row = np.arange(1, n + 1, dtype=np.int32)
    s, mt, nm, fl = hungarian(lib, n, ptr, row, d)
    assert fl == 0 and nm == n and np.array_equal(mt, np.arange(1, n + 1))
    np.testing.assert_allclose(s * s * np.abs(d), 1.0, rtol=1e-13)


def test_structurally_singular_flags(lib):
    # the second column is empty: no perfect matching
    ptr = np.array([1, 2, 2, 3], np.int64)
    row = np.array([1, 3], np.int32)
    val = np.array([2.0, 8.0])
    s, mt, nm, fl = hungarian(lib, 3, ptr, row, val, False)
    assert nm == 2 and fl == -2
    s, mt, nm, fl = hungarian(lib, 3, ptr, row, val, True)
    assert nm == 2 and fl == 1 and mt[1] == 0


# ---- matching-based elimination order (options.ordering = 2) ----------------

def match_order(lib, m):
    order = np.zeros(m.n, np.int32)
    scale = np.zeros(m.n)
    f = lib.spral_match_order(m.n, np.asarray(m.ptr, np.int64), np.asarray(m.row, np.int32),
                              np.asarray(m.val, np.float64), order, scale)
    return f, order, scale


@pytest.mark.parametrize("case", ["kkt_laplace", "kkt_random", "kkt27", "indef"])
def test_match_order_pairs_and_scaling(lib, case):
    """The order is a permutation; every matched pair of the 2-cycle split
    holds consecutive positions; the matching scaling puts every matched
    entry of S A S at modulus 1 and none above 1."""
    m = {"kkt_laplace": lambda: matrices.kkt_laplace(grid=6, m_con=60),
         "kkt_random": lambda: matrices.kkt_random(900, seed=11),
         "kkt27": lambda: matrices.kkt_stencil27(8, m_con=88),
         "indef": lambda: matrices.random_indef(300, density=0.03, seed=2)}[case]()
    f, order, scale = match_order(lib, m)
    assert f == 0
    assert np.array_equal(np.sort(order), np.arange(1, m.n + 1))
    # the matching itself (pinned to the reference above) on the same matrix
    s1, mt, nm, fl = hungarian(lib, m.n, m.ptr, m.row, m.val)
    assert nm == m.n
    partner = mt - 1
    # 2-cycles of the matching are pairs: consecutive positions
    for i in range(m.n):
        j = partner[i]
        if j != i and partner[j] == i:
            assert abs(int(order[i]) - int(order[j])) == 1, (i, j, order[i], order[j])
    col = np.repeat(np.arange(m.n), np.diff(m.ptr))
    r = np.asarray(m.row) - 1
    sv = np.abs(scale[col] * m.val * scale[r])
    assert sv.max() <= 1 + 1e-12
    full = {}
    for c_, r_, v in zip(col, r, sv):
        full[(r_, c_)] = v
        full[(c_, r_)] = v
    for i in range(m.n):
        assert abs(full[(i, partner[i])] - 1.0) <= 1e-12


def test_match_order_removes_kkt_delays(lib, oracle_lib):
    """On a saddle-point system the matched pairs (a constraint row with an H
    variable) meet in one front: the oracle's TPP factorization delays
    nothing, with the closed-form inertia, where the AMD order delays."""
    from oracle.oracle_py import OracleFactor
    from spral_amd import ssids
    m = matrices.kkt_stencil27(12, m_con=297)
    res = {}
    for ordering, sc in ((4, 0), (2, 3)):
        o = ssids.SSIDSOptions(ordering=ordering, use_gpu=False, scaling=sc)
        ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o, val=m.val)
        assert inf.flag >= 0
        of = OracleFactor(ak, m.val, scaling=(ak.scaling if sc == 3 else None))
        res[ordering] = of.stats
        assert of.stats["num_neg"] == 297 and of.stats["matrix_rank"] == m.n
    assert res[4]["num_delay"] > 100 and res[2]["num_delay"] == 0


def test_match_order_needs_values(lib):
    from spral_amd import ssids
    m = matrices.kkt_laplace(grid=4, m_con=10)
    o = ssids.SSIDSOptions(ordering=2, use_gpu=False)
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o)
    assert inf.flag == ssids.SSIDS_ERROR_VAL


# ---- auction scaling (options.scaling = 2) -----------------------------------

def auction(lib, n, ptr, row, val, max_iterations=30000, max_unchanged=(10, 100, 100),
            min_proportion=(0.90, 0.0, 0.0), eps_initial=0.01):
    ptr = np.ascontiguousarray(ptr, np.int64)
    row = np.ascontiguousarray(row, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    s = np.zeros(max(n, 1))
    mt = np.zeros(max(n, 1), np.int32)
    mu = np.ascontiguousarray(max_unchanged, np.int32)
    mp = np.ascontiguousarray(min_proportion, np.float32)
    nm, it, un = C.c_int(0), C.c_int(0), C.c_int(0)
    fl = lib.spral_auction_scale_sym(n, ptr.ctypes.data, row.ctypes.data, val.ctypes.data,
                                     s.ctypes.data, mt.ctypes.data, max_iterations,
                                     mu.ctypes.data, mp.ctypes.data, float(np.float32(eps_initial)),
                                     C.byref(nm), C.byref(it), C.byref(un))
    return s[:n], mt[:n], nm.value, it.value, fl


@pytest.mark.skipif(not osc.ref_available(), reason="reference scaling (oracle/_ref) not built")
@pytest.mark.parametrize("name,n,mat", CASES, ids=[c[0] for c in CASES])
def test_auction_matches_reference_bitwise(lib, name, n, mat):
    ptr, row, val = mat
    s, mt, nm, it, fl = auction(lib, n, ptr, row, val)
    rs, rmt, rnm, rit, rfl = osc.ref_auction_scale_sym(n, ptr, row, val)
    assert (nm, it, fl) == (rnm, rit, rfl), (name, nm, rnm, it, rit)
    assert np.array_equal(mt, rmt), name
    assert np.array_equal(s, rs), (name, np.max(np.abs(s - rs) / rs))


def test_auction_scaled_entries_bounded(lib):
    """The auction's approximate duals keep every scaled entry of S A S near
    or below 1 (within the eps the last bids paid) on a KKT system."""
    m = matrices.kkt_laplace(grid=6, m_con=60)
    s, mt, nm, it, fl = auction(lib, m.n, m.ptr, m.row, m.val)
    assert fl == 0 and nm > 0.9 * m.n
    col = np.repeat(np.arange(m.n), np.diff(m.ptr))
    sv = np.abs(s[col] * m.val * s[np.asarray(m.row) - 1])
    assert sv.max() < 10.0
