"""Host analyse: identical integer-for-integer to the reference analyse.

Golden fixtures in tests/golden/ were written by the reference's own
basic_analyse (src/core_analyse.f90, via oracle/_ref); when oracle/_ref is
built the comparison also runs live on larger inputs (config 1 and the 64^3
grid).  build_map is checked by a property: scattering A through nlist
reproduces P A P^T column by column.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from spral_amd import matrices, ssids

GOLD = Path(__file__).resolve().parent / "golden"
REF = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "libref_analyse.so"


def mine(ptr, row, order, nemin):
    n = len(ptr) - 1
    opts = ssids.SSIDSOptions(ordering=0, nemin=int(nemin), use_gpu=False)
    ak, inf = ssids.ssids_analyse(False, n, ptr, row, opts, order=order)
    assert inf.flag >= 0
    return ak


@pytest.mark.parametrize("path", sorted(GOLD.glob("analyse_*.npz")), ids=lambda p: p.stem)
def test_analyse_matches_golden(lib, path):
    g = np.load(path)
    ak = mine(g["ptr"], g["row"], g["order_in"], g["nemin"])
    assert np.array_equal(ak.sptr, g["sptr"])
    assert np.array_equal(ak.sparent, g["sparent"])
    assert np.array_equal(ak.rptr, g["rptr"])
    assert np.array_equal(ak.rlist, g["rlist"])
    assert ak.num_factor == int(g["nfact"]) and ak.num_flops == int(g["nflops"])
    # final pivot order (all variables used in these fixtures)
    assert np.array_equal(ak.order, g["order"])


@pytest.mark.skipif(not REF.exists(), reason="reference analyse shim (oracle/_ref) not built")
@pytest.mark.parametrize("case", ["kkt5120_md", "laplace64_nd", "random_nemin4"])
def test_analyse_matches_reference_live(lib, case):
    import sys
    sys.path.insert(0, str(GOLD))
    from make_analyse_golden import ref_analyse
    if case == "kkt5120_md":
        m = matrices.kkt_laplace()
        order = np.zeros(m.n, np.int32)
        lib.spral_ssids_hip_order_min_degree(m.n, m.ptr, m.row, order)
        nemin = 32
    elif case == "laplace64_nd":
        m = matrices.laplace3d(64, diag=5.7)
        order = np.zeros(m.n, np.int32)
        lib.spral_ssids_hip_order_grid_nd(64, 64, 64, order)
        nemin = 32
    else:
        m = matrices.random_indef(500, density=0.01, seed=11)
        order = np.random.default_rng(5).permutation(m.n).astype(np.int32) + 1
        nemin = 4
    ref = ref_analyse(m, order, nemin)
    ak = mine(m.ptr, m.row, order, nemin)
    for k in ("sptr", "sparent", "rptr", "rlist", "order"):
        assert np.array_equal(getattr(ak, k), ref[k]), k
    assert ak.num_flops == int(ref["nflops"])


def test_build_map_scatter_reproduces_matrix(lib):
    m = matrices.random_indef(150, density=0.05, seed=3)
    opts = ssids.SSIDSOptions(use_gpu=False)
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, opts)
    # P A P^T lower entries keyed by (pivot row, pivot col)
    pos = ak.order - 1
    cols = np.repeat(np.arange(m.n), np.diff(m.ptr))
    expect = {}
    for r, c, v in zip(m.row - 1, cols, m.val):
        pr, pc = pos[r], pos[c]
        expect[(max(pr, pc), min(pr, pc))] = v
    got = {}
    for s in range(ak.nnodes):
        rows = ak.rlist[ak.rptr[s] - 1:ak.rptr[s + 1] - 1] - 1
        blkm = len(rows)
        for k in range(ak.nptr[s] - 1, ak.nptr[s + 1] - 1):
            src, dst = ak.nlist[2 * k] - 1, ak.nlist[2 * k + 1] - 1
            c, r = divmod(dst, blkm)
            key = (rows[r], ak.sptr[s] - 1 + c)
            assert key not in got
            got[key] = m.val[src]
    assert got == expect


def test_orderings_are_permutations(lib):
    o = np.zeros(10 * 9 * 8, np.int32)
    lib.spral_ssids_hip_order_grid_nd(10, 9, 8, o)
    assert np.array_equal(np.sort(o), np.arange(1, o.size + 1))
    m = matrices.random_indef(100, density=0.05, seed=1)
    o = np.zeros(m.n, np.int32)
    lib.spral_ssids_hip_order_min_degree(m.n, m.ptr, m.row, o)
    assert np.array_equal(np.sort(o), np.arange(1, m.n + 1))


def test_bad_order_rejected(lib):
    m = matrices.random_indef(20, density=0.2, seed=1)
    order = np.ones(m.n, np.int32)
    opts = ssids.SSIDSOptions(ordering=0, use_gpu=False)
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, opts, order=order)
    assert inf.flag == ssids.SSIDS_ERROR_ORDER
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, opts)
    assert inf.flag == ssids.SSIDS_ERROR_ORDER
