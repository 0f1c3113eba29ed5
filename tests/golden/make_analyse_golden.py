"""Write golden analyse fixtures from the reference's own basic_analyse.

Runs src/core_analyse.f90 (compiled where it lies into oracle/_ref by
oracle/Makefile) on small seeded inputs and stores inputs and outputs as
.npz data.  Regenerate with:  python tests/golden/make_analyse_golden.py
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from spral_amd import matrices  # noqa: E402

REF = ROOT / "oracle" / "_ref" / "libref_analyse.so"


def ref_analyse(m, order, nemin=32):
    lib = C.CDLL(str(REF))
    # full structure from the lower-triangle pattern itself (explicit zeros kept)
    low = sp.csc_matrix((np.ones(m.row.size), m.row - 1, m.ptr - 1), shape=(m.n, m.n))
    f = sp.csc_matrix(low + sp.tril(low, k=-1).T)
    f.sort_indices()
    p = f.indptr.astype(np.int64) + 1
    r = f.indices.astype(np.int32) + 1
    nn, rl, nf, fl, info = C.c_int(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int()
    o = np.ascontiguousarray(order, np.int32).copy()
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    lib.ref_analyse_run(C.c_int(m.n), vp(p), vp(r), vp(o), C.c_int(nemin), C.byref(nn),
                        C.byref(rl), C.byref(nf), C.byref(fl), C.byref(info))
    perm = np.zeros(m.n, np.int32)
    sptr = np.zeros(nn.value + 1, np.int32)
    spar = np.zeros(nn.value, np.int32)
    rptr = np.zeros(nn.value + 1, np.int64)
    rlist = np.zeros(rl.value, np.int32)
    lib.ref_analyse_fetch(C.c_int(m.n), vp(perm), vp(sptr), vp(spar), vp(rptr), vp(rlist))
    return dict(order=perm, sptr=sptr, sparent=spar, rptr=rptr, rlist=rlist,
                nfact=np.int64(nf.value), nflops=np.int64(fl.value), info=np.int32(info.value))


def cases():
    out = []
    for seed in range(3):
        m = matrices.random_indef(200 + 50 * seed, density=0.03, seed=seed)
        order = np.random.default_rng(100 + seed).permutation(m.n).astype(np.int32) + 1
        out.append((f"random{seed}", m, order, 32))
    m = matrices.random_indef(240, density=0.02, seed=7)
    order = np.arange(1, m.n + 1, dtype=np.int32)
    out.append(("random_nemin8", m, order, 8))
    m = matrices.laplace3d(8, diag=5.7)
    order = np.zeros(m.n, np.int32)
    from spral_amd import _lib
    _lib.lib().spral_ssids_hip_order_grid_nd(8, 8, 8, order)
    out.append(("laplace8_nd", m, order, 32))
    return out


def main():
    for name, m, order, nemin in cases():
        res = ref_analyse(m, order, nemin)
        np.savez_compressed(ROOT / "tests" / "golden" / f"analyse_{name}.npz",
                            ptr=m.ptr, row=m.row, order_in=order, nemin=np.int32(nemin), **res)
        print(name, "nnodes", len(res["sptr"]) - 1)


if __name__ == "__main__":
    main()
