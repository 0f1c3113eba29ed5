"""Shared helpers for the tests (analyse wrappers, reference-style checks)."""
import numpy as np

from spral_amd import matrices, ssids


def analyse(m, grid=None, use_gpu=False, **kw):
    if grid is not None:
        opts = ssids.SSIDSOptions(ordering=3, use_gpu=use_gpu, **kw)
        return ssids.ssids_analyse(False, m.n, m.ptr, m.row, opts, grid=grid)
    opts = ssids.SSIDSOptions(use_gpu=use_gpu, **kw)
    return ssids.ssids_analyse(False, m.n, m.ptr, m.row, opts)


def rhs(m, nrhs=1, seed=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    xt = rng.uniform(-1, 1, (m.n, nrhs))
    return xt, m.full @ xt


def brute_inertia(m):
    ev = np.linalg.eigvalsh(m.full.toarray())
    return int((ev < 0).sum()), ev


def doc_example():
    """5x5 example of examples/Fortran/ssids.f90:30-42 (solution 1..5,
    docs/Fortran/ssids.rst)."""
    import scipy.sparse as sp
    ptr = np.array([1, 3, 6, 8, 9, 10], np.int64)
    row = np.array([1, 2, 2, 3, 5, 3, 4, 4, 5], np.int32)
    val = np.array([2.0, 1.0, 4.0, 1.0, 1.0, 3.0, 2.0, -1.0, 2.0])
    low = sp.csc_matrix((val, row - 1, ptr - 1), shape=(5, 5))
    full = low + sp.tril(low, k=-1).T
    m = matrices.LowerCSC.__new__(matrices.LowerCSC)
    m.n, m.ptr, m.row, m.val, m.full = 5, ptr, row, val, sp.csc_matrix(full)
    b = np.array([4.0, 17.0, 19.0, 2.0, 12.0])
    return m, b
