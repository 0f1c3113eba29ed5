"""The factorization on approximate-minimum-degree orders (ordering = 4).

Inertia does not depend on the ordering (Sylvester's law of inertia), so the
AMD-ordered factorization must give the num_neg of the dense eigenvalue count
(config-1 KKT, 1024) and of the nested-dissection order (the 32^3 stencil),
and its solve must reach the scaled backward error 1e-14 after FP64
refinement.
"""
import numpy as np
import pytest

from spral_amd import matrices, ssids
from helpers import rhs

pytestmark = pytest.mark.gpu

BWD_TOL = 1e-14


def factor(m, ordering, grid=None):
    o = ssids.SSIDSOptions(ordering=ordering)
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o, grid=grid)
    assert inf.flag >= 0, inf
    fk, finf = ssids.ssids_factor(False, m.val, ak, o)
    assert finf.flag >= 0, finf
    return ak, fk, finf, o


def refined_bwd(m, ak, fk, o, seed):
    xt, b = rhs(m, 2, seed)
    x = np.array(b, copy=True)
    inf, bwd, its = ssids.ssids_solve_refine(x, ak, fk, o, maxit=5, tol=BWD_TOL)
    assert inf.flag >= 0
    return matrices.bwd_error(m.full, x, b).max()


def test_kkt_config1_amd(lib):
    m = matrices.kkt_laplace()
    ak, fk, finf, o = factor(m, 4)
    print(f"  kkt amd: num_factor {ak.num_factor} delays {finf.num_delay}", flush=True)
    assert finf.num_neg == 1024 and finf.matrix_rank == m.n
    assert refined_bwd(m, ak, fk, o, 1) <= BWD_TOL
    ssids.ssids_free(ak, fk)


def test_stencil_amd_vs_nested_dissection(lib):
    g = 32
    m = matrices.laplace3d(g, diag=5.7)
    ak1, fk1, f1, o1 = factor(m, 4)
    ak2, fk2, f2, o2 = factor(m, 3, grid=(g, g, g))
    print(f"  32^3: amd flops {ak1.num_flops:.3e} nd flops {ak2.num_flops:.3e}", flush=True)
    assert f1.num_neg == f2.num_neg and f1.matrix_rank == f2.matrix_rank == m.n
    # closed form: eigenvalues 5.7 - 2 sum cos(pi i / (g+1))
    c = 2 * np.cos(np.pi * np.arange(1, g + 1) / (g + 1))
    ev = 5.7 - (c[:, None, None] + c[None, :, None] + c[None, None, :])
    assert f1.num_neg == int((ev < 0).sum())
    assert refined_bwd(m, ak1, fk1, o1, 2) <= BWD_TOL
    ssids.ssids_free(ak1, fk1)
    ssids.ssids_free(ak2, fk2)
