"""Matching-based scaling and order on the GPU factorization.  With the
matching order (options.ordering = 2) and its scaling (options.scaling = 3)
a saddle-point system's constraint rows are eliminated next to their
matched partners: the device factorization delays nothing, keeps the
closed-form inertia (num_neg = constraint rows, H positive definite and A of
full row rank) and solves to the 1e-14 backward error.  The Hungarian
(options.scaling = 1) and auction (2) scalings give the same inertia and
accuracy on any order."""
import numpy as np
import pytest

from spral_amd import matrices, ssids
from helpers import rhs

pytestmark = pytest.mark.gpu

BWD_TOL = 1e-14


def _run(m, ordering, scaling):
    o = ssids.SSIDSOptions(ordering=ordering, scaling=scaling)
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o, val=m.val)
    assert inf.flag >= 0, inf
    fk, finf = ssids.ssids_factor(False, m.val, ak, o)
    assert finf.flag >= 0, finf
    _, b = rhs(m, 2, 9)
    x = b.copy()
    sinf, bwd, its = ssids.ssids_solve_refine(x, ak, fk, o, maxit=5, tol=BWD_TOL)
    assert sinf.flag >= 0 and matrices.bwd_error(m.full, x, b).max() <= BWD_TOL
    res = (finf.num_neg, finf.num_delay, finf.matrix_rank, its)
    ssids.ssids_free(ak, fk)
    return res


@pytest.mark.parametrize("case", ["kkt27_16", "kkt_laplace_12", "kkt_random"])
def test_matching_order_no_delays(lib, case):
    if case == "kkt27_16":
        m, mc = matrices.kkt_stencil27(16, m_con=704), 704
    elif case == "kkt_laplace_12":
        m, mc = matrices.kkt_laplace(grid=12, shift=0.5, m_con=400), 400
    else:
        m, mc = matrices.kkt_random(3000, seed=12), int(0.2 * 3000)
    print(f"matching order case={case} n={m.n}", flush=True)
    neg, delays, rank, its = _run(m, 2, 3)
    assert neg == mc and rank == m.n
    assert delays == 0
    neg4, delays4, rank4, _ = _run(m, 4, 0)
    assert neg4 == mc and rank4 == m.n
    assert delays4 > 0


@pytest.mark.parametrize("scaling", [1, 2])
def test_matching_scalings_factor(lib, scaling):
    """Hungarian (1) and auction (2) scalings: same inertia and rank as the
    unscaled factorization, 1e-14 after refinement."""
    m = matrices.random_indef(500, density=0.02, seed=17)
    print(f"matching scaling {scaling} n={m.n}", flush=True)
    neg1, _, rank1, _ = _run(m, 4, scaling)
    neg0, _, rank0, _ = _run(m, 4, 0)
    assert neg1 == neg0 and rank1 == rank0
