"""GPU positive-definite path (ssids_factor with posdef = .true.).

The reference factors A = L L^T with pivots in order (factor_node_posdef,
src/ssids/cpu/factor.hxx:139-171) and reports the diagonal of L through
ssids_enquire_posdef (src/ssids/cpu/NumericSubtree.hxx:424-433).  The
Cholesky factor of the permuted matrix is unique, so every check here is
against numpy's dense Cholesky of P A P^T (P = the analyse order): the
diagonal of L, the forward half L y = b, the backward half L^T x = y, and
the full solve.  Tolerances are written in each test.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from spral_amd import matrices, ssids
from helpers import analyse, rhs

pytestmark = pytest.mark.gpu

BWD_TOL = 1e-14
FACT_TOL = 1e-12  # relative, entries of the Cholesky factor


def posdef_factor(m, grid=None):
    ak, inf = analyse(m, grid=grid, use_gpu=True)
    assert inf.flag >= 0, inf
    opts = ssids.SSIDSOptions() if grid is None else ssids.SSIDSOptions(ordering=3)
    fk, finf = ssids.ssids_factor(True, m.val, ak, opts)
    return ak, fk, finf, opts


def dense_chol(m, ak):
    p = np.asarray(ak.invp)  # elimination position i holds variable p[i]
    A = m.full.toarray()[np.ix_(p, p)]
    return p, np.linalg.cholesky(A)


@pytest.mark.parametrize("case", ["laplace", "random"])
def test_posdef_factor_matches_dense_cholesky(lib, case):
    if case == "laplace":
        m = matrices.laplace3d(9, diag=6.0)
        ak, fk, finf, opts = posdef_factor(m, grid=(9, 9, 9))
    else:
        m = matrices.random_posdef(400, density=0.03, seed=7)
        ak, fk, finf, opts = posdef_factor(m)
    assert finf.flag == 0, finf
    assert finf.num_neg == 0 and finf.num_two == 0 and finf.num_delay == 0
    assert finf.matrix_rank == m.n
    p, Lc = dense_chol(m, ak)
    d, inf = ssids.ssids_enquire_posdef(ak, fk)
    assert inf.flag == 0
    np.testing.assert_allclose(d, np.diag(Lc), rtol=FACT_TOL, atol=0)
    xt, b = rhs(m, 3, 5)
    # forward half: L y = P b, y returned in the original indexing
    y = np.array(b, copy=True)
    assert ssids.ssids_solve(y, ak, fk, opts, job=ssids.SSIDS_SOLVE_JOB_FWD).flag == 0
    y_ref = sla.solve_triangular(Lc, b[p], lower=True)
    np.testing.assert_allclose(y[p], y_ref, rtol=0, atol=FACT_TOL * np.abs(y_ref).max())
    # backward half on it: L^T x = y gives the solution
    x = np.array(y, copy=True)
    assert ssids.ssids_solve(x, ak, fk, opts, job=ssids.SSIDS_SOLVE_JOB_BWD).flag == 0
    assert matrices.bwd_error(m.full, x, b).max() <= BWD_TOL
    # the full solve
    x2 = np.array(b, copy=True)
    assert ssids.ssids_solve(x2, ak, fk, opts).flag == 0
    assert matrices.bwd_error(m.full, x2, b).max() <= BWD_TOL
    np.testing.assert_allclose(x2, xt, rtol=0, atol=1e-10)
    ssids.ssids_free(ak, fk)


def test_posdef_jobs_and_enquiries(lib):
    m = matrices.random_posdef(200, density=0.05, seed=3)
    ak, fk, finf, opts = posdef_factor(m)
    assert finf.flag == 0
    b = np.ones(m.n)
    for job in (ssids.SSIDS_SOLVE_JOB_DIAG, ssids.SSIDS_SOLVE_JOB_DIAG_BWD):
        inf = ssids.ssids_solve(b.copy(), ak, fk, opts, job=job)
        assert inf.flag == ssids.SSIDS_ERROR_JOB_OOR
    _, _, inf = ssids.ssids_enquire_indef(ak, fk)
    assert inf.flag == ssids.SSIDS_ERROR_NOT_LDLT
    assert ssids.ssids_alter(np.zeros((2, m.n)), ak, fk).flag == ssids.SSIDS_ERROR_NOT_LDLT
    # refinement works on the L L^T factors as on L D L^T
    xt, bb = rhs(m, 2, 9)
    x = np.array(bb, copy=True)  # b on entry, the solution on exit
    inf, bwd, steps = ssids.ssids_solve_refine(x, ak, fk, opts)
    assert inf.flag >= 0 and bwd.max() <= BWD_TOL
    assert matrices.bwd_error(m.full, x, bb).max() <= BWD_TOL
    ssids.ssids_free(ak, fk)
    # an indefinite factorization has no Cholesky diagonal
    mi = matrices.laplace3d(6, diag=5.7)
    ak2, inf2 = analyse(mi, grid=(6, 6, 6), use_gpu=True)
    fk2, finf2 = ssids.ssids_factor(False, mi.val, ak2, ssids.SSIDSOptions(ordering=3))
    assert finf2.flag >= 0
    d, inf = ssids.ssids_enquire_posdef(ak2, fk2)
    assert d is None and inf.flag == ssids.SSIDS_ERROR_NOT_LLT
    ssids.ssids_free(ak2, fk2)


def test_not_positive_definite_fails(lib):
    # the shifted stencil has 636 negative eigenvalues at 64^3 and some at 12^3
    m = matrices.laplace3d(12, diag=5.7)
    ak, fk, finf, opts = posdef_factor(m, grid=(12, 12, 12))
    assert finf.flag == ssids.SSIDS_ERROR_NOT_POS_DEF
    assert not fk.numeric
    x = np.ones(m.n)
    assert ssids.ssids_solve(x, ak, fk, opts).flag == ssids.SSIDS_ERROR_CALL_SEQUENCE
    ssids.ssids_free(ak, fk)


def test_posdef_multiple_rhs_bitwise(lib):
    m = matrices.laplace3d(10, diag=6.0)
    ak, fk, finf, opts = posdef_factor(m, grid=(10, 10, 10))
    assert finf.flag == 0
    xt, b = rhs(m, 16, 11)
    x = np.array(b, copy=True)
    assert ssids.ssids_solve(x, ak, fk, opts).flag == 0
    assert matrices.bwd_error(m.full, x, b).max() <= BWD_TOL
    x0 = np.array(b[:, 5], copy=True)
    assert ssids.ssids_solve(x0, ak, fk, opts).flag == 0
    assert np.array_equal(x0, x[:, 5])
    ssids.ssids_free(ak, fk)
