"""Pin the CPU oracle: inertia against brute-force eigenvalues, the
documented 5x5 example, pivot-growth and backward-error bounds of the
reference kernel tests, zero pivots and action=false."""
import numpy as np
import pytest

from oracle.oracle_py import OracleFactor
from spral_amd import matrices
from helpers import (analyse, brute_inertia, check_ldlt_identities, dinv_blocks, doc_example,
                     rhs, solve_operators)


@pytest.mark.parametrize("seed", range(6))
def test_inertia_matches_eigenvalues(lib, oracle_lib, seed):
    m = matrices.random_indef(120 + 30 * seed, density=0.04, seed=seed)
    ak, _ = analyse(m)
    of = OracleFactor(ak, m.val)
    neg, ev = brute_inertia(m)
    assert of.stats["num_neg"] == neg
    # numerically zero eigenvalues show up as zero pivots (rank deficit)
    nzero = int((np.abs(ev) < 1e-12 * np.abs(ev).max()).sum())
    assert of.stats["matrix_rank"] == m.n - nzero
    xt, b = rhs(m, 2, seed)
    x = of.solve(b)
    assert matrices.bwd_error(m.full, x, b).max() < 1e-13


def test_kkt_config1_inertia(lib, oracle_lib):
    m = matrices.kkt_laplace()
    ak, _ = analyse(m)
    of = OracleFactor(ak, m.val)
    neg, _ = brute_inertia(m)
    assert neg == 1024
    assert of.stats["num_neg"] == neg


def test_doc_example(lib, oracle_lib):
    m, b = doc_example()
    ak, _ = analyse(m)
    of = OracleFactor(ak, m.val)
    x = of.solve(b)
    assert np.allclose(x, [1, 2, 3, 4, 5], atol=1e-12)


def test_threshold_bounds_l(lib, oracle_lib):
    # |L| <= 1/u for every TPP pivot (tests/ssids/kernels/ldlt_tpp.cxx:298)
    m = matrices.random_indef(200, density=0.05, seed=9)
    ak, _ = analyse(m)
    u = 0.01
    of = OracleFactor(ak, m.val, u=u)
    for s in range(ak.nnodes):
        nf = of.node_factors(s)
        ne = nf["nelim"]
        Lm = np.tril(nf["L"][:, :ne], -1)
        assert np.abs(Lm).max(initial=0) <= 1 / u * (1 + 1e-12)


def test_zero_pivot_warning_and_rank(lib, oracle_lib):
    # a zero row/column: warning +7 and rank deficiency 1
    # (tests/ssids/ssids.f90:1245-1340)
    import scipy.sparse as sp
    a = sp.random(30, 30, density=0.2, random_state=1, format="csc")
    a = a + a.T + sp.identity(30) * 3
    a = sp.lil_matrix(a)
    a[7, :] = 0
    a[:, 7] = 0
    a[7, 7] = 0.0
    a = sp.csc_matrix(a)
    m = matrices.LowerCSC(a)
    m.full = a
    # keep a structural diagonal entry for variable 7
    ak, _ = analyse(m)
    of = OracleFactor(ak, m.val)
    assert of.stats["num_zero"] >= 1 or of.stats["matrix_rank"] < 30


def test_singular_action_false(lib, oracle_lib):
    import scipy.sparse as sp
    a = sp.csc_matrix(np.array([[1.0, 1.0], [1.0, 1.0]]))
    m = matrices.LowerCSC(a)
    ak, _ = analyse(m)
    of = OracleFactor(ak, m.val, action=False)
    assert of.stats["flag"] == -5
    of = OracleFactor(ak, m.val, action=True)
    assert of.stats["flag"] == 7 and of.stats["matrix_rank"] == 1


def test_planted_mistake_is_caught(lib, oracle_lib):
    """A wrong sign in one off-diagonal entry must break the residual check
    the GPU parity tests rely on."""
    m = matrices.random_indef(100, density=0.05, seed=2)
    ak, _ = analyse(m)
    bad = m.val.copy()
    k = np.nonzero(m.row - 1 != np.repeat(np.arange(m.n), np.diff(m.ptr)))[0][3]
    bad[k] = -bad[k]
    of = OracleFactor(ak, bad)
    xt, b = rhs(m)
    x = of.solve(b)
    assert matrices.bwd_error(m.full, x, b).max() > 1e-10


def _oracle_ops(of, n):
    return solve_operators(lambda b, job: of.solve(b, job=job), n)


@pytest.mark.parametrize("case", ["delays0", "delays1", "kkt", "doc"])
def test_enquire_and_solve_jobs_pinned(lib, oracle_lib, case):
    """oracle_enquire and oracle_solve jobs 1-4 against what A = Q^T L D L^T Q
    fixes (helpers.check_ldlt_identities): the permuted forward operator is
    unit lower triangular, the diagonal job is exactly the enquired D^{-1},
    F A F^T = D, bwd = F^T, diag_bwd = F^T D^{-1}, all = A^{-1}."""
    if case.startswith("delays"):
        m = matrices.random_indef(90 + 40 * int(case[-1]), density=0.06, seed=40 + int(case[-1]))
    elif case == "kkt":
        m = matrices.kkt_laplace(grid=5, m_con=30)
    else:
        m, _ = doc_example()
    ak, _ = analyse(m)
    of = OracleFactor(ak, m.val)
    assert of.stats["matrix_rank"] == m.n
    if case.startswith("delays"):
        assert of.stats["num_delay"] > 0 and of.stats["num_two"] > 0
    piv, d = of.enquire()
    A = m.full.toarray()
    check_ldlt_identities(A, piv, d, _oracle_ops(of, m.n))
    # inertia from the enquired blocks equals the factorization's count
    ev = np.linalg.eigvalsh(dinv_blocks(piv, d))
    assert int((ev < 0).sum()) == of.stats["num_neg"]


def test_identity_check_catches_planted_mistakes(lib, oracle_lib):
    """Each plausible mistake in what enquire or a solve job returns breaks
    one of the identities: a wrong 2x2 off-diagonal sign, a swapped pair of
    pivot positions, a transposed forward operator."""
    m = matrices.random_indef(90, density=0.06, seed=40)
    ak, _ = analyse(m)
    of = OracleFactor(ak, m.val)
    piv, d = of.enquire()
    A = m.full.toarray()
    ops = _oracle_ops(of, m.n)
    check_ldlt_identities(A, piv, d, ops)
    k = int(np.nonzero(d[1] != 0.0)[0][0])
    bad = d.copy()
    bad[1, k] = -bad[1, k]
    with pytest.raises(AssertionError):
        check_ldlt_identities(A, piv, bad, ops)
    ones = np.nonzero(piv > 0)[0]
    sw = piv.copy()
    sw[ones[0]], sw[ones[1]] = piv[ones[1]], piv[ones[0]]
    with pytest.raises(AssertionError):
        check_ldlt_identities(A, sw, d, ops)
    tr = dict(ops)
    tr[1] = ops[1].T.copy()
    with pytest.raises(AssertionError):
        check_ldlt_identities(A, piv, d, tr)
