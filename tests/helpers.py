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


def solve_operators(solve, n, chunk=64):
    """The five solve jobs as dense n x n operators on user-ordered vectors:
    solve(b, job) applies job to the columns of b.  Returns a dict with
    keys 0..4 (src/ssids/ssids.f90:1195-1254 job numbering)."""
    eye = np.eye(n)
    ops = {}
    for job in range(5):
        cols = [solve(eye[:, c0:c0 + chunk].copy(), job) for c0 in range(0, n, chunk)]
        ops[job] = np.hstack(cols)
    return ops


def dinv_blocks(piv, d):
    """Block-diagonal D^{-1} in user order from enquire_indef's output:
    piv(v) = +-(pivot position of v) (1-based, negative for 2x2 members),
    d(1,k) diagonal and d(2,k) subdiagonal of D^{-1} at pivot k
    (src/ssids/ssids.f90:1300-1356)."""
    n = len(piv)
    q = np.abs(piv) - 1
    inv = np.empty(n, np.int64)
    inv[q] = np.arange(n)
    neg = piv[inv] < 0  # by pivot position
    # the reference writes -(0) = 0 for the first member of a 2x2 block at
    # pivot position 0 (src/ssids/cpu/NumericSubtree.hxx:454), which the
    # 1-based conversion (src/ssids/ssids.f90:1343-1352) reports as +1: a
    # run of negative positions from 1 of odd length pairs with position 0
    if n > 1 and not neg[0] and neg[1]:
        run = int(np.argmin(np.append(neg[1:], False)))
        neg[0] = run % 2 == 1
    Dp = np.zeros((n, n))
    k = 0
    while k < n:
        Dp[k, k] = d[0, k]
        if neg[k]:
            assert k + 1 < n and neg[k + 1], "2x2 members come in pairs"
            Dp[k + 1, k] = Dp[k, k + 1] = d[1, k]
            Dp[k + 1, k + 1] = d[0, k + 1]
            assert d[1, k + 1] == 0.0
            k += 2
        else:
            assert d[1, k] == 0.0
            k += 1
    return Dp[np.ix_(q, q)]


def check_ldlt_identities(A, piv, d, ops, tol=1e-10):
    """What A = Q^T L D L^T Q fixes, with Q the pivot order from piv:
      * the forward operator F (job 1) permuted by Q is unit lower
        triangular (exactly: untouched entries stay 0, the unit diagonal 1);
      * the diagonal operator (job 2) is exactly enquire's D^{-1} blocks;
      * F A F^T = D, i.e. F A F^T D^{-1} = I;
      * the backward operator (job 3) is F^T, diag_bwd (job 4) is F^T D^{-1},
        and the full solve (job 0) is A^{-1}.
    A is dense in user order."""
    n = A.shape[0]
    q = np.abs(piv) - 1
    assert np.array_equal(np.sort(q), np.arange(n))
    F = ops[1]
    Lq = np.empty_like(F)
    Lq[np.ix_(q, q)] = F  # Lq[q(v), q(w)] = F[v, w] = L^{-1}[q(v), q(w)]
    assert np.all(np.triu(Lq, 1) == 0.0)
    assert np.all(np.diag(Lq) == 1.0)
    Dinv = dinv_blocks(piv, d)
    assert np.array_equal(ops[2], Dinv)
    scale = max(1.0, np.abs(F).max()) ** 2 * np.abs(A).max() * max(1.0, np.abs(Dinv).max())
    assert np.abs(F @ A @ F.T @ Dinv - np.eye(n)).max() <= tol * scale
    fs = max(1.0, np.abs(F).max())
    assert np.abs(ops[3] - F.T).max() <= tol * fs
    assert np.abs(ops[4] - F.T @ Dinv).max() <= tol * fs * max(1.0, np.abs(Dinv).max())
    assert np.abs(A @ ops[0] - np.eye(n)).max() <= tol * scale
