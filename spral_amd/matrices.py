"""Seeded synthetic matrices of the benchmark workloads (lower-triangle CSC,
1-based, as SSIDS takes them: src/ssids/ssids.f90:148 analyse_double).

* laplace3d(nx, ny, nz, diag)      : 7-point stencil, off-diagonals -1;
* kkt_laplace(...)                 : [H A^T; A 0] with H a shifted 7-point
                                      Laplacian and A random with exactly k
                                      distinct N(0,1) entries per row;
* kkt_random(n, seed)              : one system of the batched workload (random
                                      SPD H, 20% constraint rows);
* kkt_stencil27(grid, m_con, seed)  : the mixed-precision KKT workload: H a
                                      27-point stencil with N(0,1) couplings
                                      and diagonal 27 + |N(0,1)|, A random
                                      with 4 N(0,1) entries per row;
* random_indef / random_posdef     : small random test matrices in the spirit of
                                      gen_random_indef / gen_random_posdef
                                      (tests/ssids/ssids.f90:2830-2900).
Seeds are numpy PCG64; every generator is deterministic for a given seed.
"""
import numpy as np
import scipy.sparse as sp

SEED_BASE = 20261202


class LowerCSC:
    """Lower triangle of a symmetric matrix in 1-based CSC."""

    def __init__(self, a: sp.spmatrix):
        a = sp.csc_matrix(a)
        low = sp.tril(a, format="csc")
        low.sort_indices()
        self.n = a.shape[0]
        self.ptr = (low.indptr.astype(np.int64) + 1)
        self.row = (low.indices.astype(np.int32) + 1)
        self.val = low.data.astype(np.float64)
        self.full = a

    @property
    def nz(self):
        return int(self.ptr[-1] - 1)

    def matvec(self, x):
        return self.full @ x


def _lap7(nx, ny, nz):
    def d1(k):
        e = np.ones(k)
        return sp.diags([-e[:-1], -e[:-1]], [-1, 1], shape=(k, k))
    ix, iy, iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    off = (sp.kron(iz, sp.kron(iy, d1(nx))) + sp.kron(iz, sp.kron(d1(ny), ix))
           + sp.kron(d1(nz), sp.kron(iy, ix)))
    return sp.csc_matrix(off)


def laplace3d(nx, ny=None, nz=None, diag=6.0) -> LowerCSC:
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    off = _lap7(nx, ny, nz)
    n = nx * ny * nz
    return LowerCSC(off + diag * sp.identity(n, format="csc"))


def random_rows(rng, nrows, ncols, k):
    """A (nrows x ncols) with exactly k distinct uniformly chosen columns per
    row, N(0,1) values."""
    cols = np.empty((nrows, k), dtype=np.int64)
    for i in range(nrows):
        cols[i] = rng.choice(ncols, size=k, replace=False)
    vals = rng.standard_normal((nrows, k))
    rows = np.repeat(np.arange(nrows), k)
    return sp.csr_matrix((vals.ravel(), (rows, cols.ravel())), shape=(nrows, ncols))


def kkt_laplace(grid=16, shift=-0.5, m_con=1024, k=3, seed=SEED_BASE + 1) -> LowerCSC:
    """Saddle-point K = [H A^T; A 0]: H = 7-pt grid^3 Laplacian (diag 6) +
    shift*I, A m_con x grid^3 with k N(0,1) entries per row."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nh = grid ** 3
    h = _lap7(grid, grid, grid) + (6.0 + shift) * sp.identity(nh, format="csc")
    a = random_rows(rng, m_con, nh, k)
    kk = sp.bmat([[h, a.T], [a, None]], format="csc")
    return LowerCSC(kk)


def kkt_random(n, seed=SEED_BASE + 4) -> LowerCSC:
    """One system of the batched workload: K = [H A^T; A 0] of order n with
    m = floor(0.2 n) constraint rows.  H (n - m square) has 8 off-diagonal
    entries per row before symmetrisation (4 random N(0,1) partners per row,
    mirrored), and the diagonal 1.1 * sum|off| of its row, so H is SPD; A
    (m x (n - m)) has exactly 3 distinct N(0,1) entries per row.  The
    inertia is then m negative eigenvalues when A has full row rank."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m_con = int(0.2 * n)
    nh = n - m_con
    r = np.repeat(np.arange(nh), 4)
    c = rng.integers(0, nh, r.size)
    keep = r != c
    off = sp.coo_matrix((rng.standard_normal(int(keep.sum())), (r[keep], c[keep])),
                        shape=(nh, nh)).tocsc()
    off = off + off.T
    diag = 1.1 * np.asarray(abs(off).sum(axis=1)).ravel() + 1e-3
    h = off + sp.diags(diag)
    a = random_rows(rng, m_con, nh, 3)
    kk = sp.bmat([[h, a.T], [a, None]], format="csc")
    return LowerCSC(kk)


def kkt_stencil27(grid=80, m_con=88000, k=4, seed=SEED_BASE + 3) -> LowerCSC:
    """K = [H A^T; A 0] of order grid^3 + m_con.  H couples each grid point
    with its 26 neighbours of the 3x3x3 cube; each coupling is one N(0,1)
    value (the same in both triangles), and the diagonal is 27 + |N(0,1)|.
    The 26 couplings of a row sum to about 20.7 in absolute value, so H is
    diagonally dominant in practice, and with A (m_con x grid^3, exactly k
    distinct N(0,1) columns per row) of full row rank the inertia has m_con
    negative eigenvalues."""
    rng = np.random.Generator(np.random.PCG64(seed))
    g = grid
    nh = g ** 3
    idx = np.arange(nh).reshape(g, g, g)  # [z, y, x]
    rows, cols = [], []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if (dz, dy, dx) <= (0, 0, 0):
                    continue  # each unordered pair once: offsets after 0
                src = idx[max(0, -dz):g - max(0, dz), max(0, -dy):g - max(0, dy),
                          max(0, -dx):g - max(0, dx)]
                dst = idx[max(0, dz):g + min(0, dz) or None, max(0, dy):g + min(0, dy) or None,
                          max(0, dx):g + min(0, dx) or None]
                rows.append(dst.ravel())
                cols.append(src.ravel())
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = rng.standard_normal(r.size)
    low = sp.coo_matrix((v, (np.maximum(r, c), np.minimum(r, c))), shape=(nh, nh)).tocsc()
    h = low + low.T + sp.diags(27.0 + np.abs(rng.standard_normal(nh)))
    a = random_rows(rng, m_con, nh, k)
    kk = sp.bmat([[h, a.T], [a, None]], format="csc")
    return LowerCSC(kk)


def _lower_from_coo(n, r, c, v):
    """LowerCSC from lower-triangle coordinates, explicit zeros kept."""
    low = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsc()
    low.sum_duplicates()
    low.sort_indices()
    out = LowerCSC.__new__(LowerCSC)
    out.n = n
    out.ptr = low.indptr.astype(np.int64) + 1
    out.row = low.indices.astype(np.int32) + 1
    out.val = low.data.astype(np.float64)
    out.full = sp.csc_matrix(low + sp.tril(low, k=-1).T)
    return out


def random_indef(n, density=0.05, seed=0, zero_diag_frac=0.2, big_last=True) -> LowerCSC:
    """Random symmetric indefinite matrix: U(-1,1) strictly-lower entries, a
    structurally full diagonal of which a fraction is zero-valued, and (as the
    reference generator does, tests/ssids/ssids.f90:2830-2871) the last entry
    of each column scaled by 1000 to provoke delayed pivots."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nnz = max(n, int(density * n * n / 2))
    r = rng.integers(0, n, nnz)
    c = rng.integers(0, n, nnz)
    keep = r != c
    lo, hi = np.minimum(r, c)[keep], np.maximum(r, c)[keep]
    key = np.unique(hi.astype(np.int64) * n + lo)
    hi, lo = key // n, key % n
    v = rng.uniform(-1, 1, key.size)
    if big_last:
        # last (largest-row) entry of each column
        order = np.lexsort((hi, lo))
        hi, lo, v = hi[order], lo[order], v[order]
        last = np.ones(key.size, bool)
        last[:-1] = lo[1:] != lo[:-1]
        v[last] *= 1000.0
    d = rng.uniform(-1, 1, n)
    d[rng.random(n) < zero_diag_frac] = 0.0
    rr = np.concatenate([hi, np.arange(n)])
    cc = np.concatenate([lo, np.arange(n)])
    vv = np.concatenate([v, d])
    return _lower_from_coo(n, rr, cc, vv)


def random_posdef(n, density=0.05, seed=0) -> LowerCSC:
    rng = np.random.Generator(np.random.PCG64(seed))
    nnz = max(n, int(density * n * n / 2))
    r = rng.integers(0, n, nnz)
    c = rng.integers(0, n, nnz)
    lo, hi = np.minimum(r, c), np.maximum(r, c)
    v = rng.uniform(-1, 1, nnz)
    a = sp.coo_matrix((v, (hi, lo)), shape=(n, n)).tocsc()
    a.sum_duplicates()
    a = sp.tril(a, k=-1, format="csc")
    full = a + a.T
    dd = np.asarray(abs(full).sum(axis=1)).ravel() + 1.0
    return LowerCSC(full + sp.diags(dd))


def bwd_error(full, x, b):
    """Scaled backward error max|b-Ax| / (||A||_inf ||x||_inf + ||b||_inf)
    (tests/ssids/ssids.f90:1762-1810), per right-hand side."""
    x = np.asarray(x)
    b = np.asarray(b)
    if x.ndim == 1:
        x = x[:, None]
        b = b[:, None]
    r = b - full @ x
    anorm = abs(full).sum(axis=1).max()
    out = []
    for j in range(x.shape[1]):
        den = anorm * np.abs(x[:, j]).max() + np.abs(b[:, j]).max()
        out.append(np.abs(r[:, j]).max() / den if den > 0 else 0.0)
    return np.array(out)
