"""The synthetic workload generators (spral_amd/matrices.py).

kkt_stencil27 is the mixed-precision KKT workload. Pinned here by counts and
brute force, not by retyping its formula:
* H's pattern is the 27-point stencil: row i couples with exactly the grid
  points whose coordinates differ from its own by at most 1 in every
  dimension (closed-form count per row, checked against the generator);
* H is symmetric with diagonal 27 + |N(0,1)| >= 27, the (2,2) block is empty,
  and every row of A has exactly 4 distinct entries;
* the full-size instance (80^3 grid, 88,000 constraint rows) has 7,348,636
  lower-triangle entries: 512,000 diagonal + 6,484,636 stencil couplings
  (sum over the 13 half-space offsets of the (80-|dx|)(80-|dy|)(80-|dz|)
  pairs) + 352,000 constraint entries;
* inertia: dense eigvalsh of a small instance gives H > 0 and exactly m_con
  negative eigenvalues of K (Sylvester / Haynsworth with A of full row rank).
"""
import itertools

import numpy as np

from spral_amd import matrices


def half_offsets():
    return [o for o in itertools.product((-1, 0, 1), repeat=3) if o > (0, 0, 0)]


def test_stencil27_pattern_and_blocks():
    g, mc = 5, 21
    m = matrices.kkt_stencil27(g, mc)
    nh = g ** 3
    assert m.n == nh + mc
    f = m.full.tocsr()
    h = f[:nh, :nh].toarray()
    assert np.array_equal(h, h.T)
    assert np.all(np.diag(h) >= 27.0)
    for i in range(nh):
        z, y, x = np.unravel_index(i, (g, g, g))
        want = 1
        for c in (x, y, z):
            want *= (2 if c in (0, g - 1) else 3)
        assert np.count_nonzero(h[i]) == want, i
        for j in np.nonzero(h[i])[0]:
            zj, yj, xj = np.unravel_index(j, (g, g, g))
            assert max(abs(z - zj), abs(y - yj), abs(x - xj)) <= 1
    a = f[nh:, :nh].toarray()
    assert np.all(np.count_nonzero(a, axis=1) == 4)
    assert f[nh:, nh:].nnz == 0


def test_stencil27_full_size_count():
    g = 80
    pairs = sum((g - abs(dx)) * (g - abs(dy)) * (g - abs(dz)) for dz, dy, dx in half_offsets())
    assert pairs == 6_484_636
    m = matrices.kkt_stencil27(g, 88000)
    assert m.n == 600_000
    assert m.nz == g ** 3 + pairs + 4 * 88000 == 7_348_636


def test_stencil27_inertia_brute_force():
    g, mc = 6, 37
    m = matrices.kkt_stencil27(g, mc)
    k = m.full.toarray()
    nh = g ** 3
    assert np.linalg.eigvalsh(k[:nh, :nh]).min() > 0
    assert np.linalg.matrix_rank(k[nh:, :nh]) == mc
    ev = np.linalg.eigvalsh(k)
    assert int((ev < 0).sum()) == mc and np.abs(ev).min() > 1e-8
