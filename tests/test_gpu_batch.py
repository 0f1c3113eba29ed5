"""Many independent systems in one factor + solve call (spral_amd/batch.py).

The batch is one block-diagonal matrix whose assembly forest has one tree
per system, so every launch covers the fronts of all systems. Bars, written
in the tests:
* each system's solution is bitwise the one of factoring and solving it
  alone with the same order (alone and in the mix give the same result);
* per-system inertia equals the system's own factorization's, and the
  closed form floor(0.2 n) of the KKT generator (dense eigenvalue counts
  confirm it on the smaller systems);
* after FP64 refinement every system reaches the scaled backward error
  1e-14.
"""
import numpy as np
import pytest

from spral_amd import batch, matrices, ssids

pytestmark = pytest.mark.gpu

BWD_TOL = 1e-14


def systems(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 5 == 4:  # delays and 2x2 pivots
            out.append(matrices.random_indef(int(rng.integers(60, 200)), density=0.05,
                                             seed=seed * 100 + i))
        else:
            out.append(matrices.kkt_random(int(rng.integers(50, 500)), seed=seed * 100 + i))
    return out


def alone(m, order):
    o = ssids.SSIDSOptions(ordering=0)
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o, order=order)
    fk, finf = ssids.ssids_factor(False, m.val, ak, o)
    return ak, fk, finf, o


@pytest.mark.parametrize("seed", range(3))
def test_batch_matches_alone_bitwise(lib, seed):
    mats = systems(25, seed)
    b, inf = batch.batch_analyse(mats)
    assert inf.flag >= 0
    finf, neg, rank = batch.batch_factor(b, mats)
    assert finf.flag >= 0, finf
    print(f"batch seed={seed}: {len(mats)} systems, n={b.akeep.n}, "
          f"{b.akeep.nnodes} fronts, levels {finf.num_levels}", flush=True)
    rng = np.random.default_rng(seed + 7)
    rhs = [m.full @ rng.uniform(-1, 1, m.n) for m in mats]
    xs = batch.batch_solve(b, rhs)
    for i, m in enumerate(mats):
        ak, fk, fi, o = alone(m, b.orders[i])
        assert fi.flag >= 0
        assert neg[i] == fi.num_neg, (i, neg[i], fi.num_neg)
        assert rank[i] == fi.matrix_rank
        x = np.array(rhs[i], copy=True)
        assert ssids.ssids_solve(x, ak, fk, o).flag >= 0
        assert np.array_equal(xs[i], x), (i, np.max(np.abs(xs[i] - x)))
        if i % 5 != 4:
            assert neg[i] == int(0.2 * m.n)
            if m.n <= 200:
                assert neg[i] == int((np.linalg.eigvalsh(m.full.toarray()) < 0).sum())
        ssids.ssids_free(ak, fk)
    batch.batch_free(b)


def test_batch_refined_accuracy(lib):
    mats = [matrices.kkt_random(n, seed=900 + n) for n in (200, 800, 1500, 3000, 5000)]
    b, inf = batch.batch_analyse(mats)
    finf, neg, rank = batch.batch_factor(b, mats)
    assert finf.flag >= 0
    assert list(neg) == [int(0.2 * m.n) for m in mats]
    rng = np.random.default_rng(3)
    rhs = [m.full @ rng.uniform(-1, 1, m.n) for m in mats]
    xs = batch.batch_solve(b, rhs, refine=True)
    for m, x, r in zip(mats, xs, rhs):
        assert matrices.bwd_error(m.full, x, r).max() <= BWD_TOL
    batch.batch_free(b)


@pytest.mark.timeout(900)
def test_batch_config4_full_size(lib):
    """4,096 KKT systems, n uniform in [200, 5000], each in the matching-based
    order with its scaling (the batched workload at its full size): every
    system's inertia is the generator's floor(0.2 n) and every solution
    refines to the 1e-14 bar."""
    rng = np.random.Generator(np.random.PCG64(20261202 + 4))
    sizes = rng.integers(200, 5001, 4096)
    mats = [matrices.kkt_random(int(n), seed=20261202 + 4 + i) for i, n in enumerate(sizes)]
    b, inf = batch.batch_analyse(mats, ssids.SSIDSOptions(ordering=2, scaling=3))
    assert inf.flag >= 0, inf
    val = np.concatenate([m.val for m in mats])
    finf, neg, rank = batch.batch_factor(b, val)
    assert finf.flag >= 0, finf
    print(f"4096 systems: n={b.akeep.n}, {b.akeep.nnodes} fronts, levels {finf.num_levels}, "
          f"delays {finf.num_delay}", flush=True)
    assert list(neg) == [int(0.2 * m.n) for m in mats]
    assert list(rank) == [m.n for m in mats]
    rhs = [m.full @ rng.uniform(-1, 1, m.n) for m in mats]
    xs = batch.batch_solve(b, rhs, refine=True)
    worst = max(float(matrices.bwd_error(m.full, x, r).max()) for m, x, r in zip(mats, xs, rhs))
    print(f"  worst refined bwd err {worst:.2e}", flush=True)
    assert worst <= BWD_TOL
    batch.batch_free(b)
