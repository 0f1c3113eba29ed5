"""The schedule autotuner (ssids.ssids_autotune, no reference counterpart)
and the level_schedule option.  Every tree-level schedule factors each front
from the same inputs in the same order, so the factors and the solutions
must be bitwise the same under each schedule; the autotuner keeps the
schedule with the least device time."""
import numpy as np
import pytest

from spral_amd import batch, matrices, ssids
from helpers import rhs

pytestmark = pytest.mark.gpu


def factor_solve(m, o, b, grid=None):
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o, grid=grid)
    assert inf.flag >= 0, inf
    fk, finf = ssids.ssids_factor(False, m.val, ak, o)
    assert finf.flag >= 0, finf
    x = np.array(b, copy=True)
    assert ssids.ssids_solve(x, ak, fk, o).flag >= 0
    ssids.ssids_free(ak, fk)
    return finf, x


@pytest.mark.parametrize("case", ["grid24", "kkt_amd"])
def test_schedules_bitwise(lib, case):
    if case == "grid24":
        m, grid, base = matrices.laplace3d(24, diag=5.7), (24, 24, 24), dict(ordering=3)
    else:  # delays and TPP at the root (approximate minimum degree order)
        m, grid, base = matrices.kkt_laplace(), None, dict(ordering=4)
    xt, b = rhs(m, 2, 11)
    ref = None
    for s in (1, 2, 3):
        finf, x = factor_solve(m, ssids.SSIDSOptions(level_schedule=s, **base), b, grid)
        print(f"  {case} schedule {s}: levels {finf.num_levels}, {finf.factor_ms:.2f} ms, "
              f"delays {finf.num_delay}", flush=True)
        if ref is None:
            ref = (finf, x)
            continue
        assert finf.num_neg == ref[0].num_neg and finf.num_delay == ref[0].num_delay
        assert np.array_equal(x, ref[1]), np.max(np.abs(x - ref[1]))


def test_autotune_keeps_fastest(lib):
    m = matrices.laplace3d(24, diag=5.7)
    o = ssids.SSIDSOptions(ordering=3)
    xt, b = rhs(m, 1, 5)
    f0, x0 = factor_solve(m, o, b, (24, 24, 24))
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o, grid=(24, 24, 24))
    times = ssids.ssids_autotune(ak, m.val, o, reps=2)
    print(f"  autotune: {times} -> {ak.level_schedule}", flush=True)
    assert sorted(times) == [1, 2, 3] and all(t > 0 for t in times.values())
    assert ak.level_schedule == min(times, key=lambda s: (times[s], s))
    assert ak.schedule_ms == times
    fk, finf = ssids.ssids_factor(False, m.val, ak, o)
    assert finf.flag >= 0 and finf.num_neg == f0.num_neg
    x = np.array(b, copy=True)
    assert ssids.ssids_solve(x, ak, fk, o).flag >= 0
    assert np.array_equal(x, x0)
    ssids.ssids_free(ak, fk)


def test_autotune_batch(lib):
    """A batch of independent systems (a forest): the autotuned batch gives
    each system's inertia and solution of the batch under the default
    schedule, bitwise."""
    rng = np.random.default_rng(3)
    mats = [matrices.kkt_random(int(rng.integers(50, 400)), seed=300 + i) for i in range(40)]
    rhs_ = [m.full @ rng.uniform(-1, 1, m.n) for m in mats]
    b0, inf0 = batch.batch_analyse(mats)
    f0, neg0, rank0 = batch.batch_factor(b0, mats)
    assert f0.flag >= 0
    xs0 = batch.batch_solve(b0, rhs_)
    b1, inf1 = batch.batch_analyse(mats, autotune=True)
    assert inf1.flag >= 0 and sorted(b1.akeep.schedule_ms) == [1, 2, 3]
    print(f"  batch autotune: {b1.akeep.schedule_ms} -> {b1.akeep.level_schedule}", flush=True)
    f1, neg1, rank1 = batch.batch_factor(b1, mats)
    assert f1.flag >= 0
    assert np.array_equal(neg0, neg1) and np.array_equal(rank0, rank1)
    xs1 = batch.batch_solve(b1, rhs_)
    for a, c in zip(xs0, xs1):
        assert np.array_equal(a, c)
    batch.batch_free(b0)
    batch.batch_free(b1)
