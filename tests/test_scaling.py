"""CPU checks of the equilibration-scaling oracle (oracle/scaling.py).

The oracle restates inf_norm_equilib_sym (src/scaling.f90:480-521).  It is
pinned three ways:
* bit for bit against the reference routine itself (src/scaling.f90 compiled
  where it lies into oracle/_ref/libref_scaling.so; skipped when the
  reference tree was absent at build time);
* closed forms: a diagonal matrix reaches |s_i^2 a_ii| = 1 after one step;
  the 2x2 matrix [[a, b], [b, c]] goes to s = |b|^{-1/2} (1, 1) in one step
  when |b| dominates, and to (|a|^{-1/2}, 1/(|b| |a|^{-1/2})) when
  |a| > |b| > |c|;
* the reference test's own property (tests/scaling.f90:455-470): after the
  call every scaled row's infinity norm is within 0.05 of 1.
"""
import numpy as np
import pytest

from oracle import scaling as osc

TOL32 = np.float32(1e-8)


def random_sym(n, seed, diag=True, spread=3.0, max_nz=20000):
    """Random lower CSC (1-based) with entries of widely varying size."""
    rng = np.random.default_rng(seed)
    want = min(n + int(rng.integers(0, max(1, n * n // 2 - n) + 1)), max_nz)
    r = rng.integers(0, n, want)
    c = rng.integers(0, n, want)
    r, c = np.maximum(r, c), np.minimum(r, c)
    if diag:
        r = np.concatenate([np.arange(n), r])
        c = np.concatenate([np.arange(n), c])
    key = np.unique(c.astype(np.int64) * n + r)
    col, row0 = key // n, key % n
    row = (row0 + 1).astype(np.int32)
    ptr = np.ones(n + 1, np.int64)
    ptr[1:] += np.cumsum(np.bincount(col, minlength=n))
    val = rng.standard_normal(len(key)) * 10.0 ** rng.uniform(-spread, spread, len(key))
    return ptr, row, val


def scaled_row_inf(n, ptr, row, val, s):
    """Infinity norm of every row of S A S (symmetric), as the reference test."""
    rinf = np.zeros(n)
    col = np.repeat(np.arange(n), np.diff(ptr))
    r = row - 1
    v = np.abs(s[col] * val * s[r])
    np.maximum.at(rinf, r, v)
    np.maximum.at(rinf, col, v)
    return rinf


def test_diagonal_closed_form():
    n = 50
    rng = np.random.default_rng(3)
    d = rng.standard_normal(n) * 10.0 ** rng.uniform(-8, 8, n)
    ptr = np.arange(1, n + 2, dtype=np.int64)
    row = np.arange(1, n + 1, dtype=np.int32)
    s, its = osc.equilib_scale_sym(n, ptr, row, d)
    # one step puts every |s_i^2 d_i| at 1 to within rounding; the exit test
    # then passes at the second iteration (reported as 1)
    assert its == 1
    assert np.max(np.abs(s * s * np.abs(d) - 1.0)) < 1e-15  # a few roundings
    assert np.all(s > 0)


def test_two_by_two_fixed_points():
    ptr = np.array([1, 3, 4], np.int64)
    row = np.array([1, 2, 2], np.int32)
    # |b| >= |a|, |c|: the first step gives s1 = s2 = |b|^{-1/2}, after which
    # every scaled row's maximum is |b| s1 s2 = 1 and the next test passes
    a, b, c = 4.0, -10.0, 0.25
    s, its = osc.equilib_scale_sym(2, ptr, row, np.array([a, b, c]))
    assert its == 1
    np.testing.assert_allclose(s, [10.0 ** -0.5] * 2, rtol=2e-16)
    # |a| > |b| > |c|: s1 settles at |a|^{-1/2} and s2 -> 1/(|b| s1) = 1 by
    # square roots (s2 = 10^{-2^-k}), about 27 steps to |1 - m| < 1e-8
    a, b, c = 100.0, 10.0, 0.01
    s, its = osc.equilib_scale_sym(2, ptr, row, np.array([a, b, c]), max_iterations=100)
    assert 20 <= its <= 35
    np.testing.assert_allclose(s, [0.1, 1.0], rtol=2e-8)


@pytest.mark.parametrize("seed", range(30))
def test_reference_property(seed):
    n = seed + 1 if seed < 20 else 50 + 37 * seed
    ptr, row, val = random_sym(n, seed)
    s, its = osc.equilib_scale_sym(n, ptr, row, val)
    rinf = scaled_row_inf(n, ptr, row, val, s)
    assert np.all(1.0 - rinf <= 0.05), (seed, its, rinf.min())


def test_empty_row_never_converges():
    # column 2 empty: its maximum stays 0, |1 - 0| = 1, so all iterations run
    ptr = np.array([1, 2, 2, 3], np.int64)
    row = np.array([1, 3], np.int32)
    val = np.array([2.0, 8.0])
    s, its = osc.equilib_scale_sym(3, ptr, row, val, max_iterations=7)
    assert its == 7
    assert s[1] == 1.0


@pytest.mark.skipif(not osc.ref_available(), reason="oracle/_ref/libref_scaling.so not built")
@pytest.mark.parametrize("seed", range(40))
def test_oracle_matches_reference_bitwise(seed):
    n = seed + 1 if seed < 20 else 40 + 29 * seed
    ptr, row, val = random_sym(n, 100 + seed, diag=(seed % 3 != 0))
    for maxit in (1, 3, 10):
        s_o, it_o = osc.equilib_scale_sym(n, ptr, row, val, max_iterations=maxit)
        s_r, it_r, fl = osc.ref_equilib_scale_sym(n, ptr, row, val, max_iterations=maxit)
        assert fl == 0
        assert it_o == it_r, (seed, maxit, it_o, it_r)
        assert np.array_equal(s_o, s_r), (seed, maxit, np.max(np.abs(s_o - s_r)))
