"""Rutherford-Boeing files (spral_amd/rb.py over include/spral_rb.h) against
the reference's own rb_write / rb_read (src/rutherford_boeing.f90, compiled
where it lies into oracle/_ref/libref_rb.so; these tests skip when that was
not built).  Writes must match byte for byte; reads of either writer's files
must give the reference's arrays (the lower and upper forms entry for
entry, the expanded and diagonal-completed forms as the same matrix), and
the header fields of rb_peek.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import ref_rb
from spral_amd import rb

pytestmark = pytest.mark.skipif(not ref_rb.available(), reason="oracle/_ref/libref_rb.so not built")


def random_csc(m, n, density, seed, lower=False, spread=30.0):
    rng = np.random.default_rng(seed)
    a = sp.random(m, n, density=density, random_state=rng, format="csc")
    if lower:
        a = sp.tril(a + sp.identity(n) * (seed % 2), format="csc")
    a.sort_indices()
    val = rng.standard_normal(a.nnz) * 10.0 ** rng.uniform(-spread, spread, a.nnz)
    if a.nnz > 3:
        val[0] = 0.0
        val[1] = -0.0
        val[2] = 1e-200 if seed % 2 else -3e250
    return (a.indptr + 1).astype(np.int64), (a.indices + 1).astype(np.int32), val


CASES = [(1, 1, 1.0, 4, True), (30, 30, 0.3, 4, True), (200, 200, 0.05, 4, True),
         (50, 80, 0.1, 1, False), (120, 120, 0.05, 2, False), (3000, 3000, 0.002, 6, True),
         (40000, 40000, 0.0002, 4, True)]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("pattern", [False, True])
def test_write_matches_reference_bytes(lib, tmp_path, case, pattern):
    m, n, dens, mtype, lower = CASES[case]
    ptr, row, val = random_csc(m, n, dens, seed=case, lower=lower)
    v = None if pattern else val
    f_ours, f_ref = tmp_path / "ours.rb", tmp_path / "ref.rb"
    assert rb.rb_write(f_ours, mtype, m, n, ptr, row, v) == 0
    assert ref_rb.rb_write(f_ref, mtype, m, n, ptr, row, v) == 0
    a, b = f_ours.read_bytes(), f_ref.read_bytes()
    if a != b:
        la, lb = a.decode().splitlines(), b.decode().splitlines()
        bad = next(i for i in range(min(len(la), len(lb))) if la[i] != lb[i])
        pytest.fail(f"line {bad}:\n ours {la[bad]!r}\n ref  {lb[bad]!r}")


def canon(n, ptr, row, val):
    """Entries sorted within each column (the same matrix, any entry order)."""
    out = []
    for c in range(n):
        k = np.arange(ptr[c] - 1, ptr[c + 1] - 1)
        o = np.argsort(row[k], kind="stable")
        out.append((row[k][o], None if val is None else val[k][o]))
    return out


@pytest.mark.parametrize("case", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("mode", [(1, 0, False), (2, 0, False), (3, 0, False), (1, 1, False),
                                  (1, 0, True), (3, 0, True)])
def test_read_matches_reference(lib, tmp_path, case, mode):
    m, n, dens, mtype, lower = CASES[case]
    ptr, row, val = random_csc(m, n, dens, seed=10 + case, lower=lower)
    f = tmp_path / "a.rb"
    assert ref_rb.rb_write(f, mtype, m, n, ptr, row, val) == 0
    luf, values, diag = mode
    got = rb.rb_read(f, luf, values, diag)
    want = ref_rb.rb_read(f, luf, values, diag)
    assert got[0] == want[0] == 0
    assert got[1:3] == want[1:3] and got[6] == want[6]
    assert np.array_equal(got[3], want[3])
    if luf in (1, 2) and not diag:
        assert np.array_equal(got[4], want[4])
        if want[5] is not None:
            assert np.array_equal(got[5], want[5])
        else:
            assert got[5] is None
    else:
        for (ra, va), (rb_, vb) in zip(canon(got[2], got[3], got[4], got[5]),
                                       canon(want[2], want[3], want[4], want[5])):
            assert np.array_equal(ra, rb_)
            if vb is not None:
                assert np.array_equal(va, vb)


def test_round_trip_and_peek(lib, tmp_path):
    ptr, row, val = random_csc(500, 500, 0.01, seed=3, lower=True)
    f = tmp_path / "r.rb"
    assert rb.rb_write(f, 4, 500, 500, ptr, row, val, title="round trip", identifier="RT1") == 0
    h = rb.rb_peek(f)
    assert h["info"] == 0 and (h["m"], h["n"], h["nelt"]) == (500, 500, 0)
    assert h["nvar"] == h["nval"] == ptr[-1] - 1
    assert h["type_code"] == "rsa" and h["matrix_type"] == 4
    assert h["title"].rstrip() == "round trip" and h["identifier"].rstrip() == "RT1"
    info, m, n, p2, r2, v2, mt = rb.rb_read(f)
    assert info == 0 and np.array_equal(p2, ptr) and np.array_equal(r2, row)
    # 16 significant digits: within one part in 1e16 of the written values
    np.testing.assert_allclose(v2, val, rtol=1e-15, atol=0)


def test_errors(lib, tmp_path):
    assert rb.rb_write(tmp_path / "x.rb", 5, 2, 2, [1, 2, 3], [1, 2], [1.0, 2.0]) == -6
    assert rb.rb_read(tmp_path / "missing.rb")[0] == -1
    (tmp_path / "bad.rb").write_text("title\n1 1 1 1\nxyz 1 1 1 0\n(1i1) (1i1) (1e8.1)\n")
    assert rb.rb_read(tmp_path / "bad.rb")[0] == -2
    ptr, row, val = random_csc(10, 10, 0.3, seed=1, lower=True)
    f = tmp_path / "ok.rb"
    rb.rb_write(f, 4, 10, 10, ptr, row, val)
    assert rb.rb_read(f, lwr_upr_full=4)[0] == -11
    assert rb.rb_read(f, values=2)[0] == -13
