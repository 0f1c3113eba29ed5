"""Serialised factors (spral_ssids_hip_save_numeric / _load_numeric): a
handle rebuilt from the file on the same analyse solves, refines and
enquires bit for bit as the original, keeps its stats, and can be
refactorized; files of another analyse or damaged files are rejected."""
import numpy as np
import pytest

from spral_amd import matrices, ssids
from helpers import analyse, rhs

pytestmark = pytest.mark.gpu


def _factor(m, posdef=False, **kw):
    ak, inf = analyse(m, use_gpu=True)
    assert inf.flag >= 0
    opts = ssids.SSIDSOptions(**kw)
    fk, finf = ssids.ssids_factor(posdef, m.val, ak, opts)
    assert finf.flag >= 0, finf
    return ak, fk, opts


@pytest.mark.parametrize("case", ["delays", "scaled", "posdef", "kkt"])
def test_saved_factors_solve_bitwise(lib, tmp_path, case):
    if case == "delays":
        m, posdef, kw = matrices.random_indef(400, density=0.02, seed=3), False, {}
    elif case == "scaled":
        m, posdef, kw = matrices.random_indef(300, density=0.03, seed=8), False, {"scaling": 4}
    elif case == "posdef":
        m, posdef, kw = matrices.laplace3d(10, diag=6.5), True, {}
    else:
        m, posdef, kw = matrices.kkt_laplace(grid=8, m_con=100), False, {}
    print(f"fkeep io case={case} n={m.n}", flush=True)
    ak, fk, opts = _factor(m, posdef, **kw)
    path = tmp_path / "factors.bin"
    assert ssids.ssids_save_fkeep(fk, path) == 0
    fk2, inf2 = ssids.ssids_load_fkeep(ak, path)
    assert inf2.flag == fk.inform.flag and fk2.numeric
    for k in ("num_neg", "num_two", "num_delay", "matrix_rank", "num_factor"):
        assert getattr(inf2, k) == getattr(fk.inform, k), k
    assert fk2.pos_def == posdef
    assert (fk2.scaling is None) == (fk.scaling is None)
    if fk.scaling is not None:
        assert np.array_equal(fk2.scaling, fk.scaling)
    _, b = rhs(m, 3, 11)
    x1, x2 = b.copy(), b.copy()
    ssids.ssids_solve(x1, ak, fk, opts)
    ssids.ssids_solve(x2, ak, fk2, opts)
    assert np.array_equal(x1, x2)
    assert matrices.bwd_error(m.full, x2, b).max() <= 1e-13
    y1, y2 = b.copy(), b.copy()
    i1, bw1, it1 = ssids.ssids_solve_refine(y1, ak, fk, opts)
    i2, bw2, it2 = ssids.ssids_solve_refine(y2, ak, fk2, opts)
    assert np.array_equal(y1, y2) and it1 == it2 and np.array_equal(bw1, bw2)
    if not posdef:
        p1, d1 = ssids.ssids_enquire_indef(ak, fk)[:2]
        p2, d2 = ssids.ssids_enquire_indef(ak, fk2)[:2]
        assert np.array_equal(p1, p2) and np.array_equal(d1, d2)
    # a loaded handle refactorizes like any other (new values, same
    # pattern): its solve equals a new handle's factorization of those values
    val2 = m.val * 1.5
    fk3, inf3 = ssids.ssids_factor(posdef, val2, ak, opts, fkeep=fk2)
    assert inf3.flag >= 0
    fk4, inf4 = ssids.ssids_factor(posdef, val2, ak, opts)
    assert inf4.flag >= 0
    x3, x4 = b.copy(), b.copy()
    ssids.ssids_solve(x3, ak, fk3, opts)
    ssids.ssids_solve(x4, ak, fk4, opts)
    assert np.array_equal(x3, x4)
    ssids.ssids_free(None, fk3)
    ssids.ssids_free(None, fk4)
    ssids.ssids_free(ak, fk)


def test_load_rejects_other_analyse_and_damaged_files(lib, tmp_path):
    m = matrices.random_indef(200, density=0.03, seed=4)
    ak, fk, opts = _factor(m)
    path = tmp_path / "f.bin"
    assert ssids.ssids_save_fkeep(fk, path) == 0
    other = matrices.random_indef(200, density=0.03, seed=5)
    ak2, fk2, _ = _factor(other)
    _, inf = ssids.ssids_load_fkeep(ak2, path)
    assert inf.flag == -2
    data = path.read_bytes()
    (tmp_path / "short.bin").write_bytes(data[: len(data) // 2])
    _, inf = ssids.ssids_load_fkeep(ak, tmp_path / "short.bin")
    assert inf.flag == -1
    _, inf = ssids.ssids_load_fkeep(ak, tmp_path / "missing.bin")
    assert inf.flag == -1
    ssids.ssids_free(ak, fk)
    ssids.ssids_free(ak2, fk2)
