"""Analyse residency: an analysed pattern saved and reloaded (arrays only,
nothing pickled), and the per-pattern analyse cache of a resident service.
The reloaded arrays must equal the originals exactly; the cache must return
the stored analyse for a pattern it has seen and a new one for a new pattern
or new ordering options."""
import numpy as np
import pytest

from spral_amd import matrices, ssids


def test_save_load_round_trip(lib, tmp_path):
    m = matrices.kkt_laplace(grid=8, m_con=100)
    o = ssids.SSIDSOptions(use_gpu=False)
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o)
    f = tmp_path / "ak.npz"
    ssids.ssids_save_akeep(ak, f)
    ak2, inf2 = ssids.ssids_load_akeep(f, o)
    for k in ("ptr", "row", "order", "invp", "sptr", "sparent", "rptr", "rlist", "nptr",
              "nlist"):
        assert np.array_equal(getattr(ak, k), getattr(ak2, k)), k
    assert (ak2.n, ak2.nnodes, ak2.num_factor, ak2.num_flops) == \
        (ak.n, ak.nnodes, ak.num_factor, ak.num_flops)
    assert inf2.maxfront == inf.maxfront and inf2.matrix_rank == inf.matrix_rank


def test_analyse_cache(lib):
    cache = ssids.AnalyseCache()
    o = ssids.SSIDSOptions(use_gpu=False)
    m1 = matrices.laplace3d(6, diag=6.0)
    m2 = matrices.laplace3d(7, diag=6.0)
    a1, _ = cache.analyse(m1.n, m1.ptr, m1.row, o)
    a1b, _ = cache.analyse(m1.n, m1.ptr.copy(), m1.row.copy(), o)
    assert a1b is a1 and len(cache) == 1
    a2, _ = cache.analyse(m2.n, m2.ptr, m2.row, o)
    assert a2 is not a1 and len(cache) == 2
    a3, _ = cache.analyse(m1.n, m1.ptr, m1.row, ssids.SSIDSOptions(use_gpu=False, nemin=8))
    assert a3 is not a1 and len(cache) == 3
    cache.free()
    assert len(cache) == 0


@pytest.mark.gpu
def test_reloaded_analyse_factors_bitwise(lib, tmp_path):
    m = matrices.kkt_laplace(grid=10, m_con=200)
    o = ssids.SSIDSOptions()
    ak, _ = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o)
    f = tmp_path / "ak.npz"
    ssids.ssids_save_akeep(ak, f)
    ak2, inf2 = ssids.ssids_load_akeep(f, o)
    assert inf2.flag >= 0 and ak2.symbolic
    rng = np.random.default_rng(4)
    b = m.full @ rng.uniform(-1, 1, m.n)
    xs = []
    for a in (ak, ak2):
        fk, fi = ssids.ssids_factor(False, m.val, a, o)
        assert fi.flag >= 0
        x = b.copy()
        assert ssids.ssids_solve(x, a, fk, o).flag >= 0
        xs.append(x)
        ssids.ssids_free(a, fk)
    assert np.array_equal(xs[0], xs[1])
