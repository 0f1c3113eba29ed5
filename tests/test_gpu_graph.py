"""The static-plan factorization replayed as a captured graph.

A refactor on a warm handle replays the graph captured by the first
factorization on its plan. The eager path (SPRAL_SSIDS_HIP_NO_GRAPH=1) is
the reference: replays alternate two value sets, with a solve, host work
and synchronizations between them, and every replay's solution must equal
the eager factorization's of the same values bit for bit (a replay that
missed new values or reused stale device state would differ).
"""
import ctypes as C
import os

import numpy as np
import pytest

from spral_amd import _lib, devmem, matrices, ssids
from helpers import analyse, rhs

pytestmark = pytest.mark.gpu


def cases():
    return [("lap12", matrices.laplace3d(12, diag=5.7), (12, 12, 12)),
            ("kkt", matrices.kkt_laplace(grid=8, m_con=100), None),
            ("rnd", matrices.random_indef(300, density=0.03, seed=2), None)]


def solve(L, num, ak, b):
    x = np.array(b[ak.invp], copy=True)
    assert ssids._solve_all(L, num, 1, x, ak.n) == 0
    out = np.empty_like(x)
    out[ak.invp] = x
    return out


@pytest.mark.parametrize("case", range(3))
def test_graph_replays_equal_eager(lib, case):
    name, m, grid = cases()[case]
    L = _lib.lib()
    ak, _ = analyse(m, grid=grid, use_gpu=True)
    opts = (ssids.SSIDSOptions(ordering=3) if grid else ssids.SSIDSOptions()).to_c()
    rng = np.random.default_rng(case)
    vals = [m.val.copy(), m.val * (1.0 + 0.01 * rng.standard_normal(m.val.size))]
    d_vals = [devmem.DeviceArray.from_host(v) for v in vals]
    b = rhs(m, 1, 3)[1][:, 0]

    def run(graph):
        if graph:
            os.environ.pop("SPRAL_SSIDS_HIP_NO_GRAPH", None)
        else:
            os.environ["SPRAL_SSIDS_HIP_NO_GRAPH"] = "1"
        st = _lib.Stats()
        num = L.spral_ssids_hip_factor_device(False, ak.symbolic, d_vals[0].ptr, m.val.size,
                                              None, C.byref(opts), C.byref(st))
        assert num and st.flag >= 0
        out = []
        for i in (0, 1, 0, 1, 1, 0):  # first: the factor above; then refactors
            if out or i:
                s = _lib.Stats()
                assert L.spral_ssids_hip_refactor_device(num, d_vals[i].ptr, None,
                                                         C.byref(opts), C.byref(s)) >= 0
            out.append(solve(L, num, ak, b))
            devmem.synchronize()
            np.linalg.norm(out[-1])  # host work between replays
        L.spral_ssids_hip_destroy_numeric(num)
        return out

    try:
        eager = run(False)
        graph = run(True)
    finally:
        os.environ.pop("SPRAL_SSIDS_HIP_NO_GRAPH", None)
    print(f"  {name}: {len(graph)} factorizations compared", flush=True)
    for k, (xe, xg) in enumerate(zip(eager, graph)):
        assert np.array_equal(xe, xg), (name, k)
    assert not np.array_equal(eager[0], eager[1])  # the two value sets differ
    ssids.ssids_free(ak)
