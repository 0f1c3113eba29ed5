"""The resident solver service (spral_amd/serve.py).

Three patterns get interleaved refactor and solve requests (1, 4 or 16
right-hand sides) from one producer thread each. The HBM budget is small
enough to force evictions, so some patterns are refactored from scratch on
their next request. Bars, written below: every result is bitwise the one of
replaying its pattern's requests alone (the refactor, then that solve by
itself, with the same FP64 refinement), and every solve reaches the scaled
backward error 1e-14. Solves were batched (fewer rounds than requests), and
p50/p99 latencies are reported.
"""
import threading

import numpy as np
import pytest
import scipy.sparse as sp

from spral_amd import matrices, serve, ssids

pytestmark = pytest.mark.gpu

BWD_TOL = 1e-14


def scripts(seed):
    rng = np.random.default_rng(seed)
    pats = {"lap": matrices.laplace3d(10, diag=5.7),
            "kkt": matrices.kkt_laplace(grid=8, m_con=100),
            "rnd": matrices.random_indef(300, density=0.03, seed=5)}
    plan = {}
    for name, m in pats.items():
        steps = [("refactor", m.val.copy())]
        for i in range(14):
            if rng.random() < 0.2:
                # new values on the same pattern: a refactor
                v = m.val * (1.0 + 0.01 * rng.standard_normal(m.val.size))
                steps.append(("refactor", v))
            else:
                k = int(rng.choice([1, 4, 16]))
                b = rng.uniform(-1, 1, (m.n, k))
                steps.append(("solve", b if k > 1 else b[:, 0]))
        plan[name] = steps
    return pats, plan


def test_service_results_equal_alone(lib):
    pats, plan = scripts(11)
    opts = ssids.SSIDSOptions()
    svc = serve.SolverService(hbm_budget_bytes=3 << 20, options=opts)  # forces evictions
    for name, m in pats.items():
        svc.register(name, m.n, m.ptr, m.row)
    svc.start()
    futures = {name: [] for name in plan}

    def produce(name):
        for kind, data in plan[name]:
            f = svc.submit_refactor(name, data) if kind == "refactor" else \
                svc.submit_solve(name, data)
            futures[name].append(f)

    th = [threading.Thread(target=produce, args=(n,)) for n in plan]
    for t in th:
        t.start()
    for t in th:
        t.join()
    results = {n: [f.result(timeout=120) for f in fs] for n, fs in futures.items()}
    nreq = sum(len(v) for v in plan.values())
    print(f"  {nreq} requests in {svc.rounds} rounds; solve latency {svc.latency_ms('solve')}",
          flush=True)
    svc.stop()
    # replay alone
    for name, m in pats.items():
        o = ssids.SSIDSOptions()
        ak, _ = ssids.ssids_analyse(False, m.n, m.ptr, m.row, o)
        fk = None
        vals = None
        for (kind, data), got in zip(plan[name], results[name]):
            if kind == "refactor":
                vals = data
                if fk is not None:
                    ssids.ssids_free(None, fk)
                fk, inf = ssids.ssids_factor(False, vals, ak, o)
                assert got.flag == inf.flag and got.num_neg == inf.num_neg
                continue
            cols = data.reshape(m.n, -1)
            low = sp.csc_matrix((vals, m.row - 1, m.ptr - 1), shape=(m.n, m.n))
            full = low + sp.tril(low, k=-1).T
            for j in range(cols.shape[1]):
                x = cols[:, j].copy()
                ssids.ssids_solve_refine(x, ak, fk, o, tol=BWD_TOL)
                assert np.array_equal(got.reshape(m.n, -1)[:, j], x), (name, j)
                assert matrices.bwd_error(full, x, cols[:, j]).max() <= BWD_TOL
        ssids.ssids_free(ak, fk)
    assert svc.rounds < nreq


def test_refactor_and_solves_overlap(lib):
    """A long refactor of one pattern and solves of another are in flight
    together: solves submitted after the refactor complete before it does,
    and a solve of the resident version of the refactoring pattern itself is
    served while its next version is built.  Results equal replays alone."""
    import time
    big = matrices.laplace3d(40, diag=5.7)          # ~30 ms factorizations
    small = matrices.random_indef(300, density=0.03, seed=5)
    svc = serve.SolverService(options=ssids.SSIDSOptions())
    svc.register("big", big.n, big.ptr, big.row)
    svc.register("small", small.n, small.ptr, small.row)
    svc.start()
    svc.submit_refactor("big", big.val).result(timeout=120)
    svc.submit_refactor("small", small.val).result(timeout=120)
    done = {}

    def stamp(key):
        return lambda f: done.setdefault(key, time.perf_counter())

    rng = np.random.default_rng(3)
    b_big = rng.uniform(-1, 1, big.n)
    f_old = svc.submit_solve("big", b_big)          # version 1 of big
    f_ref = svc.submit_refactor("big", big.val * 1.25)
    f_ref.add_done_callback(stamp("refactor"))
    f_new = svc.submit_solve("big", b_big)          # version 2: waits for the refactor
    small_f = []
    for i in range(6):
        f = svc.submit_solve("small", rng.uniform(-1, 1, small.n))
        f.add_done_callback(stamp(f"small{i}"))
        small_f.append(f)
    x_old = f_old.result(timeout=120)
    x_new = f_new.result(timeout=120)
    for f in small_f:
        f.result(timeout=120)
    f_ref.result(timeout=120)
    svc.stop()
    early = [k for k in done if k.startswith("small") and done[k] < done["refactor"]]
    print(f"  small solves done before the big refactor: {len(early)} of 6", flush=True)
    assert early, done
    # version semantics: the solve submitted before the refactor used the old
    # values, the one after it the new ones
    o = ssids.SSIDSOptions()
    ak, _ = ssids.ssids_analyse(False, big.n, big.ptr, big.row, o)
    for vals, got in ((big.val, x_old), (big.val * 1.25, x_new)):
        fk, _ = ssids.ssids_factor(False, vals, ak, o)
        x = b_big.copy()
        ssids.ssids_solve_refine(x, ak, fk, o, tol=BWD_TOL)
        assert np.array_equal(got, x)
        ssids.ssids_free(None, fk)
    ssids.ssids_free(ak)


@pytest.mark.timeout(600)
def test_service_config5_patterns(lib):
    """The serving workload's patterns at their sizes (KKT: 7-point H on 16^3
    and 28^3 plus grid^3/8 constraint rows, n = 4,608 and 24,696), matching-
    based orders from the registered values and the matching scaling per
    refactor, refactors with log-uniform [1e-8, 1e8] barrier terms on H's
    diagonal (thousands of delayed pivots at 28^3) mixed with solves of 1 to
    64 right-hand sides: every result is bitwise the one of replaying the
    pattern alone, and every solve refines to 1e-14."""
    rng = np.random.default_rng(20261203)
    grids = (16, 28)
    pats = {f"kkt{g}": matrices.kkt_laplace(grid=g, m_con=g ** 3 // 8, seed=20261202 + 5 + g)
            for g in grids}
    opts = ssids.SSIDSOptions(ordering=2, scaling=1)
    plan = {}
    for g, (name, m) in zip(grids, pats.items()):
        col = np.repeat(np.arange(m.n), np.diff(m.ptr))
        dpos = np.nonzero((m.row - 1 == col) & (col < g ** 3))[0]
        steps = [("refactor", m.val.copy())]
        for i in range(8):
            if i % 3 == 1:
                v = m.val.copy()
                v[dpos] += 10.0 ** rng.uniform(-8, 8, dpos.size)
                steps.append(("refactor", v))
            else:
                k = int(rng.choice([1, 4, 16, 64]))
                steps.append(("solve", rng.uniform(-1, 1, (m.n, k))))
        plan[name] = steps
    svc = serve.SolverService(hbm_budget_bytes=64 << 30, options=opts)
    for name, m in pats.items():
        svc.register(name, m.n, m.ptr, m.row, val=m.val)
    svc.start()
    futures = {name: [] for name in plan}

    def produce(name):
        for kind, data in plan[name]:
            f = svc.submit_refactor(name, data) if kind == "refactor" else \
                svc.submit_solve(name, data)
            futures[name].append(f)

    th = [threading.Thread(target=produce, args=(n,)) for n in plan]
    for t in th:
        t.start()
    for t in th:
        t.join()
    results = {n: [f.result(timeout=300) for f in fs] for n, fs in futures.items()}
    print(f"  refactor latency {svc.latency_ms('refactor')}, "
          f"solve latency {svc.latency_ms('solve')}", flush=True)
    svc.stop()
    for name, m in pats.items():
        ak, _ = ssids.ssids_analyse(False, m.n, m.ptr, m.row, opts, val=m.val)
        fk = None
        vals = None
        for (kind, data), got in zip(plan[name], results[name]):
            if kind == "refactor":
                vals = data
                if fk is not None:
                    ssids.ssids_free(None, fk)
                fk, inf = ssids.ssids_factor(False, vals, ak, opts)
                assert got.flag == inf.flag and got.num_neg == inf.num_neg, name
                print(f"  {name}: refactor delays {inf.num_delay}", flush=True)
                continue
            low = sp.csc_matrix((vals, m.row - 1, m.ptr - 1), shape=(m.n, m.n))
            full = low + sp.tril(low, k=-1).T
            for j in range(data.shape[1]):
                x = data[:, j].copy()
                ssids.ssids_solve_refine(x, ak, fk, opts, tol=BWD_TOL)
                assert np.array_equal(got.reshape(m.n, -1)[:, j], x), (name, j)
                assert matrices.bwd_error(full, x, data[:, j]).max() <= BWD_TOL
        ssids.ssids_free(ak, fk)
