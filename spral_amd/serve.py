"""A resident solver service: refactor and solve requests on patterns whose
analyse and factors stay in HBM.

* Each pattern is analysed once (ssids.AnalyseCache), and its factorization
  handle is reused by every refactor of new values (its pool, static plan and
  the delays seen so far persist).
* Admission by HBM footprint: a pattern's factors stay resident while the
  sum of resident footprints (the pool's peak bytes) fits the budget; else
  the least recently used pattern is evicted (its last values are kept on the
  host, and it is refactored when next needed).
* Continuous batching: a worker thread takes every request that has arrived,
  keeps each pattern's requests in arrival order, and merges each run of
  consecutive solves into one multi-RHS solve; requests that arrive while
  the GPU works join the next round. A multi-RHS solve gives each column
  bitwise what a solve of that column alone gives, so a result does not
  depend on what it was batched with.
* Latency (submission to result) is recorded per request kind for p50/p99.
"""
import queue
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import ssids


@dataclass
class _Pattern:
    n: int
    ptr: np.ndarray
    row: np.ndarray
    akeep: ssids.AKeep
    values: np.ndarray = None
    fkeep: ssids.FKeep = None
    footprint: int = 0
    last_use: float = 0.0


@dataclass
class _Request:
    kind: str            # "refactor" or "solve"
    pattern: str
    data: np.ndarray
    future: Future
    t_submit: float = field(default_factory=time.perf_counter)


class SolverService:
    def __init__(self, hbm_budget_bytes=200 << 30, options=None, refine=True, tol=1e-14):
        self.options = options or ssids.SSIDSOptions()
        self.budget = int(hbm_budget_bytes)
        self.refine, self.tol = refine, tol
        self.cache = ssids.AnalyseCache()
        self.patterns = {}
        self.q = queue.Queue()
        self.lat = {"refactor": [], "solve": []}
        self.rounds = 0
        self._stop = threading.Event()
        self._thread = None

    # ---- client side
    def register(self, name, n, ptr, row, grid=None):
        ak, inf = self.cache.analyse(n, ptr, row, self.options, grid=grid)
        if inf.flag < 0:
            raise RuntimeError(f"analyse of {name} failed ({inf.flag})")
        self.patterns[name] = _Pattern(n=n, ptr=ptr, row=row, akeep=ak)

    def submit_refactor(self, name, values):
        return self._submit("refactor", name, np.array(values, np.float64, copy=True))

    def submit_solve(self, name, rhs):
        return self._submit("solve", name, np.array(rhs, np.float64, copy=True))

    def _submit(self, kind, name, data):
        if name not in self.patterns:
            raise KeyError(name)
        f = Future()
        self.q.put(_Request(kind, name, data, f))
        return f

    def start(self):
        self._thread = threading.Thread(target=self._run, daemon=True)
        self._thread.start()

    def stop(self):
        self._stop.set()
        self.q.put(None)
        if self._thread:
            self._thread.join()
        for p in self.patterns.values():
            if p.fkeep is not None:
                ssids.ssids_free(None, p.fkeep)
                p.fkeep = None
        self.cache.free()

    def latency_ms(self, kind):
        a = np.array(self.lat[kind]) * 1e3
        if a.size == 0:
            return None
        return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99)),
                "count": int(a.size)}

    # ---- worker
    def _run(self):
        while not self._stop.is_set():
            first = self.q.get()
            if first is None:
                break
            batch = [first]
            while True:  # everything that has arrived joins this round
                try:
                    r = self.q.get_nowait()
                except queue.Empty:
                    break
                if r is None:
                    self._stop.set()
                    break
                batch.append(r)
            self.rounds += 1
            by_pattern = {}
            for r in batch:
                by_pattern.setdefault(r.pattern, []).append(r)
            for name, reqs in by_pattern.items():
                self._serve_pattern(name, reqs)

    def _ensure_resident(self, p):
        if p.fkeep is not None and p.fkeep.numeric:
            return
        # admit: evict least recently used patterns until the footprint fits
        resident = sorted((q for q in self.patterns.values()
                           if q is not p and q.fkeep is not None), key=lambda q: q.last_use)
        used = sum(q.footprint for q in resident)
        while resident and used + p.footprint > self.budget:
            q = resident.pop(0)
            ssids.ssids_free(None, q.fkeep)
            q.fkeep = None
            used -= q.footprint
        self._factor(p)

    def _factor(self, p):
        fk, inf = ssids.ssids_factor(False, p.values, p.akeep, self.options, fkeep=p.fkeep)
        p.fkeep = fk
        return inf

    def _serve_pattern(self, name, reqs):
        p = self.patterns[name]
        i = 0
        while i < len(reqs):
            r = reqs[i]
            if r.kind == "refactor":
                p.values = r.data
                try:
                    if p.fkeep is None:
                        self._ensure_resident(p)
                        inf = p.fkeep.inform
                    else:
                        inf = self._factor(p)
                    p.footprint = max(p.footprint, self._footprint(p))
                    p.last_use = time.perf_counter()
                    r.future.set_result(inf)
                except Exception as e:  # the request fails, the service goes on
                    r.future.set_exception(e)
                self.lat["refactor"].append(time.perf_counter() - r.t_submit)
                i += 1
                continue
            # a run of solves: one multi-RHS call
            j = i
            while j < len(reqs) and reqs[j].kind == "solve":
                j += 1
            run = reqs[i:j]
            try:
                if p.values is None:
                    raise RuntimeError(f"{name}: solve before any refactor")
                self._ensure_resident(p)
                cols = [r.data.reshape(p.n, -1) for r in run]
                x = np.concatenate(cols, axis=1)
                if self.refine:
                    inf, bwd, its = ssids.ssids_solve_refine(x, p.akeep, p.fkeep, self.options,
                                                             tol=self.tol)
                else:
                    inf = ssids.ssids_solve(x, p.akeep, p.fkeep, self.options)
                if inf.flag < 0:
                    raise RuntimeError(f"{name}: solve failed ({inf.flag})")
                p.last_use = time.perf_counter()
                c = 0
                for r in run:
                    k = r.data.reshape(p.n, -1).shape[1]
                    out = x[:, c:c + k]
                    r.future.set_result(out.reshape(r.data.shape).copy())
                    c += k
            except Exception as e:
                for r in run:
                    r.future.set_exception(e)
            now = time.perf_counter()
            for r in run:
                self.lat["solve"].append(now - r.t_submit)
            i = j

    @staticmethod
    def _footprint(p):
        """HBM held by a resident pattern: the factorization pool's peak
        (factors, contribution blocks, workspaces) plus the solve workspace
        for the largest right-hand-side count served (64)."""
        peak = int(p.fkeep.inform.pool_peak_bytes)
        # per column: every front's rows (the row lists, twice for the
        # backward partial sums) and four vectors of n
        solve_ws = 8 * 64 * (2 * len(p.akeep.rlist) + 4 * int(p.akeep.n))
        return max(peak, 8 * int(p.akeep.num_factor)) + solve_ws
