"""A resident solver service: refactor and solve requests on patterns whose
analyse and factors stay in HBM, with refactors and solves overlapping on
the device.

* Each pattern is analysed once (ssids.AnalyseCache; the matching-based
  order takes the values given at registration), and its factorization
  handles are reused by every refactor of new values (their pools, static
  plans and the delays seen so far persist).
* Versions: every refactor submitted for a pattern starts a new version of
  its factors; a solve is answered with the version current when it was
  submitted.  So a result does not depend on what else is in flight.
* Two lanes, each a host thread driving its own HIP streams (every
  factorization handle owns its streams): the factor lane runs refactors,
  the solve lane runs solves.  A pattern holds up to two factorization
  handles, so a refactor builds version v+1 in one while solves of version v
  still run on the other; refactors of one pattern and solves of another
  always overlap.  A solve whose version is still being factorized waits,
  and the solve lane serves other work meanwhile.
* Continuous batching: the solve lane takes every solve that has arrived and
  merges those of one pattern and version into one multi-RHS solve; a
  multi-RHS solve gives each column bitwise what a solve of that column
  alone gives.
* Admission by HBM footprint: a new factorization is admitted while the sum
  of resident footprints fits the budget; otherwise the least recently used
  idle patterns are evicted (their latest values stay on the host, and they
  are refactorized when next needed).
* Latency (submission to result) is recorded per request kind for p50/p99.
"""
import threading
import time
from collections import Counter, deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from . import ssids


@dataclass
class _Pattern:
    name: str
    n: int
    akeep: ssids.AKeep
    version: int = 0              # versions submitted (refactors)
    values: np.ndarray = None     # values of the latest version (host)
    ready: dict = field(default_factory=dict)     # version -> resident FKeep
    failed: dict = field(default_factory=dict)    # version -> why its refactor failed
    spare: list = field(default_factory=list)     # idle handles kept for reuse
    pending: Counter = field(default_factory=Counter)  # version -> solves not done
    queued_refactors: int = 0
    footprint: int = 0            # bytes per resident handle
    last_use: float = 0.0


@dataclass
class _Request:
    kind: str            # "refactor", "solve" or "restore" (internal)
    pattern: str
    version: int
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
        self.lat = {"refactor": [], "solve": []}
        self.rounds = 0
        self._cv = threading.Condition()
        self._fq = deque()       # factor lane queue
        self._sq = deque()       # solve lane queue (arrived, not yet taken)
        self._waiting = []       # solves whose version is not resident yet
        self._stop = False
        self._threads = []

    # ---- client side
    def register(self, name, n, ptr, row, grid=None, val=None):
        ak, inf = self.cache.analyse(n, ptr, row, self.options, grid=grid, val=val)
        if inf.flag < 0:
            raise RuntimeError(f"analyse of {name} failed ({inf.flag})")
        self.patterns[name] = _Pattern(name=name, n=n, akeep=ak)

    def submit_refactor(self, name, values):
        values = np.array(values, np.float64, copy=True)
        with self._cv:
            p = self._pattern(name)
            p.version += 1
            p.values = values
            p.queued_refactors += 1
            r = _Request("refactor", name, p.version, values, Future())
            self._fq.append(r)
            self._cv.notify_all()
        return r.future

    def submit_solve(self, name, rhs):
        rhs = np.array(rhs, np.float64, copy=True)
        with self._cv:
            p = self._pattern(name)
            if p.version == 0:
                raise RuntimeError(f"{name}: solve before any refactor")
            p.pending[p.version] += 1
            r = _Request("solve", name, p.version, rhs, Future())
            self._sq.append(r)
            self._cv.notify_all()
        return r.future

    def _pattern(self, name):
        if name not in self.patterns:
            raise KeyError(name)
        return self.patterns[name]

    def start(self):
        self._threads = [threading.Thread(target=self._factor_lane, daemon=True),
                         threading.Thread(target=self._solve_lane, daemon=True)]
        for t in self._threads:
            t.start()

    def stop(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        for p in self.patterns.values():
            for fk in list(p.ready.values()) + p.spare:
                ssids.ssids_free(None, fk)
            p.ready.clear()
            p.spare.clear()
        self.cache.free()

    def latency_ms(self, kind):
        a = np.array(self.lat[kind]) * 1e3
        if a.size == 0:
            return None
        return {"p50": float(np.percentile(a, 50)), "p99": float(np.percentile(a, 99)),
                "count": int(a.size)}

    # ---- factor lane
    def _factor_lane(self):
        while True:
            with self._cv:
                while not self._fq and not self._stop:
                    self._cv.wait()
                if self._stop and not self._fq:
                    return
                r = self._fq.popleft()
                p = self.patterns[r.pattern]
                # a handle to build into: a spare one, else the oldest
                # resident version once none of its solves is outstanding
                # (at most two handles per pattern)
                while True:
                    fk = self._take_handle(p)
                    if fk is not False:
                        break
                    self._cv.wait()
                if fk is None:
                    self._admit(p)
            try:
                fk, inf = ssids.ssids_factor(False, r.data, p.akeep, self.options, fkeep=fk)
                ok = inf.flag >= 0 and fk.numeric
            except Exception as e:  # the request fails, the service goes on
                inf, ok = e, False
            with self._cv:
                if ok:
                    p.ready[r.version] = fk
                    p.footprint = max(p.footprint, self._footprint(p, fk))
                    # older versions nobody will solve with again become spare
                    for v in [v for v in p.ready if v < r.version and p.pending[v] == 0]:
                        p.spare.append(p.ready.pop(v))
                else:
                    p.failed[r.version] = (inf if isinstance(inf, Exception) else
                                           RuntimeError(f"{p.name}: refactor failed "
                                                        f"(flag {inf.flag}, HIP error "
                                                        f"{inf.cuda_error})"))
                    if fk is not None and getattr(fk, "numeric", None):
                        p.spare.append(fk)
                p.last_use = time.perf_counter()
                if r.kind == "refactor":
                    p.queued_refactors -= 1
                self._cv.notify_all()
            if r.kind == "refactor":
                if isinstance(inf, Exception):
                    r.future.set_exception(inf)
                else:
                    r.future.set_result(inf)
                self.lat["refactor"].append(time.perf_counter() - r.t_submit)

    def _take_handle(self, p):
        """A handle for the next factorization of p (None: a new one; False:
        wait for a resident version's solves to finish)."""
        if p.spare:
            return p.spare.pop()
        if len(p.ready) < 2:
            return None
        oldest = min(p.ready)
        if p.pending[oldest] == 0:
            return p.ready.pop(oldest)
        return False

    def _admit(self, p):
        """Evict least recently used idle patterns until one more handle of p
        fits the budget."""
        def used():
            return sum(q.footprint * (len(q.ready) + len(q.spare))
                       for q in self.patterns.values())
        idle = sorted((q for q in self.patterns.values()
                       if q is not p and not q.queued_refactors and not sum(q.pending.values())
                       and (q.ready or q.spare)), key=lambda q: q.last_use)
        while idle and used() + p.footprint > self.budget:
            q = idle.pop(0)
            for fk in list(q.ready.values()) + q.spare:
                ssids.ssids_free(None, fk)
            q.ready.clear()
            q.spare.clear()

    # ---- solve lane
    def _solve_lane(self):
        while True:
            with self._cv:
                while not self._sq and not self._stop and not self._servable():
                    self._cv.wait()
                if self._stop and not self._sq and not self._waiting:
                    return
                self._waiting.extend(self._sq)
                self._sq.clear()
                groups, keep = {}, []
                for r in self._waiting:
                    p = self.patterns[r.pattern]
                    if r.version in p.failed:
                        p.pending[r.version] -= 1
                        r.future.set_exception(p.failed[r.version])
                        self.lat["solve"].append(time.perf_counter() - r.t_submit)
                        continue
                    if r.version in p.ready:
                        groups.setdefault((r.pattern, r.version), []).append(r)
                        continue
                    keep.append(r)
                    # an evicted version (the latest one) is factorized again
                    if (r.version == p.version and not p.queued_refactors
                            and not any(q.kind == "restore" and q.pattern == r.pattern
                                        for q in self._fq)):
                        self._fq.append(_Request("restore", r.pattern, r.version,
                                                 p.values, Future()))
                        self._cv.notify_all()
                self._waiting = keep
                if not groups:
                    if self._stop and not self._fq:
                        for r in keep:
                            r.future.set_exception(RuntimeError("service stopped"))
                        return
                    self._cv.wait(0.05)
                    continue
                self.rounds += 1
                work = [(self.patterns[k[0]], k[1], rs) for k, rs in groups.items()]
            for p, v, rs in work:
                self._solve_group(p, v, rs)

    def _servable(self):
        return any(r.version in self.patterns[r.pattern].ready or
                   r.version in self.patterns[r.pattern].failed for r in self._waiting)

    def _solve_group(self, p, v, run):
        fk = p.ready[v]  # stays resident: its pending count is above zero
        try:
            cols = [r.data.reshape(p.n, -1) for r in run]
            x = np.concatenate(cols, axis=1)
            if self.refine:
                inf, bwd, its = ssids.ssids_solve_refine(x, p.akeep, fk, self.options,
                                                         tol=self.tol)
            else:
                inf = ssids.ssids_solve(x, p.akeep, fk, self.options)
            if inf.flag < 0:
                raise RuntimeError(f"{p.name}: solve failed ({inf.flag})")
            c = 0
            for r in run:
                k = r.data.reshape(p.n, -1).shape[1]
                r.future.set_result(x[:, c:c + k].reshape(r.data.shape).copy())
                c += k
        except Exception as e:
            for r in run:
                r.future.set_exception(e)
        now = time.perf_counter()
        for r in run:
            self.lat["solve"].append(now - r.t_submit)
        with self._cv:
            p.pending[v] -= len(run)
            p.last_use = now
            if p.pending[v] == 0 and v < p.version and v in p.ready and \
                    any(w > v for w in p.ready):
                p.spare.append(p.ready.pop(v))  # superseded and idle
            self._cv.notify_all()

    @staticmethod
    def _footprint(p, fk):
        """HBM held by one resident handle: the factorization pool's peak
        (factors, contribution blocks, workspaces) plus the solve workspace
        for the largest right-hand-side count served (64)."""
        peak = int(fk.inform.pool_peak_bytes)
        solve_ws = 8 * 64 * (2 * len(p.akeep.rlist) + 4 * int(p.akeep.n))
        return max(peak, 8 * int(p.akeep.num_factor)) + solve_ws
