"""Many independent sparse symmetric systems in one factor + solve call.

The systems are put together as one block-diagonal matrix. Each block keeps
its own fill-reducing order (exact minimum degree, or approximate minimum
degree for large blocks, or the matching-based order with its scaling), and
the blocks are contiguous in the elimination order. The assembly tree is then a forest with one tree per system, and every
level's launches cover the fronts of all systems at once: one launch sequence
per tree level for the whole batch, instead of one per level per system.

A front's arithmetic depends only on its own data and its children's, so
each system's factors and solution are bitwise those of factoring it alone
with the same order. Per-system inertia is read from D^{-1} in pivot order
(ssids_enquire_indef), with the counting rule of the reference's
collect_stats (src/ssids/gpu/factor.f90:1156-1198): a 1x1 pivot is negative
if d < 0; a 2x2 block with det < 0 has one negative eigenvalue, else two if
its trace is negative; a zero pivot lowers the rank.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _lib, ssids


@dataclass
class Batch:
    sizes: np.ndarray            # n of each system
    offsets: np.ndarray          # first variable of each system (len = count + 1)
    akeep: ssids.AKeep = None
    fkeep: ssids.FKeep = None
    options: ssids.SSIDSOptions = None
    orders: list = field(default_factory=list)   # per-system order (1-based)
    scale: np.ndarray = None     # matching scalings of ordering 2 + scaling 3, concatenated


def block_diagonal(mats):
    """Lower CSC (1-based) of the block-diagonal matrix of LowerCSC systems."""
    sizes = np.array([m.n for m in mats], np.int64)
    offsets = np.zeros(len(mats) + 1, np.int64)
    offsets[1:] = np.cumsum(sizes)
    ptrs, rows, vals = [np.ones(1, np.int64)], [], []
    base = 0
    for m, off in zip(mats, offsets[:-1]):
        p = np.asarray(m.ptr, np.int64)
        ptrs.append(p[1:] - 1 + base + 1)
        rows.append(np.asarray(m.row, np.int64) + off)
        vals.append(np.asarray(m.val, np.float64))
        base += int(p[-1] - 1)
    ptr = np.concatenate(ptrs)
    row = np.concatenate(rows).astype(np.int32) if rows else np.zeros(0, np.int32)
    val = np.concatenate(vals) if vals else np.zeros(0)
    return int(offsets[-1]), ptr, row, val, sizes, offsets


def system_order(m, options):
    """Fill-reducing order of one system, as ssids_analyse would pick it
    (ordering 2: the matching-based order, returned with its scaling)."""
    L = _lib.lib()
    o = np.zeros(m.n, np.int32)
    ptr = np.ascontiguousarray(m.ptr, np.int64)
    row = np.ascontiguousarray(m.row, np.int32)
    if options.ordering == 2:
        sc = np.zeros(m.n)
        L.spral_match_order(m.n, ptr, row, np.ascontiguousarray(m.val, np.float64), o, sc)
        return o, sc
    if options.ordering == 4 or m.n > ssids.AMD_THRESHOLD:
        L.spral_ssids_hip_order_amd(m.n, ptr, row, o)
    else:
        L.spral_ssids_hip_order_min_degree(m.n, ptr, row, o)
    return o, None


def batch_analyse(mats, options=None, autotune=False):
    """One analyse of the block-diagonal pattern; returns (Batch, inform).
    autotune: pick the forest's tree-level schedule by timing each on the
    device with the systems' values (ssids.ssids_autotune); the choice and
    the times are in batch.akeep.level_schedule and .schedule_ms."""
    options = options or ssids.SSIDSOptions()
    n, ptr, row, val, sizes, offsets = block_diagonal(mats)
    res = [system_order(m, options) for m in mats]
    orders = [o for o, _ in res]
    order = np.concatenate([o + int(off) for o, off in zip(orders, offsets[:-1])]) \
        if orders else np.zeros(0, np.int32)
    scale = None
    over = {"ordering": 0}
    if options.ordering == 2 and options.scaling == 3:
        # each system's matching scaling, passed as a user scaling of the whole
        scale = np.concatenate([sc for _, sc in res]) if res else np.zeros(0)
        over["scaling"] = 0
    opts = ssids.SSIDSOptions(**{**options.__dict__, **over})
    ak, inf = ssids.ssids_analyse(False, n, ptr, row, opts, order=order)
    if autotune and inf.flag >= 0:
        ssids.ssids_autotune(ak, val, opts, scale=scale)
    return Batch(sizes=sizes, offsets=offsets, akeep=ak, options=opts,
                 orders=orders, scale=scale), inf


def batch_factor(batch, mats_or_vals, posdef=False):
    """Factor every system in one call. mats_or_vals: the systems (values in
    the analysed pattern) or their values concatenated. Returns (inform of
    the whole batch, per-system num_neg, per-system rank)."""
    if isinstance(mats_or_vals, np.ndarray):
        val = mats_or_vals
    else:
        val = np.concatenate([np.asarray(m.val, np.float64) for m in mats_or_vals])
    fk, inf = ssids.ssids_factor(posdef, val, batch.akeep, batch.options, fkeep=batch.fkeep,
                                 scale=None if batch.scale is None else batch.scale.copy())
    batch.fkeep = fk
    if inf.flag < 0:
        return inf, None, None
    if posdef:
        return inf, np.zeros(len(batch.sizes), np.int64), batch.sizes.copy()
    neg, rank = system_inertia(batch)
    return inf, neg, rank


def system_inertia(batch):
    """Per-system (num_neg, rank) from D^{-1} in pivot order."""
    piv, d, inf = ssids.ssids_enquire_indef(batch.akeep, batch.fkeep)
    n = batch.akeep.n
    pos = np.abs(piv) - 1                       # pivot position of each variable
    var_at = np.empty(n, np.int64)
    var_at[pos] = np.arange(n)
    sysid = np.searchsorted(batch.offsets, var_at, side="right") - 1
    two = piv[var_at] < 0                       # 2x2 members, by position
    # pair up consecutive 2x2 members: first members sit at even offsets of a run
    start = two & ~np.concatenate([[False], two[:-1]])
    run_start = np.maximum.accumulate(np.where(start, np.arange(n), 0))
    first = two & ((np.arange(n) - run_start) % 2 == 0)
    neg_pos = np.zeros(n, np.int64)
    zero_pos = np.zeros(n, np.int64)
    one = ~two
    neg_pos[one] = d[0, one] < 0
    zero_pos[one] = d[0, one] == 0
    k = np.nonzero(first)[0]
    a, b, c = d[0, k], d[1, k], d[0, k + 1]
    det = a * c - b * b
    neg_pos[k] = np.where(det < 0, 1, np.where(a + c < 0, 2, 0))
    count = len(batch.sizes)
    neg = np.bincount(sysid, weights=neg_pos, minlength=count).astype(np.int64)
    zero = np.bincount(sysid, weights=zero_pos, minlength=count).astype(np.int64)
    return neg, batch.sizes - zero


def batch_solve(batch, rhs, refine=False, tol=1e-14, maxit=5):
    """Solve every system; rhs is a list of per-system right-hand sides (n_b,)
    or (n_b, k) with the same k, or their concatenation. Returns the
    solutions in the same form."""
    as_list = isinstance(rhs, (list, tuple))
    x = np.concatenate([np.asarray(r, np.float64) for r in rhs]) if as_list \
        else np.array(rhs, np.float64, copy=True)
    if refine:
        inf, bwd, its = ssids.ssids_solve_refine(x, batch.akeep, batch.fkeep, batch.options,
                                                 maxit=maxit, tol=tol)
    else:
        inf = ssids.ssids_solve(x, batch.akeep, batch.fkeep, batch.options)
    if inf.flag < 0:
        raise RuntimeError(f"batch solve failed ({inf.flag})")
    if not as_list:
        return x
    return [x[batch.offsets[i]:batch.offsets[i + 1]] for i in range(len(batch.sizes))]


def batch_free(batch):
    ssids.ssids_free(batch.akeep, batch.fkeep)
    batch.akeep = batch.fkeep = None
