"""Tree levels of the factorization (spral_ssids_hip_tree_levels): every
front above its children, large fronts of unequal subtrees lined up on one
level, and the per-level step count of the 64^3 workload's tree."""
import numpy as np

from spral_amd import _lib, matrices, ssids


def tree_levels(parent, ncols, mode=0):
    parent = np.ascontiguousarray(parent, np.int32)
    ncols = np.ascontiguousarray(ncols, np.int32)
    lev = np.zeros(parent.size, np.int32)
    nl = _lib.lib().spral_ssids_hip_tree_levels(parent.size, parent.ctypes.data,
                                                ncols.ctypes.data, mode, lev.ctypes.data)
    return nl, lev


def steps(ncols):
    return (np.asarray(ncols) + 63) // 64


def height_levels(parent):
    lev = np.zeros(len(parent), int)
    for i, p in enumerate(parent):
        if p >= 0:
            lev[p] = max(lev[p], lev[i] + 1)
    return lev


def level_steps(lev, w):
    """APP steps in sequence: the most steps of each level's fronts, summed."""
    return sum(int(w[lev == L].max()) for L in range(lev.max() + 1) if (lev == L).any())


def check_valid(parent, lev, nl):
    assert nl == lev.max() + 1
    assert lev.min() >= 0
    for i, p in enumerate(parent):
        if p >= 0:
            assert lev[i] < lev[p], (i, p)


def test_unequal_chains_share_a_level(lib):
    # two 10-step fronts under a root: A above a small leaf and below one
    # small front, B a leaf below two small fronts.  By height from the
    # leaves A and B sit on different levels (31 steps in sequence); lined
    # up they share one (23)
    #   0 (32 cols) -> 1 A (640) -> 2 (32) -> 6 root (640)
    #   3 B (640) -> 4 (32) -> 5 (32) -> 6
    parent = [1, 2, 6, 4, 5, 6, -1]
    ncols = [32, 640, 32, 640, 32, 32, 640]
    nl, lev = tree_levels(parent, ncols)
    check_valid(parent, lev, nl)
    assert lev[1] == lev[3]
    w = steps(ncols)
    assert level_steps(height_levels(parent), w) == 10 + 10 + 1 + 10
    assert level_steps(lev, w) == 1 + 10 + 1 + 1 + 10


def test_forest_roots_line_up(lib):
    # independent trees of unequal height (a batch): each root has the most
    # steps of its tree, and all roots land on the top level
    parent, ncols = [], []
    for depth in (1, 3, 5, 2):
        base = len(parent)
        for d in range(depth):
            parent.append(base + d + 1)
            ncols.append(16)
        parent.append(-1)
        ncols.append(512)
    nl, lev = tree_levels(parent, ncols)
    check_valid(parent, lev, nl)
    roots = [i for i, p in enumerate(parent) if p < 0]
    assert len({int(lev[r]) for r in roots}) == 1
    assert level_steps(lev, steps(ncols)) < level_steps(height_levels(parent), steps(ncols))


def test_rejects_non_postorder(lib):
    nl, _ = tree_levels([0, -1], [1, 1])
    assert nl == -1
    nl, _ = tree_levels([2, 5, -1], [1, 1, 1])
    assert nl == -1


def test_grid64_tree(lib):
    """The 64^3 nested-dissection tree: 215 steps in sequence by height from
    the leaves (its two 2,000-column separators one level apart), 168 with
    the level assignment the analyse uses (the longest chain of fronts from a
    leaf to the root has 161)."""
    m = matrices.laplace3d(64, diag=5.7)
    opts = ssids.SSIDSOptions(ordering=3, use_gpu=False)
    ak, inf = ssids.ssids_analyse(False, m.n, m.ptr, m.row, opts, grid=(64, 64, 64))
    assert inf.flag >= 0
    ns = ak.nnodes
    sptr = np.asarray(ak.sptr) - 1
    parent = np.asarray(ak.sparent)[:ns] - 1
    parent[(parent < 0) | (parent >= ns)] = -1
    ncols = sptr[1:ns + 1] - sptr[:ns]
    nl, lev = tree_levels(parent, ncols)
    check_valid(parent, lev, nl)
    w = steps(ncols)
    assert level_steps(height_levels(parent), w) == 215
    assert level_steps(lev, w) == 168
    path = np.zeros(ns, int)
    for i in range(ns):
        path[i] += w[i]
        if parent[i] >= 0:
            path[parent[i]] = max(path[parent[i]], path[i])
    assert path.max() == 161


def test_forced_schedules(lib):
    """Modes 1-3 force height from the leaves, greedy and layered; mode 0
    keeps the cheapest of them (4 per step in sequence, 1 per level); other
    modes are rejected."""
    parent = [1, 2, 6, 4, 5, 6, -1]
    ncols = [32, 640, 32, 640, 32, 32, 640]
    w = steps(ncols)
    levs = {}
    for mode in (1, 2, 3):
        nl, lev = tree_levels(parent, ncols, mode)
        check_valid(parent, lev, nl)
        levs[mode] = lev
    assert (levs[1] == height_levels(parent)).all()
    cost = lambda lev: 4 * level_steps(lev, w) + int(lev.max()) + 1
    nl0, lev0 = tree_levels(parent, ncols, 0)
    assert cost(lev0) == min(cost(lv) for lv in levs.values()) < cost(levs[1])
    assert tree_levels(parent, ncols, 4)[0] == -1
    assert tree_levels(parent, ncols, -1)[0] == -1
