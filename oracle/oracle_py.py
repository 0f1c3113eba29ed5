"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST
INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product."""
import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError("oracle not built: make -C oracle")
        L = C.CDLL(str(path))
        vp = C.c_void_p
        L.oracle_factor.restype = vp
        L.oracle_factor.argtypes = [C.c_int, C.c_int] + [vp] * 8 + [C.c_int, C.c_double,
                                                                    C.c_double, vp]
        L.oracle_factor_prefix.restype = vp
        L.oracle_factor_prefix.argtypes = ([C.c_int, C.c_int] + [vp] * 8 +
                                           [C.c_int, C.c_double, C.c_double, C.c_int, vp])
        L.oracle_factor_app.restype = vp
        L.oracle_factor_app.argtypes = L.oracle_factor.argtypes
        L.oracle_node_app.argtypes = [vp, vp]
        L.oracle_solve.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int]
        L.oracle_enquire.argtypes = [vp, vp, vp]
        L.oracle_node_info.argtypes = [vp, vp, vp]
        L.oracle_free.argtypes = [vp]
        L.oracle_node_factors.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleFactor:
    """Factorize with the oracle given the symbolic arrays of an analyse."""

    # u defaults to the reference's value: the single-precision literal 0.01
    # widened to double (src/ssids/datatypes.f90:262)
    def __init__(self, akeep, val, scaling=None, action=True, u=float(np.float32(0.01)),
                 small=1e-20, max_nodes=None, app=False):
        """max_nodes: factor only the first max_nodes fronts in postorder (a
        prefix of the same factorization, for a bounded timing sample; such a
        handle takes no solves).  app: the device's pivoting (a posteriori
        pivoting in 64-column blocks, then TPP on the failed columns) and
        front layout (delayed columns first) instead of TPP throughout."""
        L = lib()
        self.akeep = akeep
        self.val = np.ascontiguousarray(val, np.float64)
        self.scaling = None if scaling is None else np.ascontiguousarray(
            np.asarray(scaling, np.float64)[akeep.invp])
        st = np.zeros(10, np.int64)
        args = (akeep.n, akeep.nnodes, _p(akeep.sptr), _p(akeep.sparent), _p(akeep.rptr),
                _p(akeep.rlist), _p(akeep.nptr), _p(akeep.nlist), _p(self.val),
                _p(self.scaling), 1 if action else 0, u, small)
        if app:
            assert max_nodes is None
            self.h = L.oracle_factor_app(*args, _p(st))
        elif max_nodes is None:
            self.h = L.oracle_factor(*args, _p(st))
        else:
            self.h = L.oracle_factor_prefix(*args, int(max_nodes), _p(st))
        keys = ["flag", "num_delay", "num_neg", "num_two", "num_zero", "maxfront",
                "not_first_pass", "num_factor", "num_flops", "matrix_rank"]
        self.stats = {k: int(v) for k, v in zip(keys, st)}

    def solve(self, b, job=0):
        """Solve with b in the user's variable order (permutation applied here
        exactly as src/ssids/fkeep.F90:252-315)."""
        ak = self.akeep
        B = np.asarray(b, np.float64).reshape(ak.n, -1)
        x2 = np.asfortranarray(B[ak.invp, :])
        if self.scaling is not None and job in (0, 1):
            x2 *= self.scaling[:, None]
        lib().oracle_solve(self.h, job, x2.shape[1], _p(x2), ak.n)
        if self.scaling is not None and job in (0, 3, 4):
            x2 *= self.scaling[:, None]
        out = np.empty_like(x2)
        out[ak.invp, :] = x2
        return out.reshape(np.shape(b)) if np.ndim(b) == 1 else out

    def enquire(self):
        n = self.akeep.n
        po = np.zeros(n, np.int32)
        d = np.zeros(2 * n, np.float64)
        lib().oracle_enquire(self.h, _p(po), _p(d))
        piv = np.zeros(n, np.int32)
        piv[self.akeep.invp] = po
        piv = np.where(piv >= 0, piv + 1, piv - 1)
        return piv, d.reshape(n, 2).T.copy()

    def node_info(self):
        nn = self.akeep.nnodes
        ne = np.zeros(nn, np.int32)
        nd = np.zeros(nn, np.int32)
        lib().oracle_node_info(self.h, _p(ne), _p(nd))
        return ne, nd

    def node_app(self):
        ne = np.zeros(self.akeep.nnodes, np.int32)
        lib().oracle_node_app(self.h, _p(ne))
        return ne

    def node_factors(self, node):
        info = np.zeros(6, np.int32)
        lib().oracle_node_factors(self.h, node, _p(info), None, None, None)
        m, n, nd, ne, ld, sn = (int(v) for v in info)
        Lm = np.zeros((n, ld))
        D = np.zeros(2 * n)
        perm = np.zeros(n, np.int32)
        lib().oracle_node_factors(self.h, node, _p(info), _p(Lm), _p(D), _p(perm))
        return dict(m=m, n=n, nd=nd, nelim=ne, sn=sn, L=Lm.T, D=D, perm=perm)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_free(self.h)
            self.h = None
