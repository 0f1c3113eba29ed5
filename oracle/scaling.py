"""CPU restatement of the reference's symmetric scalings -- TEST
INFRASTRUCTURE ONLY: imported by tests/, never by the product.

equilib_scale_sym restates inf_norm_equilib_sym (src/scaling.f90:480-521),
the routine behind equilib_scale_sym (src/scaling.f90:403-416) and
ssids_factor's options%scaling >= 4 (src/ssids/ssids.f90:1003-1028).  Every
operation is exactly rounded IEEE arithmetic (products in the reference's
operand order, max, sqrt, division), so the restatement reproduces the
reference bit for bit.  Pinned by tests/test_scaling.py against the
reference routine itself (oracle/_ref/libref_scaling.so, built from
src/scaling.f90 where it lies), against closed forms (a diagonal matrix is
equilibrated in one step to 1/sqrt|a_ii|; a 2x2 matrix has a closed-form
fixed point) and against the reference test's property (every scaled row's
infinity norm within 0.05 of 1, tests/scaling.f90:455-470).

ref_equilib_scale_sym / ref_auction_scale_sym / ref_hungarian_scale_sym call
the reference routines
through the shim oracle/ref_scaling_shim.f90 when oracle/_ref is built.
"""
import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_LIB = HERE / "_ref" / "libref_scaling.so"


def equilib_scale_sym(n, ptr, row, val, max_iterations=10, tol=np.float32(1e-8)):
    """Returns (scaling, iterations) for the lower CSC (1-based ptr, row).

    Follows src/scaling.f90:500-520: s = 1; per iteration the maximum of
    |(s_r a) s_c| over every stored entry touching row/column i, then
    s_i /= sqrt(m_i) where m_i > 0, then the exit test max|1 - m| < tol
    (tol a default-kind real, compared in double precision)."""
    ptr = np.asarray(ptr, np.int64)
    row = np.asarray(row, np.int64) - 1
    val = np.asarray(val, np.float64)
    nz = int(ptr[n] - 1)
    col = np.repeat(np.arange(n), np.diff(ptr[:n + 1]))
    row, val = row[:nz], val[:nz]
    s = np.ones(n)
    tolv = float(np.float32(tol))
    itr = 1
    for itr in range(1, max_iterations + 1):
        v = np.abs((s[row] * val) * s[col])
        m = np.zeros(n)
        np.maximum.at(m, row, v)
        np.maximum.at(m, col, v)
        pos = m > 0
        s[pos] = s[pos] / np.sqrt(m[pos])
        if n == 0 or np.max(np.abs(1.0 - m)) < tolv:
            return s, itr - 1
    return s, (max_iterations if max_iterations > 0 else 0)


def ref_available():
    return REF_LIB.exists()


_ref = None


def _reflib():
    global _ref
    if _ref is None:
        L = C.CDLL(str(REF_LIB))
        vp = C.c_void_p
        L.ref_equilib_scale_sym.argtypes = [C.c_int, vp, vp, vp, vp, C.c_int, C.c_float, vp, vp]
        L.ref_auction_scale_sym.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp]
        L.ref_hungarian_scale_sym.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_int, vp, vp]
        _ref = L
    return _ref


def ref_equilib_scale_sym(n, ptr, row, val, max_iterations=10, tol=np.float32(1e-8)):
    """The reference's equilib_scale_sym itself: (scaling, iterations, flag)."""
    L = _reflib()
    ptr = np.ascontiguousarray(ptr, np.int64)
    row = np.ascontiguousarray(row, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    s = np.zeros(max(n, 1))
    it, fl = C.c_int(0), C.c_int(0)
    L.ref_equilib_scale_sym(n, ptr.ctypes.data, row.ctypes.data, val.ctypes.data,
                            s.ctypes.data, max_iterations, float(np.float32(tol)),
                            C.byref(it), C.byref(fl))
    return s[:n], it.value, fl.value


def ref_hungarian_scale_sym(n, ptr, row, val, scale_if_singular=False):
    """The reference's hungarian_scale_sym itself: (scaling, match, matched,
    flag); match 1-based, 0 unmatched."""
    L = _reflib()
    ptr = np.ascontiguousarray(ptr, np.int64)
    row = np.ascontiguousarray(row, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    s = np.zeros(max(n, 1))
    mt = np.zeros(max(n, 1), np.int32)
    nm, fl = C.c_int(0), C.c_int(0)
    L.ref_hungarian_scale_sym(n, ptr.ctypes.data, row.ctypes.data, val.ctypes.data,
                              s.ctypes.data, mt.ctypes.data, 1 if scale_if_singular else 0,
                              C.byref(nm), C.byref(fl))
    return s[:n], mt[:n], nm.value, fl.value


def ref_auction_scale_sym(n, ptr, row, val):
    """The reference's auction_scale_sym itself with default auction_options:
    (scaling, match, matched, iterations, flag); match 1-based, 0 unmatched."""
    L = _reflib()
    ptr = np.ascontiguousarray(ptr, np.int64)
    row = np.ascontiguousarray(row, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    s = np.zeros(max(n, 1))
    mt = np.zeros(max(n, 1), np.int32)
    nm, it, fl = C.c_int(0), C.c_int(0), C.c_int(0)
    L.ref_auction_scale_sym(n, ptr.ctypes.data, row.ctypes.data, val.ctypes.data,
                            s.ctypes.data, mt.ctypes.data, C.byref(nm), C.byref(it),
                            C.byref(fl))
    return s[:n], mt[:n], nm.value, it.value, fl.value
