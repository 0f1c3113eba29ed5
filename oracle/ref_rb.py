"""The reference's Rutherford-Boeing routines (rb_write / rb_read of
src/rutherford_boeing.f90, compiled where they lie into
oracle/_ref/libref_rb.so behind oracle/ref_rb_shim.f90) -- TEST
INFRASTRUCTURE ONLY: imported by tests/, never by the product."""
import ctypes as C
from pathlib import Path

import numpy as np

REF_LIB = Path(__file__).resolve().parent / "_ref" / "libref_rb.so"
_ref = None


def available():
    return REF_LIB.exists()


def _lib():
    global _ref
    if _ref is None:
        L = C.CDLL(str(REF_LIB))
        vp = C.c_void_p
        L.ref_rb_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp,
                                   C.c_int, vp]
        L.ref_rb_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int] + [vp] * 6
        L.ref_rb_fetch.argtypes = [vp, vp, vp]
        _ref = L
    return _ref


def rb_write(filename, matrix_type, m, n, ptr, row, val=None):
    L = _lib()
    f = str(filename).encode()
    ptr = np.ascontiguousarray(ptr, np.int64)
    row = np.ascontiguousarray(row, np.int32)
    v = np.ascontiguousarray(val if val is not None else np.zeros(1), np.float64)
    info = C.c_int()
    L.ref_rb_write(f, len(f), matrix_type, m, n, ptr.ctypes.data, row.ctypes.data,
                   v.ctypes.data, int(val is not None), C.byref(info))
    return info.value


def rb_read(filename, lwr_upr_full=1, values=0, add_diagonal=False):
    L = _lib()
    f = str(filename).encode()
    m, n, hv, mt, info = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
    nz = C.c_int64()
    L.ref_rb_read(f, len(f), lwr_upr_full, values, int(add_diagonal), C.byref(m),
                  C.byref(n), C.byref(nz), C.byref(hv), C.byref(mt), C.byref(info))
    if info.value < 0:
        return info.value, 0, 0, None, None, None, 0
    ptr = np.zeros(n.value + 1, np.int64)
    row = np.zeros(max(nz.value, 1), np.int32)
    val = np.zeros(max(nz.value, 1))
    L.ref_rb_fetch(ptr.ctypes.data, row.ctypes.data, val.ctypes.data)
    return (info.value, m.value, n.value, ptr, row[:nz.value],
            val[:nz.value] if hv.value else None, mt.value)
