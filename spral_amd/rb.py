"""Rutherford-Boeing files (rb_peek / rb_write / rb_read of
src/rutherford_boeing.f90) over the library's host C code (include/spral_rb.h).
Indices are 1-based, as the reference's."""
import ctypes as C

import numpy as np

from . import _lib

TRI_LWR, TRI_UPR, TRI_FULL = 1, 2, 3
VALUES_FILE, VALUES_PATTERN = 0, 1


def rb_peek(filename):
    """Header fields: dict(info, m, n, nelt, nvar, nval, matrix_type,
    type_code, title, identifier)."""
    L = _lib.lib()
    m, n, mt = C.c_int(), C.c_int(), C.c_int()
    nelt, nvar, nval = C.c_int64(), C.c_int64(), C.c_int64()
    tc, title, ident = C.create_string_buffer(3), C.create_string_buffer(72), \
        C.create_string_buffer(8)
    info = L.spral_rb_peek(str(filename).encode(), C.byref(m), C.byref(n), C.byref(nelt),
                           C.byref(nvar), C.byref(nval), C.byref(mt), tc, title, ident)
    return dict(info=info, m=m.value, n=n.value, nelt=nelt.value, nvar=nvar.value,
                nval=nval.value, matrix_type=mt.value, type_code=tc.raw.decode(),
                title=title.raw.decode(), identifier=ident.raw.decode())


def rb_write(filename, matrix_type, m, n, ptr, row, val=None, val_format=None, title=None,
             identifier=None):
    """Writes an assembled matrix (val None: pattern only); returns info."""
    L = _lib.lib()
    ptr = np.ascontiguousarray(ptr, np.int64)
    row = np.ascontiguousarray(row, np.int32)
    vp = None
    if val is not None:
        val = np.ascontiguousarray(val, np.float64)
        vp = val.ctypes.data
    enc = (lambda s: None if s is None else s.encode())
    return L.spral_rb_write(str(filename).encode(), matrix_type, m, n, ptr.ctypes.data,
                            row.ctypes.data, vp, enc(val_format), enc(title), enc(identifier))


def rb_read(filename, lwr_upr_full=TRI_LWR, values=VALUES_FILE, add_diagonal=False):
    """Returns (info, m, n, ptr, row, val or None, matrix_type)."""
    L = _lib.lib()
    m, n, hv, mt, info = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
    nz = C.c_int64()
    h = L.spral_rb_read(str(filename).encode(), lwr_upr_full, values, int(add_diagonal),
                        C.byref(m), C.byref(n), C.byref(nz), C.byref(hv), C.byref(mt),
                        C.byref(info))
    if not h:
        return info.value, 0, 0, None, None, None, 0
    ptr = np.zeros(n.value + 1, np.int64)
    row = np.zeros(max(nz.value, 1), np.int32)
    val = np.zeros(max(nz.value, 1)) if hv.value else None
    L.spral_rb_fetch(h, ptr.ctypes.data, row.ctypes.data,
                     val.ctypes.data if val is not None else None)
    return (info.value, m.value, n.value, ptr, row[:nz.value],
            None if val is None else val[:nz.value], mt.value)
