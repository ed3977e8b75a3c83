"""The reference's output and backup files (C ABI in include/mlau_io.h).

Host-only: these calls need the library but no GPU.  Fields are float64 arrays
of shape (7, nx) in 1D or (8, [nz,] ny, nx), argument order, ghosts included.
"""
import ctypes as C
import os

import numpy as np

from . import Grid, MLAUError, _dp, _ptrs, lib

F64, F32 = 0, 1
NO_BACKUP = 5
_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        i, d, ll, cp = C.c_int, C.c_double, C.c_longlong, C.c_char_p
        L.mlau_io_write_grid.argtypes = [cp, C.POINTER(Grid), _dp, _dp, _dp, i, i, i, i]
        L.mlau_io_write_output.argtypes = [cp, C.POINTER(Grid), i, i, d, C.POINTER(_dp), i, i]
        L.mlau_io_read_output.argtypes = [cp, C.POINTER(Grid), i, i, i, C.POINTER(_dp)]
        L.mlau_io_bkup_save.argtypes = [cp, ll, i, i, d, d, d, i, C.POINTER(_dp)]
        L.mlau_io_bkup_load.argtypes = [cp, ll, i, C.POINTER(i), C.POINTER(i), C.POINTER(d), C.POINTER(d),
                                        C.POINTER(d), C.POINTER(_dp)]
        _bound = True
    return L


def _check(rc):
    if rc != 0:
        raise MLAUError(lib().mlau_last_error().decode() or "error %d" % rc)


def grid_of(fields, xoff=4, gamma=5.0 / 3.0, bx=0.0):
    f = np.asarray(fields)
    g = Grid()
    g.dim = f.ndim - 1
    g.nx = f.shape[-1]
    g.ny = f.shape[-2] if g.dim >= 2 else 1
    g.nz = f.shape[-3] if g.dim == 3 else 1
    g.xoff = xoff
    g.yoff = xoff if g.dim >= 2 else 0
    g.zoff = xoff if g.dim == 3 else 0
    g.dx = g.dy = g.dz = 1.0
    g.gamma, g.bx = gamma, bx
    return g


def _rows(f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    return f, _ptrs([f[n] for n in range(f.shape[0])])


def write_grid(directory, g, x, y=None, z=None, rank=-1, ix=0, iy=0, iz=0):
    arr = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
    ptr = [None if v is None else v.ctypes.data_as(_dp) for v in arr]
    _check(_lib().mlau_io_write_grid(os.fsencode(directory), C.byref(g), *ptr, rank, ix, iy, iz))


def write_output(directory, g, first, cnt, tim, fields, fmt=F64, rank=-1):
    f, p = _rows(fields)
    _check(_lib().mlau_io_write_output(os.fsencode(directory), C.byref(g), int(bool(first)), cnt, tim, p, fmt, rank))


def read_output(directory, g, cnt, fmt=F64, rank=-1):
    shape = (7, g.nx) if g.dim == 1 else ((8, g.ny, g.nx) if g.dim == 2 else (8, g.nz, g.ny, g.nx))
    out = np.empty(shape)
    _check(_lib().mlau_io_read_output(os.fsencode(directory), C.byref(g), cnt, fmt, rank,
                                      _ptrs([out[n] for n in range(shape[0])])))
    return out


def bkup_save(directory, fields, n, cnt, tim, dt, trec, rank=0):
    f, p = _rows(fields)
    _check(_lib().mlau_io_bkup_save(os.fsencode(directory), f[0].size, n, cnt, tim, dt, trec, rank, p))


def bkup_load(directory, shape, rank=0):
    """(fields, dict(n, cnt, tim, dt, trec)) or None when there is no backup."""
    out = np.empty(shape)
    n, cnt = C.c_int(), C.c_int()
    tim, dt, trec = C.c_double(), C.c_double(), C.c_double()
    rc = _lib().mlau_io_bkup_load(os.fsencode(directory), out[0].size, rank, C.byref(n), C.byref(cnt), C.byref(tim),
                                  C.byref(dt), C.byref(trec), _ptrs([out[k] for k in range(shape[0])]))
    if rc == NO_BACKUP:
        return None
    _check(rc)
    return out, dict(n=n.value, cnt=cnt.value, tim=tim.value, dt=dt.value, trec=trec.value)
