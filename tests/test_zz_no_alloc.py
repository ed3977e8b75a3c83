"""Collected last on purpose.  The no-allocation entry point mlau_fd_ct_2d_noalloc (device arrays in
place, scratch in the library's fixed device array) gives exactly the output
of the existing one-step entry point mlau_fd_ct_2d on the same input, on two
calls in a row, the second with a smaller grid (so stale scratch from the
first call would show), and allocates no device memory."""
import ctypes as C

import numpy as np
import pytest

import mlau_amd as M
from mlau_amd import problems as P

pytestmark = pytest.mark.gpu

DP = C.POINTER(C.c_double)


def hip():
    H = C.CDLL("libamdhip64.so")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipDeviceSynchronize.argtypes = []
    H.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return H


H2D, D2H = 1, 2


def noalloc_fn():
    L = M.lib()
    f = L.mlau_fd_ct_2d_noalloc
    f.argtypes = [C.c_void_p] * 8 + [C.c_double] * 3 + [C.c_int] * 4 + [C.c_double, C.POINTER(M.Scheme), C.c_void_p]
    f.restype = C.c_int
    L.mlau_noalloc_max_cells.restype = C.c_longlong
    return f


def run_noalloc(f0, p, sch, dt, nsteps, H, fn, check_free=False):
    ny, nx = f0.shape[1:]
    nbytes = nx * ny * 8
    bufs = []
    for n in range(8):
        d = C.c_void_p()
        assert H.hipMalloc(C.byref(d), nbytes) == 0
        assert H.hipMemcpy(d, f0[n].ctypes.data, nbytes, H2D) == 0
        bufs.append(d)
    free0, tot = C.c_size_t(), C.c_size_t()
    H.hipDeviceSynchronize()
    H.hipMemGetInfo(C.byref(free0), C.byref(tot))
    for _ in range(nsteps):
        rc = fn(*bufs, dt, p["dx"], p["dy"], nx, ny, 4, 4, p["gamma"], C.byref(sch), None)
        assert rc == 0, M.lib().mlau_last_error().decode()
    assert H.hipDeviceSynchronize() == 0
    free1 = C.c_size_t()
    H.hipMemGetInfo(C.byref(free1), C.byref(tot))
    if check_free:
        assert free1.value == free0.value, (free0.value, free1.value)
    out = np.empty_like(f0)
    for n in range(8):
        assert H.hipMemcpy(out[n].ctypes.data, bufs[n], nbytes, D2H) == 0
        H.hipFree(bufs[n])
    return out


def run_existing(f0, p, sch, dt, nsteps):
    f = f0.copy()
    ny, nx = f0.shape[1:]
    for _ in range(nsteps):
        M.mhd_fd_ct_2d(*[f[n] for n in range(8)], dt, p["dx"], p["dy"], nx, ny, 4, 4, p["gamma"], sch)
    return f


def test_noalloc_equals_existing_entry_point(gpu):
    H, fn = hip(), noalloc_fn()
    cases = [(P.vortex_2d(160, 4), M.scheme(4, 3, "lhlld", 1)),   # first call: larger grid
             (P.blast_2d(48, 100.0, 0.5, 0.3), M.scheme(2, 3, "hlld", 1))]  # second call: smaller grid
    for k, ((f0, p), sch) in enumerate(cases):
        dt = 1e-3 if k == 0 else 2e-4
        got = run_noalloc(f0, p, sch, dt, 3, H, fn, check_free=(k == 1))
        want = run_existing(f0, p, sch, dt, 3)
        assert np.all(np.isfinite(want))
        assert np.array_equal(got, want), (k, int((got != want).sum()))


def test_noalloc_rejects_grids_beyond_the_scratch(gpu):
    H, fn = hip(), noalloc_fn()
    cap = M.lib().mlau_noalloc_max_cells()
    assert cap > 500 * 500
    n = int(np.sqrt(cap)) + 8
    null = [None] * 8
    dummy = [C.c_void_p(16)] * 8  # never dereferenced: the size check comes first
    rc = fn(*dummy, 1e-3, 0.01, 0.01, n, n, 4, 4, 5.0 / 3.0, C.byref(M.scheme(2, 3, "hlld", 1)), None)
    assert rc == 1 and b"exceeds the no-allocation scratch" in M.lib().mlau_last_error()
    rc = fn(*null, 1e-3, 0.01, 0.01, 64, 64, 4, 4, 5.0 / 3.0, C.byref(M.scheme(2, 3, "hlld", 1)), None)
    assert rc == 1
