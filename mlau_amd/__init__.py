"""mlau_amd -- MI355X (gfx950) implementation of MLAU's per-step ideal-MHD update.

The compute path is the hand-written HIP library ``_lib/libmlau_hip.so``
(C ABI declared in ``include/mlau.h``); this module is a thin ctypes host layer
over it.  It mirrors the reference's per-step entry points
(``mhd_fd4c_1d``, ``mhd_fd_ct_2d``, ``mhd_fd_ct_3d``: same argument meaning,
arrays updated in place) and adds :class:`Solver`, a handle that keeps the
state resident in HBM across steps.

There is no CPU fallback: importing works without a GPU, but every compute
call raises :class:`MLAUError` when the library or the device is missing.
"""
import ctypes as C
import os

import numpy as np

from . import problems  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmlau_hip.so")

FLUX = {"mlau": 0, "roe": 1, "hlld": 2, "lhlld": 3}
BC1D = {"open": 0, "periodic": 1}
PRECISION = {"fp64": 0, "mixed": 1, "mixed_recon": 2}


class MLAUError(RuntimeError):
    pass


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("xoff", C.c_int), ("yoff", C.c_int), ("zoff", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("gamma", C.c_double), ("bx", C.c_double)]


class Scheme(C.Structure):
    _fields_ = [("odr", C.c_int), ("rk", C.c_int), ("flux", C.c_int), ("ctw", C.c_int), ("bc1d", C.c_int),
                ("dnxs", C.c_int * 8), ("dnys", C.c_int * 8), ("dnzs", C.c_int * 8), ("precision", C.c_int)]


def scheme(odr=4, rk=3, flux="lhlld", ctw=1, bc1d="open", dnxs=(0,) * 8, dnys=None, dnzs=None, precision="fp64"):
    """Scheme selection the reference makes with #defines (ODR, R_K, CTW,
    func_flux) and driver globals (dnxs/dnys/dnzs boundary flags), plus the
    arithmetic: "fp64" (the reference's), "mixed" (FP32 reconstruction and
    flux with FP64 pressure terms, accumulation and CT) or "mixed_recon" (FP32
    reconstruction with the FP64 Riemann solvers, for long runs)."""
    s = Scheme()
    s.odr, s.rk, s.ctw = int(odr), int(rk), int(ctw)
    s.flux = FLUX[flux] if isinstance(flux, str) else int(flux)
    s.bc1d = BC1D[bc1d] if isinstance(bc1d, str) else int(bc1d)
    s.precision = PRECISION[precision] if isinstance(precision, str) else int(precision)
    for n in range(8):
        s.dnxs[n] = dnxs[n]
        s.dnys[n] = (dnxs if dnys is None else dnys)[n]
        s.dnzs[n] = (dnxs if dnzs is None else dnzs)[n]
    return s


_lib = None
_dp = C.POINTER(C.c_double)


def lib():
    """Load libmlau_hip.so (raises MLAUError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MLAUError("libmlau_hip.so is not built (run __graft_entry__.build()); no CPU fallback exists")
        L = C.CDLL(LIB_PATH)
        i, d, vp = C.c_int, C.c_double, C.c_void_p
        L.mlau_last_error.restype = C.c_char_p
        L.mlau_create.argtypes = [C.POINTER(Grid), C.POINTER(Scheme), C.POINTER(vp)]
        L.mlau_destroy.argtypes = [vp]
        L.mlau_destroy.restype = None
        L.mlau_upload.argtypes = [vp, C.POINTER(_dp)]
        L.mlau_download.argtypes = [vp, C.POINTER(_dp)]
        L.mlau_set_dt.argtypes = [vp, d]
        L.mlau_step.argtypes = [vp, i]
        L.mlau_cfl.argtypes = [vp, d, d, _dp, _dp]
        L.mlau_step_cfl.argtypes = [vp, i, d, d]
        L.mlau_get_time.argtypes = [vp, _dp, _dp, C.POINTER(C.c_longlong)]
        L.mlau_set_time.argtypes = [vp, d, C.c_longlong]
        L.mlau_sync.argtypes = [vp]
        L.mlau_device_fields.argtypes = [vp, C.POINTER(_dp)]
        L.mlau_kernel_timing.argtypes = [vp, i]
        L.mlau_kernel_stats.argtypes = [vp, i, _dp, C.POINTER(C.c_longlong), C.POINTER(C.c_char_p)]
        L.mlau_num_kernels.restype = i
        L.mlau_fd4c_1d.argtypes = [_dp] * 7 + [d, d, d, i, i, d, C.POINTER(Scheme)]
        L.mlau_fd_ct_2d.argtypes = [_dp] * 8 + [d, d, d, i, i, i, i, d, C.POINTER(Scheme)]
        L.mlau_fd_ct_3d.argtypes = [_dp] * 8 + [d, d, d, d, i, i, i, i, i, i, d, C.POINTER(Scheme)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise MLAUError(lib().mlau_last_error().decode() or "error %d" % rc)


def _ptrs(arrs):
    for a in arrs:
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise MLAUError("fields must be C-contiguous float64 arrays")
    return (_dp * len(arrs))(*[a.ctypes.data_as(_dp) for a in arrs])


class Solver:
    """Resident solver: one problem's state kept in HBM across steps.

    ``fields`` is a float64 array (7, nx) in 1D or (8, [nz,] ny, nx) with the
    reference's argument order and ghost cells.
    """

    def __init__(self, fields, sch, dx, dy=None, dz=None, gamma=5.0 / 3.0, bx=0.0, xoff=4):
        f = np.ascontiguousarray(fields, dtype=np.float64)
        dim = f.ndim - 1
        g = Grid()
        g.dim = dim
        g.nx = f.shape[-1]
        g.ny = f.shape[-2] if dim >= 2 else 1
        g.nz = f.shape[-3] if dim == 3 else 1
        g.xoff = xoff
        g.yoff = xoff if dim >= 2 else 0
        g.zoff = xoff if dim == 3 else 0
        g.dx = dx
        g.dy = dy if dy is not None else (dx if dim >= 2 else 1.0)
        g.dz = dz if dz is not None else (dx if dim == 3 else 1.0)
        g.gamma = gamma
        g.bx = bx
        self.grid, self.scheme, self.shape, self.dim = g, sch, f.shape, dim
        self._h = C.c_void_p()
        _check(lib().mlau_create(C.byref(g), C.byref(sch), C.byref(self._h)))
        self.upload(f)

    def close(self):
        if self._h:
            lib().mlau_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        if f.shape != self.shape:
            raise MLAUError("shape mismatch %s vs %s" % (f.shape, self.shape))
        _check(lib().mlau_upload(self._h, _ptrs([f[n] for n in range(f.shape[0])])))

    def download(self, out=None):
        if out is None:
            out = np.empty(self.shape)
        _check(lib().mlau_download(self._h, _ptrs([out[n] for n in range(out.shape[0])])))
        return out

    def set_dt(self, dt):
        _check(lib().mlau_set_dt(self._h, dt))

    def step(self, n=1):
        _check(lib().mlau_step(self._h, int(n)))

    def step_cfl(self, n, cfl=0.4, vfloor=0.0):
        _check(lib().mlau_step_cfl(self._h, int(n), cfl, vfloor))

    def cfl(self, cfl=0.4, vfloor=0.0):
        dt, vm = C.c_double(), C.c_double()
        _check(lib().mlau_cfl(self._h, cfl, vfloor, C.byref(dt), C.byref(vm)))
        return dt.value, vm.value

    def time(self):
        t, dt, n = C.c_double(), C.c_double(), C.c_longlong()
        _check(lib().mlau_get_time(self._h, C.byref(t), C.byref(dt), C.byref(n)))
        return t.value, dt.value, n.value

    def set_time(self, tim, nstep):
        _check(lib().mlau_set_time(self._h, float(tim), int(nstep)))

    def sync(self):
        _check(lib().mlau_sync(self._h))

    def checkpoint(self, directory, cnt=0, trec=0.0, rank=0):
        """Write the reference's backup files (bkup_stamp/time/data) for the
        current state, time, step count and next dt (2D/MPI/OTvortex/bkup.hpp)."""
        from . import io
        tim, dt, n = self.time()
        io.bkup_save(directory, self.download(), n, cnt, tim, dt, trec, rank)

    def resume(self, directory, rank=0):
        """Load a backup written by checkpoint() (or by the reference) and
        continue from it; returns its stamp, or None when there is none."""
        from . import io
        got = io.bkup_load(directory, self.shape, rank)
        if got is None:
            return None
        f, st = got
        self.upload(f)
        self.set_dt(st["dt"])
        self.set_time(st["tim"], st["n"])
        return st

    def kernel_timing(self, enable=True):
        _check(lib().mlau_kernel_timing(self._h, 1 if enable else 0))

    def kernel_stats(self):
        out = {}
        for k in range(lib().mlau_num_kernels()):
            ms, cnt, nm = C.c_double(), C.c_longlong(), C.c_char_p()
            _check(lib().mlau_kernel_stats(self._h, k, C.byref(ms), C.byref(cnt), C.byref(nm)))
            if cnt.value:
                out[nm.value.decode()] = (ms.value, cnt.value)
        return out


# ---- reference-named one-step entry points (arrays updated in place) ----
def mhd_fd4c_1d(ro, mx, my, mz, en, by, bz, bx, dt, dx, nx, xoff, gamma, sch):
    """One step of 1D/SERIAL/shock/mhd_fd4c_1d.c:60-63 on the GPU."""
    p = _ptrs([ro, mx, my, mz, en, by, bz])
    _check(lib().mlau_fd4c_1d(*p, bx, dt, dx, nx, xoff, gamma, C.byref(sch)))


def mhd_fd_ct_2d(ro, mx, my, mz, en, bx, by, bz, dt, dx, dy, nx, ny, xoff, yoff, gamma, sch):
    """One step of 2D/SERIAL/OTvortex/mhd_fd_ct_2d.c:91-94 on the GPU."""
    p = _ptrs([ro, mx, my, mz, en, bx, by, bz])
    _check(lib().mlau_fd_ct_2d(*p, dt, dx, dy, nx, ny, xoff, yoff, gamma, C.byref(sch)))


def mhd_fd_ct_3d(ro, mx, my, mz, en, bx, by, bz, dt, dx, dy, dz, nx, ny, nz, xoff, yoff, zoff, gamma, sch):
    """One step of 3D/MPI/OTvortex/mhd_fd_ct_3d.c:99-103 (single rank) on the GPU."""
    p = _ptrs([ro, mx, my, mz, en, bx, by, bz])
    _check(lib().mlau_fd_ct_3d(*p, dt, dx, dy, dz, nx, ny, nz, xoff, yoff, zoff, gamma, C.byref(sch)))
