"""Batched ensemble engine (C ABI in include/mlau_ensemble.h).

Many independent 1D, 2D and 3D problems advance together on one GPU, each with
its own grid, Riemann solver, boundary flags, CFL time step and end time;
members can be added and removed between calls.  Per member the result is
bit-identical to running it alone.
"""
import ctypes as C

import numpy as np

from . import FLUX, Grid, MLAUError, Scheme, _dp, _ptrs, lib, scheme

RUNNING, DONE, FAILED, CAPPED = 0, 1, 2, 3
GRAPH, BALANCE, UNFUSED = 1, 2, 4


class Member(C.Structure):
    _fields_ = [("grid", Grid), ("scheme", Scheme), ("tend", C.c_double), ("cfl", C.c_double),
                ("vfloor", C.c_double), ("max_steps", C.c_longlong)]


_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        i, d, vp, ll = C.c_int, C.c_double, C.c_void_p, C.c_longlong
        L.mlau_ens_create.argtypes = [ll, i, C.POINTER(vp)]
        L.mlau_ens_destroy.argtypes = [vp]
        L.mlau_ens_destroy.restype = None
        L.mlau_ens_member_bytes.argtypes = [C.POINTER(Member)]
        L.mlau_ens_member_bytes.restype = ll
        L.mlau_ens_used_bytes.argtypes = [vp]
        L.mlau_ens_used_bytes.restype = ll
        L.mlau_ens_reserved_bytes.argtypes = [vp]
        L.mlau_ens_reserved_bytes.restype = ll
        L.mlau_ens_add.argtypes = [vp, C.POINTER(Member), C.POINTER(_dp), C.POINTER(i)]
        L.mlau_ens_remove.argtypes = [vp, i]
        L.mlau_ens_run.argtypes = [vp, i, C.POINTER(i)]
        L.mlau_ens_launch.argtypes = [vp, i]
        L.mlau_ens_sync.argtypes = [vp, C.POINTER(i)]
        L.mlau_ens_ready.argtypes = [vp]
        L.mlau_ens_status.argtypes = [vp, i, C.POINTER(i), C.POINTER(d), C.POINTER(d), C.POINTER(ll)]
        L.mlau_ens_finished.argtypes = [vp, C.POINTER(i), i, C.POINTER(i)]
        L.mlau_ens_download.argtypes = [vp, i, C.POINTER(_dp)]
        L.mlau_ens_set_tune.argtypes = [vp, i]
        L.mlau_ens_autotune.argtypes = [vp, i, C.POINTER(i)]
        L.mlau_ens_kernel_timing.argtypes = [vp, i]
        L.mlau_ens_kernel_stats.argtypes = [vp, i, C.POINTER(d), C.POINTER(ll), C.POINTER(C.c_char_p)]
        L.mlau_ens_num_kernels.restype = i
        L.mlau_ens_zone_updates.argtypes = [vp]
        L.mlau_ens_zone_updates.restype = ll
        _bound = True
    return L


def _check(rc):
    if rc != 0:
        raise MLAUError(lib().mlau_last_error().decode() or "error %d" % rc)


def member(fields, sch, dx, dy=None, dz=None, tend=0.1, cfl=0.4, vfloor=0.0, gamma=5.0 / 3.0, max_steps=0, xoff=4,
           bx=0.0):
    """Member description from its initial fields (mlau_upload layout): shape
    (7, nx) in 1D (bx: the constant normal field), (8, ny, nx) in 2D and
    (8, nz, ny, nx) in 3D."""
    dim = fields.ndim - 1
    m = Member()
    m.grid.dim = dim
    m.grid.nx = fields.shape[-1]
    m.grid.ny = fields.shape[-2] if dim >= 2 else 1
    m.grid.nz = fields.shape[-3] if dim == 3 else 1
    m.grid.xoff = xoff
    m.grid.yoff = xoff if dim >= 2 else 0
    m.grid.zoff = xoff if dim == 3 else 0
    m.grid.dx = dx
    m.grid.dy = dy if dy is not None else (dx if dim >= 2 else 1.0)
    m.grid.dz = dz if dz is not None else (dx if dim == 3 else 1.0)
    m.grid.gamma = gamma
    m.grid.bx = bx
    m.scheme = sch
    m.tend, m.cfl, m.vfloor, m.max_steps = tend, cfl, vfloor, int(max_steps)
    return m


class Ensemble:
    """Resident batch of members (see include/mlau_ensemble.h)."""

    def __init__(self, budget_bytes=16 << 30, graph=True, balance=False, fused=True):
        self._h = C.c_void_p()
        self._shapes = {}
        flags = (GRAPH if graph else 0) | (BALANCE if balance else 0) | (0 if fused else UNFUSED)
        _check(_lib().mlau_ens_create(int(budget_bytes), flags, C.byref(self._h)))

    def close(self):
        if self._h:
            _lib().mlau_ens_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def member_bytes(m):
        return _lib().mlau_ens_member_bytes(C.byref(m))

    def used_bytes(self):
        return _lib().mlau_ens_used_bytes(self._h)

    def reserved_bytes(self):
        """Device bytes held for member memory (never above the budget)."""
        return _lib().mlau_ens_reserved_bytes(self._h)

    def add(self, fields, m):
        f = np.ascontiguousarray(fields, dtype=np.float64)
        i = C.c_int()
        _check(_lib().mlau_ens_add(self._h, C.byref(m), _ptrs([f[n] for n in range(f.shape[0])]), C.byref(i)))
        self._shapes[i.value] = f.shape
        return i.value

    def remove(self, mid):
        _check(_lib().mlau_ens_remove(self._h, int(mid)))
        self._shapes.pop(mid, None)

    def run(self, nsteps):
        n = C.c_int()
        _check(_lib().mlau_ens_run(self._h, int(nsteps), C.byref(n)))
        return n.value

    def launch(self, nsteps):
        """Queue nsteps steps and return at once (see sync)."""
        _check(_lib().mlau_ens_launch(self._h, int(nsteps)))

    def ready(self):
        """True when sync would not wait."""
        r = _lib().mlau_ens_ready(self._h)
        if r < 0:
            _check(-r)
        return r == 1

    def sync(self):
        """Wait for launched steps; returns the number of running members."""
        n = C.c_int()
        _check(_lib().mlau_ens_sync(self._h, C.byref(n)))
        return n.value

    def run_all(self, chunk=64, max_calls=1 << 20):
        for _ in range(max_calls):
            if self.run(chunk) == 0:
                return
        raise MLAUError("members still running after %d calls" % max_calls)

    def status(self, mid):
        s, t, dt, n = C.c_int(), C.c_double(), C.c_double(), C.c_longlong()
        _check(_lib().mlau_ens_status(self._h, int(mid), C.byref(s), C.byref(t), C.byref(dt), C.byref(n)))
        return s.value, t.value, dt.value, n.value

    def finished(self):
        n = C.c_int()
        _check(_lib().mlau_ens_finished(self._h, None, 0, C.byref(n)))
        ids = (C.c_int * max(n.value, 1))()
        _check(_lib().mlau_ens_finished(self._h, ids, n.value, C.byref(n)))
        return list(ids[:n.value])

    def download(self, mid, out=None):
        if out is None:
            out = np.empty(self._shapes[mid])
        _check(_lib().mlau_ens_download(self._h, int(mid), _ptrs([out[n] for n in range(out.shape[0])])))
        return out

    def set_tune(self, tune):
        _check(_lib().mlau_ens_set_tune(self._h, int(tune)))

    def autotune(self, steps=16):
        t = C.c_int()
        _check(_lib().mlau_ens_autotune(self._h, int(steps), C.byref(t)))
        return t.value

    def kernel_timing(self, enable=True):
        _check(_lib().mlau_ens_kernel_timing(self._h, 1 if enable else 0))

    def kernel_stats(self):
        out = {}
        L = _lib()
        for k in range(L.mlau_ens_num_kernels()):
            ms, cnt, nm = C.c_double(), C.c_longlong(), C.c_char_p()
            _check(L.mlau_ens_kernel_stats(self._h, k, C.byref(ms), C.byref(cnt), C.byref(nm)))
            if cnt.value:
                out[nm.value.decode()] = (ms.value, cnt.value)
        return out

    def zone_updates(self):
        return _lib().mlau_ens_zone_updates(self._h)


__all__ = ["Ensemble", "Member", "member", "scheme", "FLUX", "RUNNING", "DONE", "FAILED", "CAPPED"]
