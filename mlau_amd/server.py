"""Solve server (C ABI in include/mlau_server.h): continuous batching of
independent 1D/2D/3D MHD problems on one GPU.

A worker thread inside the library owns the GPU; ``submit`` may be called
from any thread while earlier requests run, and newly arrived requests join
the running batch at the next step boundary as device memory allows.
"""
import ctypes as C

import numpy as np

from . import MLAUError, _dp, _ptrs, lib
from .ensemble import Member, _lib as _ens_lib

_bound = False


class Times(C.Structure):
    _fields_ = [("submit_s", C.c_double), ("admit_s", C.c_double), ("finish_s", C.c_double),
                ("nstep", C.c_longlong), ("zone_updates", C.c_longlong)]


def _lib():
    global _bound
    L = _ens_lib()
    if not _bound:
        i, d, vp, ll = C.c_int, C.c_double, C.c_void_p, C.c_longlong
        L.mlau_srv_create.argtypes = [ll, i, C.POINTER(vp)]
        L.mlau_srv_destroy.argtypes = [vp]
        L.mlau_srv_destroy.restype = None
        L.mlau_srv_submit.argtypes = [vp, C.POINTER(Member), C.POINTER(_dp), C.POINTER(i)]
        L.mlau_srv_wait.argtypes = [vp, i, d, C.POINTER(i)]
        L.mlau_srv_result.argtypes = [vp, i, C.POINTER(_dp), C.POINTER(Times)]
        L.mlau_srv_load.argtypes = [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(ll)]
        L.mlau_srv_clock.argtypes = [vp]
        L.mlau_srv_clock.restype = d
        _bound = True
    return L


def _check(rc):
    if rc != 0:
        raise MLAUError(lib().mlau_last_error().decode() or "error %d" % rc)


class Server:
    def __init__(self, budget_bytes=200 << 30, steps_per_round=8):
        self._h = C.c_void_p()
        self._shapes = {}
        _check(_lib().mlau_srv_create(int(budget_bytes), int(steps_per_round), C.byref(self._h)))

    def close(self):
        """Finish every queued and running request, then stop."""
        if self._h:
            _lib().mlau_srv_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, fields, member):
        f = np.ascontiguousarray(fields, dtype=np.float64)
        t = C.c_int()
        _check(_lib().mlau_srv_submit(self._h, C.byref(member), _ptrs([f[n] for n in range(f.shape[0])]),
                                      C.byref(t)))
        self._shapes[t.value] = f.shape
        return t.value

    def wait(self, ticket, timeout=600.0):
        s = C.c_int()
        _check(_lib().mlau_srv_wait(self._h, int(ticket), float(timeout), C.byref(s)))
        return s.value

    def result(self, ticket):
        """(fields, times) of a finished request; raises MLAUError for a
        failed one (after releasing it)."""
        out = np.empty(self._shapes.pop(ticket))
        tm = Times()
        _check(_lib().mlau_srv_result(self._h, int(ticket), _ptrs([out[n] for n in range(out.shape[0])]),
                                      C.byref(tm)))
        return out, {k: getattr(tm, k) for k, _ in Times._fields_}

    def load(self):
        q, r, f, b = C.c_int(), C.c_int(), C.c_int(), C.c_longlong()
        _check(_lib().mlau_srv_load(self._h, C.byref(q), C.byref(r), C.byref(f), C.byref(b)))
        return {"queued": q.value, "running": r.value, "finished": f.value, "bytes": b.value}

    def clock(self):
        return _lib().mlau_srv_clock(self._h)
