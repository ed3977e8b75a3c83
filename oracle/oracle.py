"""ctypes bindings for the CPU restatement (``libmlau_oracle.so``) and for the
reference checkers built from the reference sources (``_ref/*.so``).

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke check and the
CPU-baseline leg of bench.py -- never by the product package ``mlau_amd``.
All state arrays follow the reference layout (SoA float64 with ghost cells,
x fastest); step functions update them in place like the reference.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

FLUX = {"mlau": 0, "roe": 1, "hlld": 2, "lhlld": 3}


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _ia(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32))
    return a, a.ctypes.data_as(_ip)


class Scheme(C.Structure):
    _fields_ = [("odr", C.c_int), ("rk", C.c_int), ("flux", C.c_int), ("ctw", C.c_int)]


def scheme(odr=4, rk=3, flux="lhlld", ctw=1):
    return Scheme(odr, rk, FLUX[flux] if isinstance(flux, str) else int(flux), ctw)


_lib = None


def lib():
    """Load (building if needed) the oracle shared library."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libmlau_oracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(path)
        d, i = C.c_double, C.c_int
        L.oc_minmod.restype = d
        L.oc_minmod.argtypes = [d, d]
        L.oc_minmod3.restype = d
        L.oc_minmod3.argtypes = [d, d, d]
        for n in ("oc_muscl_mm", "oc_muscl_mc", "oc_muscl_kr"):
            getattr(L, n).argtypes = [d, d, d, _dp, _dp]
        L.oc_wcns3.argtypes = [d, d, d, _dp, _dp]
        L.oc_wcns4.argtypes = [d, d, d, d, d, _dp, _dp]
        L.oc_eigen_st.argtypes = [_dp, d, _dp, _dp, _dp]
        L.oc_eigen_ae.argtypes = [_dp, d, _dp, _dp]
        L.oc_reconst.argtypes = [i, i, _dp, d, d, _dp, _dp]
        L.oc_flux.argtypes = [i, _dp, d, d, _dp, _dp]
        L.oc_step_1d.argtypes = [_dp] * 7 + [d, d, d, i, i, d, C.POINTER(Scheme), i]
        L.oc_step_2d.argtypes = [_dp] * 8 + [d, d, d, i, i, i, i, d, C.POINTER(Scheme)] + [_ip] * 4
        L.oc_step_3d.argtypes = [_dp] * 8 + [d, d, d, d, i, i, i, i, i, i, d, C.POINTER(Scheme)] + [_ip] * 6
        L.oc_bc2d.argtypes = [_dp, i, i, i, i, i, i, i, i]
        L.oc_cfl_vmax_2d.restype = d
        L.oc_cfl_vmax_2d.argtypes = [_dp] * 8 + [i, i, i, i, d, d]
        L.oc_cfl_vmax_3d.restype = d
        L.oc_cfl_vmax_3d.argtypes = [_dp] * 8 + [i, i, i, i, i, i, d, d]
        L.oc_vmax_1d.restype = d
        L.oc_vmax_1d.argtypes = [_dp] * 7 + [d, i, d, d]
        L.oc_num_threads.restype = i
        L.oc_set_num_threads.argtypes = [i]
        _lib = L
    return _lib


# ---------------------------------------------------------------- point API
def flux(kind, q, bn, gamma, dv, L=None):
    L = L or lib()
    q = np.ascontiguousarray(q, dtype=np.float64)
    dv = np.ascontiguousarray(dv, dtype=np.float64)
    out = np.zeros(7)
    fn = L.oc_flux if hasattr(L, "oc_flux") else L.ref_flux
    fn(FLUX.get(kind, kind) if isinstance(kind, str) else kind, _p(q), bn, gamma, _p(dv), _p(out))
    return out


def reconst(kind, odr, st, bn, gamma, L=None):
    L = L or lib()
    st = np.ascontiguousarray(st, dtype=np.float64).reshape(35)
    ul, ur = np.zeros(7), np.zeros(7)
    fn = L.oc_reconst if hasattr(L, "oc_reconst") else L.ref_reconst
    fn(kind, odr, _p(st), bn, gamma, _p(ul), _p(ur))
    return ul, ur


# ---------------------------------------------------------------- step API
def step_2d(f, dt, dx, dy, xoff, gamma, sch, dn=(0,) * 8, dny=None, stx=(0, 0, 0, 0, 0, 1, 0, 0),
            sty=(0, 0, 0, 0, 0, 0, 1, 0)):
    """f: float64 array (8, ny, nx) in argument order ro,mx,my,mz,en,bx,by,bz."""
    L = lib()
    ny, nx = f.shape[1:]
    a = [_ia(dn), _ia(dn if dny is None else dny), _ia(stx), _ia(sty)]
    L.oc_step_2d(*[_p(f[n]) for n in range(8)], dt, dx, dy, nx, ny, xoff, xoff, gamma, C.byref(sch),
                 *[x[1] for x in a])


def step_3d(f, dt, dx, dy, dz, xoff, gamma, sch, dn=(0,) * 8):
    L = lib()
    nz, ny, nx = f.shape[1:]
    a = [_ia(dn), _ia(dn), _ia(dn), _ia((0, 0, 0, 0, 0, 1, 0, 0)), _ia((0, 0, 0, 0, 0, 0, 1, 0)),
         _ia((0, 0, 0, 0, 0, 0, 0, 1))]
    L.oc_step_3d(*[_p(f[n]) for n in range(8)], dt, dx, dy, dz, nx, ny, nz, xoff, xoff, xoff, gamma,
                 C.byref(sch), *[x[1] for x in a])


def bc3d_state(f, xoff, dn=(0,) * 8):
    """Ghost cells of a 3D state under the boundary flags dn on every axis
    (oc_bc3d_state; 3D/MPI/OTvortex/mhd_fd_ct_3d.c:1161-1175)."""
    nz, ny, nx = f.shape[1:]
    ptrs = (_dp * 8)(*[_p(f[n]) for n in range(8)])
    a = [_ia(dn), _ia(dn), _ia(dn), _ia((0, 0, 0, 0, 0, 1, 0, 0)), _ia((0, 0, 0, 0, 0, 0, 1, 0)),
         _ia((0, 0, 0, 0, 0, 0, 0, 1))]
    lib().oc_bc3d_state(ptrs, 8, nx, ny, nz, xoff, xoff, xoff, *[x[1] for x in a])


def step_1d(f, bx, dt, dx, xoff, gamma, sch, periodic=False):
    """f: float64 array (7, nx) in argument order ro,mx,my,mz,en,by,bz."""
    L = lib()
    nx = f.shape[1]
    L.oc_step_1d(*[_p(f[n]) for n in range(7)], bx, dt, dx, nx, xoff, gamma, C.byref(sch), 1 if periodic else 0)


def cfl_vmax_2d(f, xoff, gamma, vfloor):
    ny, nx = f.shape[1:]
    return lib().oc_cfl_vmax_2d(*[_p(f[n]) for n in range(8)], nx, ny, xoff, xoff, gamma, vfloor)


def cfl_vmax_3d(f, xoff, gamma, vfloor):
    nz, ny, nx = f.shape[1:]
    return lib().oc_cfl_vmax_3d(*[_p(f[n]) for n in range(8)], nx, ny, nz, xoff, xoff, xoff, gamma, vfloor)


def vmax_1d(f, bx, gamma, vfloor=1.0):
    return lib().oc_vmax_1d(*[_p(f[n]) for n in range(7)], bx, f.shape[1], gamma, vfloor)


def _finite_state(f, xoff):
    """Failure rule of the batched engine (the reference has none): a member
    whose density or energy is not finite in a cell the CFL check visits fails
    at that step, before it is advanced.  The speed itself follows the
    reference's "if (vtmp > vmax)", so a NaN speed of a finite cell is skipped."""
    if f.ndim == 2:  # 1D: the driver's estimate runs over all cells
        return bool(np.isfinite(f[0]).all() and np.isfinite(f[4]).all())
    inner = (slice(xoff, -xoff),) * (f.ndim - 1)
    return bool(np.isfinite(f[0][inner]).all() and np.isfinite(f[4][inner]).all())


def run_to_2d(f, dx, dy, xoff, gamma, sch, tend, cfl, vfloor, dn=(0,) * 8, dny=None, max_steps=0):
    """Run one 2D problem to time tend, in place: the driver loop
    cflcheck -> mhd_fd_ct_2d -> t += dt of 2D/SERIAL/blast/main.cpp:20-35 and
    cflcheck.hpp:1-32 (vmax from vfloor, dt = cfl*min(dx,dy)/vmax), with the
    end-time rule of the batched engine: when tim + dt >= tend the step uses
    dt = tend - tim and ends at tim = tend.  Returns (tim, nstep, status) with
    status 1 done, 2 failed (vmax not finite, or a non-finite density or
    energy: _finite_state), 3 capped by max_steps."""
    tim, n = 0.0, 0
    while True:
        vmax = cfl_vmax_2d(f, xoff, gamma, vfloor)
        if not np.isfinite(vmax) or not _finite_state(f, xoff):
            return tim, n, 2
        dt = cfl * min(dx, dy) / vmax
        last = tim + dt >= tend
        if last:
            dt = tend - tim
        step_2d(f, dt, dx, dy, xoff, gamma, sch, dn=dn, dny=dny)
        tim = tend if last else tim + dt
        n += 1
        if last:
            return tim, n, 1
        if max_steps and n >= max_steps:
            return tim, n, 3


def run_to_3d(f, dx, dy, dz, xoff, gamma, sch, tend, cfl, vfloor, dn=(0,) * 8, max_steps=0):
    """3D counterpart of run_to_2d: cflcheck of 3D/MPI/OTvortex/cflcheck.hpp:1-39
    (dt = cfl*min(dx,dy,dz)/vmax) before every step, same end-time rule."""
    tim, n = 0.0, 0
    while True:
        vmax = cfl_vmax_3d(f, xoff, gamma, vfloor)
        if not np.isfinite(vmax) or not _finite_state(f, xoff):
            return tim, n, 2
        dt = cfl * min(dx, dy, dz) / vmax
        last = tim + dt >= tend
        if last:
            dt = tend - tim
        step_3d(f, dt, dx, dy, dz, xoff, gamma, sch, dn=dn)
        tim = tend if last else tim + dt
        n += 1
        if last:
            return tim, n, 1
        if max_steps and n >= max_steps:
            return tim, n, 3


def run_to_1d(f, bx, dx, xoff, gamma, sch, tend, cfl, vfloor, periodic=False, max_steps=0):
    """1D driver rule (1D/SERIAL/shock/main.cpp:113-130): one dt = cfl*dx/vmax
    from the initial state over all cells, then steps with that dt; the last
    step is clipped at tend as in run_to_2d."""
    vmax = vmax_1d(f, bx, gamma, vfloor)
    if not np.isfinite(vmax) or not _finite_state(f, xoff):
        return 0.0, 0, 2
    dtc = cfl * dx / vmax
    tim, n = 0.0, 0
    while True:
        dt = dtc
        last = tim + dt >= tend
        if last:
            dt = tend - tim
        step_1d(f, bx, dt, dx, xoff, gamma, sch, periodic=periodic)
        tim = tend if last else tim + dt
        n += 1
        if last:
            return tim, n, 1
        if max_steps and n >= max_steps:
            return tim, n, 3


def set_threads(n):
    lib().oc_set_num_threads(int(n))


def threads():
    return lib().oc_num_threads()


# ------------------------------------------------------ reference checkers
def ref_available():
    return os.path.exists(os.path.join(REF_DIR, "libref_kat.so"))


_ref_cache = {}


def ref_lib(name):
    """Load oracle/_ref/lib<name>.so (built by ref_build.sh)."""
    if name not in _ref_cache:
        L = C.CDLL(os.path.join(REF_DIR, "lib%s.so" % name), mode=C.RTLD_LOCAL)
        d, i = C.c_double, C.c_int
        if name == "ref_kat":
            L.ref_flux.argtypes = [i, _dp, d, d, _dp, _dp]
            L.ref_reconst.argtypes = [i, i, _dp, d, d, _dp, _dp]
            L.ref_eigen.argtypes = [i, _dp, d, _dp, _dp, _dp]
            L.ref_wcns4.argtypes = [d, d, d, d, d, _dp, _dp]
            L.ref_wcns3.argtypes = [d, d, d, _dp, _dp]
            L.ref_muscl.argtypes = [i, d, d, d, _dp, _dp]
            L.ref_minmod.restype = d
            L.ref_minmod.argtypes = [d, d]
            L.ref_minmod3.restype = d
            L.ref_minmod3.argtypes = [d, d, d]
        else:
            L.ref_set_bc.argtypes = [_ip, _ip, _ip]
            if name.startswith("ref2d"):
                L.ref_step_2d.argtypes = [_dp] * 8 + [d, d, d, i, i, i, i, d]
            elif name.startswith("ref1d"):
                L.ref_step_1d.argtypes = [_dp] * 7 + [d, d, d, i, i, d]
            else:
                L.ref_step_3d.argtypes = [_dp] * 8 + [d, d, d, d, i, i, i, i, i, i, d]
        _ref_cache[name] = L
    return _ref_cache[name]


def ref_step_2d(name, f, dt, dx, dy, xoff, gamma, dn=(0,) * 8):
    L = ref_lib(name)
    a = _ia(dn)
    L.ref_set_bc(a[1], a[1], a[1])
    ny, nx = f.shape[1:]
    L.ref_step_2d(*[_p(f[n]) for n in range(8)], dt, dx, dy, nx, ny, xoff, xoff, gamma)


def ref_step_3d(name, f, dt, dx, dy, dz, xoff, gamma, dn=(0,) * 8):
    L = ref_lib(name)
    a = _ia(dn)
    L.ref_set_bc(a[1], a[1], a[1])
    nz, ny, nx = f.shape[1:]
    L.ref_step_3d(*[_p(f[n]) for n in range(8)], dt, dx, dy, dz, nx, ny, nz, xoff, xoff, xoff, gamma)


def ref_step_1d(name, f, bx, dt, dx, xoff, gamma):
    L = ref_lib(name)
    L.ref_step_1d(*[_p(f[n]) for n in range(7)], bx, dt, dx, f.shape[1], xoff, gamma)
