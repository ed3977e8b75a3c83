"""The drop-in boundary on the GPU: the one-step C-ABI forms mlau_fd4c_1d and
mlau_fd_ct_3d, the reference-named C++ shims of libmlau_dropin for 1D and 3D
(mhd_fd4c_1d, mhd_fd_ct_3d with the driver's dnxs/dnys/dnzs), and the
reference's own unmodified 2D driver linked against libmlau_dropin.  Every
result equals the oracle (or the reference binary) value for value.  The 2D
one-step forms are tested in test_gpu_2d.py."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import mlau_amd as M
from mlau_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DP = C.POINTER(C.c_double)
LIBDIR = os.path.dirname(M.LIB_PATH)
REF_DIR = O.REF_DIR


def assert_same(a, b, what):
    assert np.all(np.isfinite(b)), what + ": oracle produced non-finite values"
    bad = a != b
    if bad.any():
        idx = np.argwhere(bad)[:5]
        raise AssertionError("%s: %d mismatches, first %s gpu=%r oracle=%r" % (
            what, bad.sum(), idx.tolist(), a[tuple(idx[0])], b[tuple(idx[0])]))


def dropin():
    L = C.CDLL(os.path.join(LIBDIR, "libmlau_dropin.so"))
    L.mlau_dropin_set_scheme.argtypes = [C.c_int, C.POINTER(M.Scheme)]
    return L


@pytest.mark.parametrize("bc", ["open", "periodic"])
def test_fd4c_1d_one_step_form(gpu, bc):
    """mlau_fd4c_1d (1D/SERIAL/shock/mhd_fd4c_1d.c:60-63 argument list plus a
    scheme), called once per step on host arrays, equals the oracle."""
    f0, p = P.brio_wu(200)
    dt = 0.4 * p["dx"] / O.vmax_1d(f0, p["bx"], p["gamma"])
    f_g, f_o = f0.copy(), f0.copy()
    sch = M.scheme(4, 3, "mlau", 1, bc1d=bc)
    nx = f0.shape[1]
    for _ in range(12):
        M.mhd_fd4c_1d(*[f_g[n] for n in range(7)], p["bx"], dt, p["dx"], nx, 4, p["gamma"], sch)
        O.step_1d(f_o, p["bx"], dt, p["dx"], 4, p["gamma"], O.scheme(4, 3, "mlau", 1), periodic=(bc == "periodic"))
    assert_same(f_g, f_o, "mlau_fd4c_1d " + bc)


def test_fd_ct_3d_one_step_form(gpu):
    """mlau_fd_ct_3d (3D/MPI/OTvortex/mhd_fd_ct_3d.c:99-103 for one rank) on a
    non-cubic periodic grid equals the oracle."""
    g, p = P.shear_layer_3d(16, 0.05)
    f0 = np.ascontiguousarray(g[:, :, :, :])
    dt = 0.4 * p["dx"] / O.cfl_vmax_3d(f0, 4, p["gamma"], 0.0)
    f_g, f_o = f0.copy(), f0.copy()
    sch = M.scheme(4, 3, "mlau", 1)
    nz, ny, nx = f0.shape[1:]
    for _ in range(3):
        M.mhd_fd_ct_3d(*[f_g[n] for n in range(8)], dt, p["dx"], p["dy"], p["dz"], nx, ny, nz, 4, 4, 4, p["gamma"],
                       sch)
        O.step_3d(f_o, dt, p["dx"], p["dy"], p["dz"], 4, p["gamma"], O.scheme(4, 3, "mlau", 1))
    assert_same(f_g, f_o, "mlau_fd_ct_3d")


def test_reference_named_1d_shim(gpu):
    """_Z11mhd_fd4c_1d... of libmlau_dropin (the reference's name, C++ linkage
    and argument list) with its default scheme: the reference file's own
    #defines (ODR=1, R_K=2, LHLLD, open boundary, mhd_fd4c_1d.c:25-26,87)."""
    L = dropin()
    fn = getattr(L, "_Z11mhd_fd4c_1dPdS_S_S_S_S_S_dddiid")
    fn.argtypes = [DP] * 7 + [C.c_double] * 3 + [C.c_int] * 2 + [C.c_double]
    fn.restype = None
    f0, p = P.brio_wu(160)
    dt = 0.4 * p["dx"] / O.vmax_1d(f0, p["bx"], p["gamma"])
    f_g, f_o = f0.copy(), f0.copy()
    nx = f0.shape[1]
    for _ in range(20):
        fn(*[f_g[n].ctypes.data_as(DP) for n in range(7)], p["bx"], dt, p["dx"], nx, 4, p["gamma"])
        O.step_1d(f_o, p["bx"], dt, p["dx"], 4, p["gamma"], O.scheme(1, 2, "lhlld", 1))
    assert_same(f_g, f_o, "1D shim (reference defaults)")
    # another scheme through mlau_dropin_set_scheme
    assert L.mlau_dropin_set_scheme(1, C.byref(M.scheme(2, 3, "roe", 1, bc1d="periodic"))) == 0
    f_g, f_o = f0.copy(), f0.copy()
    for _ in range(20):
        fn(*[f_g[n].ctypes.data_as(DP) for n in range(7)], p["bx"], dt, p["dx"], nx, 4, p["gamma"])
        O.step_1d(f_o, p["bx"], dt, p["dx"], 4, p["gamma"], O.scheme(2, 3, "roe", 1), periodic=True)
    assert L.mlau_dropin_set_scheme(1, C.byref(M.scheme(1, 2, "lhlld", 0, bc1d="open"))) == 0
    assert_same(f_g, f_o, "1D shim (Roe, periodic)")


def test_reference_named_3d_shim(gpu):
    """_Z12mhd_fd_ct_3d... with the boundary flags read from the driver-style
    globals dnxs/dnys/dnzs (walls with sign flips on every axis), default
    scheme = the reference file's #defines (ODR=2, R_K=3, CTW=1, LHLLD)."""
    L = dropin()
    fn = getattr(L, "_Z12mhd_fd_ct_3dPdS_S_S_S_S_S_S_ddddiiiiiidiiii")
    fn.argtypes = [DP] * 8 + [C.c_double] * 4 + [C.c_int] * 6 + [C.c_double] + [C.c_int] * 4
    fn.restype = None
    flags = (1, -1, 1, -1, 1, 1, -1, 1)
    arrs = [(C.c_int * 8).in_dll(L, nm) for nm in ("dnxs", "dnys", "dnzs")]
    g, p = P.ot_vortex_3d(16)
    f0 = np.ascontiguousarray(g[:, :17, :21, :])  # nz=17, ny=21, nx=24
    O.bc3d_state(f0, 4, flags)
    f_g, f_o = f0.copy(), f0.copy()
    nz, ny, nx = f0.shape[1:]
    try:
        for a in arrs:
            for n in range(8):
                a[n] = flags[n]
        for _ in range(2):
            fn(*[f_g[n].ctypes.data_as(DP) for n in range(8)], 2e-3, p["dx"], p["dy"], p["dz"], nx, ny, nz, 4, 4, 4,
               p["gamma"], 0, 1, 1, 1)
            O.step_3d(f_o, 2e-3, p["dx"], p["dy"], p["dz"], 4, p["gamma"], O.scheme(2, 3, "lhlld", 1), dn=flags)
    finally:
        for a in arrs:
            for n in range(8):
                a[n] = 0
    assert_same(f_g, f_o, "3D shim (walls)")


def _run_driver(exe, workdir):
    os.makedirs(os.path.join(workdir, "dat"))
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", "16")
    r = subprocess.run([exe], cwd=workdir, env=env, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, (exe, r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def test_unmodified_reference_driver_on_the_gpu(gpu, tmp_path):
    """The reference's own 2D driver (2D/SERIAL/OTvortex/main.cpp:10-49 with
    boundary.c, unmodified; built by oracle/ref_build.sh) linked against
    libmlau_dropin instead of mhd_fd_ct_2d.o runs its 750 steps on the GPU and
    writes the same dat/ files as the same driver linked with the reference
    solver: text files byte for byte, every outdat record value for value
    (any byte difference may only be the sign of an exact zero)."""
    ref_exe = os.path.join(REF_DIR, "otvortex2d_ref")
    gpu_exe = os.path.join(REF_DIR, "otvortex2d_gpu")
    if not (os.path.exists(ref_exe) and os.path.exists(gpu_exe)):
        pytest.skip("oracle/_ref has no driver builds (made by oracle/ref_build.sh where the reference tree is)")
    out_g = _run_driver(gpu_exe, str(tmp_path / "gpu"))
    _run_driver(ref_exe, str(tmp_path / "ref"))
    assert "750 / 750 iterations finished." in out_g
    dg, dr = tmp_path / "gpu" / "dat", tmp_path / "ref" / "dat"
    names = sorted(os.listdir(dr))
    assert names == sorted(os.listdir(dg))
    outdat = [n for n in names if n.startswith("outdat_")]
    assert len(outdat) == 6  # t = 0 and every 150 steps
    for n in names:
        a, b = (dg / n).read_bytes(), (dr / n).read_bytes()
        if not n.startswith("outdat_"):
            assert a == b, n
            continue
        x, y = np.frombuffer(a, np.float64), np.frombuffer(b, np.float64)
        assert x.shape == y.shape and np.all(np.isfinite(y)), n
        assert np.array_equal(x, y), (n, int((x != y).sum()))
        diff = np.frombuffer(a, np.uint64) != np.frombuffer(b, np.uint64)
        assert np.all(y[diff] == 0.0), n
