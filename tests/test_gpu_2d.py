"""2D parity of the HIP path against the oracle (CPU restatement pinned to the
reference).  Results must be equal value for value (==; the GPU may differ from
the reference only in the sign of an exact zero)."""
import numpy as np
import pytest

import mlau_amd as M
from mlau_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SCHEMES = [  # (odr, rk, flux, ctw)
    (4, 3, "lhlld", 1), (2, 3, "mlau", 1), (2, 3, "hlld", 1), (2, 3, "roe", 1), (1, 3, "lhlld", 1),
    (3, 3, "mlau", 1), (4, 2, "mlau", 1), (2, 1, "hlld", 1), (2, 3, "lhlld", 0), (4, 3, "roe", 1),
]


def assert_same(a, b, what):
    assert np.all(np.isfinite(b)), what + ": oracle produced non-finite values"
    bad = a != b
    if bad.any():
        idx = np.argwhere(bad)[:5]
        raise AssertionError("%s: %d mismatches, first %s gpu=%r oracle=%r maxabs=%g" % (
            what, bad.sum(), idx.tolist(), a[tuple(idx[0])], b[tuple(idx[0])], np.abs(a - b).max()))


def run_pair(f0, p, odr, rk, flux, ctw, nsteps, dn=(0,) * 8):
    dt = 0.4 * min(p["dx"], p["dy"]) / O.cfl_vmax_2d(f0, 4, p["gamma"], p["vfloor"])
    f_o = f0.copy()
    sch_o = O.scheme(odr, rk, flux, ctw)
    for _ in range(nsteps):
        O.step_2d(f_o, dt, p["dx"], p["dy"], 4, p["gamma"], sch_o, dn=dn)
    s = M.Solver(f0, M.scheme(odr, rk, flux, ctw, dnxs=dn), p["dx"], p["dy"], gamma=p["gamma"])
    s.set_dt(dt)
    s.step(nsteps)
    f_g = s.download()
    s.close()
    return f_g, f_o


@pytest.mark.parametrize("odr,rk,flux,ctw", SCHEMES)
def test_vortex_parity(gpu, odr, rk, flux, ctw):
    f0, p = P.vortex_2d(64, odr)
    f_g, f_o = run_pair(f0, p, odr, rk, flux, ctw, 6)
    assert_same(f_g, f_o, "vortex odr=%d rk=%d %s ctw=%d" % (odr, rk, flux, ctw))


@pytest.mark.parametrize("flux", ["mlau", "hlld", "lhlld", "roe"])
def test_blast_parity(gpu, flux):
    f0, p = P.blast_2d(48, 300.0, 0.7, 0.6)
    f_g, f_o = run_pair(f0, p, 2, 3, flux, 1, 8)
    assert_same(f_g, f_o, "blast " + flux)


def test_nonsquare_parity(gpu):
    """nx != ny with a grid wider than one x tile (63 faces) and one y chunk."""
    f0, p = P.blast_2d(40, 50.0, 0.5, 1.1)
    g, _ = P.blast_2d(140, 50.0, 0.5, 1.1)
    f0 = np.ascontiguousarray(g[:, :48, :])  # 48 rows, 148 columns
    P.periodic_fill(f0)
    f_g, f_o = run_pair(f0, p, 4, 3, "lhlld", 1, 4)
    assert_same(f_g, f_o, "non-square")


WALLS = {  # boundary flag sets used by the reference drivers (2D/MPI/*/global.hpp)
    "neumann": ((1,) * 8, (1,) * 8),
    "khi": ((0,) * 8, (1, 1, -1, 1, 1, 1, -1, 1)),   # 2D/MPI/KHI/global.hpp:21-22
    "rmi": ((0,) * 8, (1,) * 8),                     # 2D/MPI/RMI/global.hpp:21-22
}


@pytest.mark.parametrize("name", sorted(WALLS))
def test_wall_boundaries(gpu, name):
    """Neumann / Dirichlet ghost rules (boundary.c:36-48,58-72) incl. stagger."""
    dnx, dny = WALLS[name]
    f0, p = P.blast_2d(40, 100.0, 0.6, 0.4)
    for n in range(8):  # make the ghosts consistent with the wall rule first
        O.lib().oc_bc2d(O._p(f0[n]), f0.shape[2], f0.shape[1], 4, 4, int(n == 5), dnx[n], int(n == 6), dny[n])
    dt = 0.4 * p["dx"] / O.cfl_vmax_2d(f0, 4, p["gamma"], p["vfloor"])
    f_o = f0.copy()
    for _ in range(5):
        O.step_2d(f_o, dt, p["dx"], p["dy"], 4, p["gamma"], O.scheme(2, 3, "hlld", 1), dn=dnx, dny=dny)
    s = M.Solver(f0, M.scheme(2, 3, "hlld", 1, dnxs=dnx, dnys=dny), p["dx"], p["dy"], gamma=p["gamma"])
    s.set_dt(dt)
    s.step(5)
    assert_same(s.download(), f_o, "walls " + name)


def test_cfl_loop(gpu):
    """Device-side cflcheck after every step (2D/SERIAL/blast/main.cpp:20-35)."""
    f0, p = P.vortex_2d(64, 4)
    sch_o = O.scheme(4, 3, "lhlld", 1)
    f_o = f0.copy()
    dt = 0.4 * p["dx"] / O.cfl_vmax_2d(f_o, 4, p["gamma"], 1.0)
    dts = []
    for _ in range(10):
        O.step_2d(f_o, dt, p["dx"], p["dy"], 4, p["gamma"], sch_o)
        dts.append(dt)
        vm = O.cfl_vmax_2d(f_o, 4, p["gamma"], 1.0)
        if np.isfinite(vm):
            dt = 0.4 * p["dx"] / vm
    s = M.Solver(f0, M.scheme(4, 3, "lhlld", 1), p["dx"], p["dy"], gamma=p["gamma"])
    dt0, _ = s.cfl(0.4, 1.0)
    assert dt0 == dts[0]
    s.step_cfl(10, 0.4, 1.0)
    tim, dtg, n = s.time()
    assert n == 10
    assert dtg == dt
    assert tim == sum(dts[:1] + dts[1:]) or abs(tim - sum(dts)) < 1e-15
    assert_same(s.download(), f_o, "cfl loop")


def test_dropin_one_step(gpu):
    """mhd_fd_ct_2d with the reference argument list, arrays in place."""
    f0, p = P.vortex_2d(32, 2)
    f_g = f0.copy()
    f_o = f0.copy()
    dt = 1e-3
    sch = M.scheme(2, 3, "lhlld", 1)
    ny, nx = f0.shape[1:]
    for _ in range(3):
        M.mhd_fd_ct_2d(*[f_g[n] for n in range(8)], dt, p["dx"], p["dy"], nx, ny, 4, 4, p["gamma"], sch)
        O.step_2d(f_o, dt, p["dx"], p["dy"], 4, p["gamma"], O.scheme(2, 3, "lhlld", 1))
    assert_same(f_g, f_o, "drop-in")


def test_reference_named_shim(gpu):
    """libmlau_dropin's C++ mhd_fd_ct_2d (reference name and signature) with the
    boundary flags read from the driver-style globals dnxs/dnys."""
    import ctypes as C
    import os
    L = C.CDLL(os.path.join(os.path.dirname(M.LIB_PATH), "libmlau_dropin.so"))
    fn = getattr(L, "_Z12mhd_fd_ct_2dPdS_S_S_S_S_S_S_dddiiiid")
    dp = C.POINTER(C.c_double)
    fn.argtypes = [dp] * 8 + [C.c_double] * 3 + [C.c_int] * 4 + [C.c_double]
    fn.restype = None
    L.mlau_dropin_set_scheme.argtypes = [C.c_int, C.POINTER(M.Scheme)]
    assert L.mlau_dropin_set_scheme(2, C.byref(M.scheme(4, 3, "mlau", 1))) == 0
    dny = (C.c_int * 8).in_dll(L, "dnys")
    flags = (1, 1, -1, 1, 1, 1, -1, 1)
    for n in range(8):
        dny[n] = flags[n]
    f0, p = P.blast_2d(24, 30.0, 0.5, 0.3)
    for n in range(8):
        O.lib().oc_bc2d(O._p(f0[n]), f0.shape[2], f0.shape[1], 4, 4, int(n == 5), 0, int(n == 6), flags[n])
    f_g, f_o = f0.copy(), f0.copy()
    ny, nx = f0.shape[1:]
    for _ in range(3):
        fn(*[f_g[n].ctypes.data_as(dp) for n in range(8)], 2e-3, p["dx"], p["dy"], nx, ny, 4, 4, p["gamma"])
        O.step_2d(f_o, 2e-3, p["dx"], p["dy"], 4, p["gamma"], O.scheme(4, 3, "mlau", 1), dn=(0,) * 8, dny=flags)
    for n in range(8):
        dny[n] = 0
    assert_same(f_g, f_o, "reference-named shim")


@pytest.mark.parametrize("tune", [1, 2, 4, 5, 16, 32, 33, 36, 64, 96])
def test_launch_variants_parity(gpu, monkeypatch, tune):
    """Launch-shape variants give the oracle's values, periodic and with
    walls: bits 0-1 choose the y-sweep chunk (16 or 4 rows); with walls (and
    on periodic grids with bit 5 or 4) the stage runs the primitive and
    auxiliary passes (bit 2: as one pass; bit 4: a separate boundary pass),
    otherwise the sweeps form the primitives themselves; bit 6 keeps the y
    sweep on the handle's stream instead of a side stream beside the x sweep."""
    monkeypatch.setenv("MLAU_TUNE", str(tune))
    f0, p = P.vortex_2d(64, 4)
    f_g, f_o = run_pair(f0, p, 4, 3, "lhlld", 1, 4)
    assert_same(f_g, f_o, "vortex tune=%d" % tune)
    dnx, dny = WALLS["khi"]
    f0, p = P.blast_2d(40, 100.0, 0.6, 0.4)
    for n in range(8):
        O.lib().oc_bc2d(O._p(f0[n]), f0.shape[2], f0.shape[1], 4, 4, int(n == 5), dnx[n], int(n == 6), dny[n])
    dt = 0.4 * p["dx"] / O.cfl_vmax_2d(f0, 4, p["gamma"], p["vfloor"])
    f_o = f0.copy()
    for _ in range(4):
        O.step_2d(f_o, dt, p["dx"], p["dy"], 4, p["gamma"], O.scheme(2, 3, "mlau", 1), dn=dnx, dny=dny)
    s = M.Solver(f0, M.scheme(2, 3, "mlau", 1, dnxs=dnx, dnys=dny), p["dx"], p["dy"], gamma=p["gamma"])
    s.set_dt(dt)
    s.step(4)
    assert_same(s.download(), f_o, "khi walls tune=%d" % tune)


@pytest.mark.parametrize("dim", [2, 3])
def test_step_graph_replays_equal_eager(gpu, monkeypatch, dim):
    """Step loops of 8+ steps replay a captured HIP graph; three replays in a
    row, with host synchronisation and downloads between them, equal the
    eager launches of the same steps (state and time)."""
    if dim == 2:
        f0, p = P.vortex_2d(64, 4)
        sch = M.scheme(4, 3, "lhlld", 1)
    else:
        f0, p = P.ot_vortex_3d(16)
        sch = M.scheme(2, 3, "hlld", 1)

    def run():
        s = M.Solver(f0, sch, p["dx"], p.get("dy"), p.get("dz"), gamma=p["gamma"])
        s.cfl(0.4, p["vfloor"])
        out = []
        for _ in range(3):
            s.step_cfl(8, 0.4, p["vfloor"])
            s.sync()
            out.append((s.download(), s.time()))
        s.step(9)
        out.append((s.download(), s.time()))
        s.close()
        return out

    graph = run()
    monkeypatch.setenv("MLAU_NO_GRAPH", "1")
    eager = run()
    for k, ((a, ta), (b, tb)) in enumerate(zip(graph, eager)):
        assert ta == tb
        assert_same(a, b, "%dD replay %d" % (dim, k))


@pytest.mark.parametrize("tune", [0, 1, 2])
@pytest.mark.parametrize("scheme", [(4, 3, "lhlld", 1), (4, 3, "mlau", 0), (2, 2, "hlld", 1), (3, 1, "roe", 1)])
def test_inline_primitive_stage_tiles(gpu, monkeypatch, tune, scheme):
    """The periodic stage whose sweeps form the primitives themselves, on a
    non-square grid (interior 520 x 300: partial wavefronts, blocks and row
    chunks) with y-sweep chunks of 8, 16 and 4 rows: the oracle's values."""
    monkeypatch.setenv("MLAU_TUNE", str(tune))
    odr, rk, flux, ctw = scheme
    g, p = P.vortex_2d(512, odr)
    f0 = np.ascontiguousarray(g[:, :308, :528])
    dt = 2e-4
    f_o = f0.copy()
    for _ in range(2):
        O.step_2d(f_o, dt, p["dx"], p["dy"], 4, p["gamma"], O.scheme(odr, rk, flux, ctw))
    s = M.Solver(f0, M.scheme(odr, rk, flux, ctw), p["dx"], p["dy"], gamma=p["gamma"])
    s.set_dt(dt)
    s.step(2)
    assert_same(s.download(), f_o, "inline-primitive stage %s tune=%d" % (scheme, tune))
    s.close()
