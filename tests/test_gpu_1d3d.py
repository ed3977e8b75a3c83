"""1D and 3D parity of the HIP path against the oracle (value for value)."""
import numpy as np
import pytest

import mlau_amd as M
from mlau_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def assert_same(a, b, what):
    assert np.all(np.isfinite(b)), what + ": oracle produced non-finite values"
    bad = a != b
    if bad.any():
        idx = np.argwhere(bad)[:5]
        raise AssertionError("%s: %d mismatches, first %s gpu=%r oracle=%r maxabs=%g" % (
            what, bad.sum(), idx.tolist(), a[tuple(idx[0])], b[tuple(idx[0])], np.abs(a - b).max()))


S1 = [(2, 3, "mlau", "open"), (1, 2, "lhlld", "open"), (4, 3, "lhlld", "open"), (3, 3, "hlld", "open"),
      (2, 3, "roe", "open"), (2, 3, "lhlld", "periodic"), (4, 3, "mlau", "periodic"), (1, 2, "hlld", "periodic")]


@pytest.mark.parametrize("odr,rk,flux,bc", S1)
def test_brio_wu_parity(gpu, odr, rk, flux, bc):
    f0, p = P.brio_wu(256)
    dt = 0.4 * p["dx"] / O.vmax_1d(f0, p["bx"], p["gamma"])
    f_o = f0.copy()
    for _ in range(60):
        O.step_1d(f_o, p["bx"], dt, p["dx"], 4, p["gamma"], O.scheme(odr, rk, flux, 1), periodic=(bc == "periodic"))
    s = M.Solver(f0, M.scheme(odr, rk, flux, 1, bc1d=bc), p["dx"], gamma=p["gamma"], bx=p["bx"])
    dtg, _ = s.cfl(0.4, 1.0)
    assert dtg == dt
    s.step(60)
    assert_same(s.download(), f_o, "1D odr=%d rk=%d %s %s" % (odr, rk, flux, bc))


def test_brio_wu_full_run(gpu):
    """1D Brio-Wu shock tube, nx=1024, MLAU + MUSCL + SSP-RK3 to the reference driver's t=0.1 output."""
    f0, p = P.brio_wu(1024)
    dt = 0.4 * p["dx"] / O.vmax_1d(f0, p["bx"], p["gamma"])
    n = P.brio_wu_steps(dt)
    f_o = f0.copy()
    sch = O.scheme(2, 3, "mlau", 1)
    for _ in range(n):
        O.step_1d(f_o, p["bx"], dt, p["dx"], 4, p["gamma"], sch)
    s = M.Solver(f0, M.scheme(2, 3, "mlau", 1), p["dx"], gamma=p["gamma"], bx=p["bx"])
    s.set_dt(dt)
    s.step(n)
    assert_same(s.download(), f_o, "Brio-Wu full run")


S3 = [(4, 3, "mlau", "shear"), (2, 3, "lhlld", "ot"), (2, 3, "hlld", "ot"), (3, 3, "roe", "ot"), (4, 3, "lhlld", "ot"),
      (1, 2, "mlau", "ot")]


@pytest.mark.parametrize("odr,rk,flux,prob", S3)
def test_3d_parity(gpu, odr, rk, flux, prob):
    f0, p = P.shear_layer_3d(16, 0.05) if prob == "shear" else P.ot_vortex_3d(16)
    dt = 0.4 * p["dx"] / O.cfl_vmax_3d(f0, 4, p["gamma"], p["vfloor"])
    f_o = f0.copy()
    for _ in range(3):
        O.step_3d(f_o, dt, p["dx"], p["dy"], p["dz"], 4, p["gamma"], O.scheme(odr, rk, flux, 1))
    s = M.Solver(f0, M.scheme(odr, rk, flux, 1), p["dx"], p["dy"], p["dz"], gamma=p["gamma"])
    dtg, _ = s.cfl(0.4, p["vfloor"])
    assert dtg == dt
    s.step(3)
    assert_same(s.download(), f_o, "3D %s odr=%d rk=%d %s" % (prob, odr, rk, flux))


def test_3d_noncubic(gpu):
    """nx != ny != nz with x wider than one 63-face tile and z longer than one chunk."""
    f0, p = P.ot_vortex_3d(16)
    g, _ = P.ot_vortex_3d(72)
    f0 = np.ascontiguousarray(g[:, :28, :20, :])  # nz=28, ny=20, nx=80
    P.periodic_fill(f0)
    dt = 1e-3
    f_o = f0.copy()
    for _ in range(2):
        O.step_3d(f_o, dt, p["dx"], p["dy"], p["dz"], 4, p["gamma"], O.scheme(4, 3, "lhlld", 1))
    s = M.Solver(f0, M.scheme(4, 3, "lhlld", 1), p["dx"], p["dy"], p["dz"], gamma=p["gamma"])
    s.set_dt(dt)
    s.step(2)
    assert_same(s.download(), f_o, "3D non-cubic")


def test_3d_cfl_loop(gpu):
    f0, p = P.ot_vortex_3d(16)
    f_o = f0.copy()
    sch = O.scheme(2, 3, "lhlld", 1)
    dt = 0.4 * p["dx"] / O.cfl_vmax_3d(f_o, 4, p["gamma"], 1.0)
    for _ in range(4):
        O.step_3d(f_o, dt, p["dx"], p["dy"], p["dz"], 4, p["gamma"], sch)
        dt = 0.4 * p["dx"] / O.cfl_vmax_3d(f_o, 4, p["gamma"], 1.0)
    s = M.Solver(f0, M.scheme(2, 3, "lhlld", 1), p["dx"], p["dy"], p["dz"], gamma=p["gamma"])
    s.cfl(0.4, 1.0)
    s.step_cfl(4, 0.4, 1.0)
    assert s.time()[1] == dt
    assert_same(s.download(), f_o, "3D cfl loop")


@pytest.mark.parametrize("tune", [0, 8, 64])
def test_3d_edge_update_variants(gpu, monkeypatch, tune):
    """z-marching (default) and plane-parallel edge/update kernels both give the
    oracle's values on a non-cubic grid whose z extent is not a multiple of
    the marching chunk, with the 4th-order difference stencils; with tune bit
    6 the y and z sweeps run on the handle's stream instead of side streams."""
    monkeypatch.setenv("MLAU_TUNE", str(tune))
    g, p = P.ot_vortex_3d(72)
    f0 = np.ascontiguousarray(g[:, :29, :21, :])  # nz=29, ny=21, nx=80
    P.periodic_fill(f0)
    dt = 1e-3
    f_o = f0.copy()
    for _ in range(2):
        O.step_3d(f_o, dt, p["dx"], p["dy"], p["dz"], 4, p["gamma"], O.scheme(4, 3, "mlau", 1))
    s = M.Solver(f0, M.scheme(4, 3, "mlau", 1), p["dx"], p["dy"], p["dz"], gamma=p["gamma"])
    s.set_dt(dt)
    s.step(2)
    assert_same(s.download(), f_o, "3D tune=%d" % tune)


@pytest.mark.parametrize("tune", [0, 8, 64])
def test_3d_walls(gpu, monkeypatch, tune):
    """Wall flags with sign flips on all three axes (the mirror rules of
    3D/MPI/OTvortex/boundary.c:227-309, applied to every axis) on a non-cubic
    grid give the oracle's values, including the mirrored face-B centring and
    velocity differences of the ghost cells, with the z-marching (default) and
    the plane-parallel edge/update kernels, and with the y and z sweeps on the
    handle's stream (tune bit 6) instead of side streams."""
    monkeypatch.setenv("MLAU_TUNE", str(tune))
    g, p = P.ot_vortex_3d(72)
    f0 = np.ascontiguousarray(g[:, :29, :21, :])  # nz=29, ny=21, nx=80
    dn = (1, -1, 1, -1, 1, 1, -1, 1)
    O.bc3d_state(f0, 4, dn)
    dt = 1e-3
    f_o = f0.copy()
    for _ in range(2):
        O.step_3d(f_o, dt, p["dx"], p["dy"], p["dz"], 4, p["gamma"], O.scheme(4, 3, "mlau", 1), dn=dn)
    assert np.all(np.isfinite(f_o))
    s = M.Solver(f0, M.scheme(4, 3, "mlau", 1, dnxs=dn), p["dx"], p["dy"], p["dz"], gamma=p["gamma"])
    s.set_dt(dt)
    s.step(2)
    assert_same(s.download(), f_o, "3D walls tune=%d" % tune)


@pytest.mark.parametrize("bc", ["open", "periodic"])
def test_1d_separate_launch_variant(gpu, monkeypatch, bc):
    """MLAU_TUNE bit 3 in 1D: separate flux and update launches in every stage
    (instead of one launch per stage after the first) give the oracle's values."""
    monkeypatch.setenv("MLAU_TUNE", "8")
    f0, p = P.brio_wu(256)
    dt = 0.4 * p["dx"] / O.vmax_1d(f0, p["bx"], p["gamma"])
    f_o = f0.copy()
    for _ in range(30):
        O.step_1d(f_o, p["bx"], dt, p["dx"], 4, p["gamma"], O.scheme(2, 3, "mlau", 1), periodic=(bc == "periodic"))
    s = M.Solver(f0, M.scheme(2, 3, "mlau", 1, bc1d=bc), p["dx"], gamma=p["gamma"], bx=p["bx"])
    s.set_dt(dt)
    s.step(30)
    assert_same(s.download(), f_o, "1D tune=8 " + bc)
