"""CPU checks of the oracle (the restatement the GPU path is tested against).

Pinned three ways: (1) bit-for-bit against golden fixtures produced by the
reference's own sources (tests/golden/gen_golden.py); (2) against the live
reference builds in oracle/_ref when this container has them; (3) against
closed forms and invariants that a transcription slip would break.
"""
import glob
import math
import os

import numpy as np
import pytest

from mlau_amd import problems as P
from oracle import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


# ----------------------------------------------------------- golden: kernels
def test_flux_kat():
    d = load("kat_flux.npz")
    for k in range(4):
        for t in range(d["q"].shape[0]):
            out = O.flux(k, d["q"][t], d["bn"][t], d["gamma"][t], d["dv"][t])
            assert np.array_equal(bits(out), bits(d["out"][k, t])), (k, t)


def test_recon_kat():
    d = load("kat_recon.npz")
    for kind in range(2):
        for odr in range(1, 5):
            for t in range(d["st"].shape[0]):
                ul, ur = O.reconst(kind, odr, d["st"][t], d["bn"][t], 5 / 3)
                assert np.array_equal(bits(ul), bits(d["ul"][kind, odr - 1, t]))
                assert np.array_equal(bits(ur), bits(d["ur"][kind, odr - 1, t]))


def test_interp_kat():
    import ctypes as C
    d = load("kat_interp.npz")
    L = O.lib()
    a, b = C.c_double(), C.c_double()
    fns = [lambda x: L.oc_wcns4(*x, C.byref(a), C.byref(b)), lambda x: L.oc_wcns3(*x[1:4], C.byref(a), C.byref(b)),
           lambda x: L.oc_muscl_mm(*x[1:4], C.byref(a), C.byref(b)),
           lambda x: L.oc_muscl_mc(*x[1:4], C.byref(a), C.byref(b)),
           lambda x: L.oc_muscl_kr(*x[1:4], C.byref(a), C.byref(b))]
    for w, fn in enumerate(fns):
        for t in range(d["f"].shape[0]):
            fn(tuple(d["f"][t]))
            assert (a.value, b.value) == tuple(d["out"][w, t]) or np.array_equal(
                bits(np.array([a.value, b.value])), bits(d["out"][w, t])), (w, t)


def test_eigen_kat():
    d = load("kat_eigen.npz")
    for t in range(d["p"].shape[0]):
        al, L, R = np.zeros(7), np.zeros(49), np.zeros(49)
        O.lib().oc_eigen_st(O._p(np.ascontiguousarray(d["p"][t])), 5 / 3, O._p(al), O._p(L), O._p(R))
        assert np.array_equal(bits(al), bits(d["al0"][t]))
        assert np.array_equal(bits(L), bits(d["L0"][t])) and np.array_equal(bits(R), bits(d["R0"][t]))
        O.lib().oc_eigen_ae(O._p(np.ascontiguousarray(d["p"][t])), 5 / 3, O._p(L), O._p(R))
        assert np.array_equal(bits(L), bits(d["L1"][t])) and np.array_equal(bits(R), bits(d["R1"][t]))


# ------------------------------------------------------------- golden: steps
def _scheme_of(name):
    # names like o4r3_lhlld[_ctw0] or shock_o2r3_mlau / wave_o4r3_mlau
    parts = name.split("_")
    i = [k for k, p in enumerate(parts) if p.startswith("o") and "r" in p and p[1].isdigit()][0]
    odr, rk = int(parts[i][1]), int(parts[i][3])
    flux = parts[i + 1]
    ctw = 0 if name.endswith("ctw0") else 1
    return odr, rk, flux, ctw


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(G, "step1d_*.npz"))), ids=os.path.basename)
def test_step1d_golden(path):
    d = load(path)
    name = os.path.basename(path)[7:-4]
    odr, rk, flux, _ = _scheme_of(name.replace("brio_wu", "shock"))
    f = d["f0"].copy()
    for _ in range(int(d["nsteps"])):
        O.step_1d(f, float(d["bx"]), float(d["dt"]), float(d["dx"]), 4, float(d["gamma"]), O.scheme(odr, rk, flux),
                  periodic=name.startswith("wave"))
    assert np.array_equal(bits(f), bits(d["f"]))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(G, "step2d_*.npz"))), ids=os.path.basename)
def test_step2d_golden(path):
    d = load(path)
    odr, rk, flux, ctw = _scheme_of(os.path.basename(path)[7:-4].split("_", 1)[1])
    f = d["f0"].copy()
    dn = tuple(int(x) for x in d["dn"])
    for _ in range(int(d["nsteps"])):
        O.step_2d(f, float(d["dt"]), float(d["dx"]), float(d["dy"]), 4, float(d["gamma"]),
                  O.scheme(odr, rk, flux, ctw), dn=dn)
    assert np.array_equal(bits(f), bits(d["f"]))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(G, "step3d_*.npz"))), ids=os.path.basename)
def test_step3d_golden(path):
    d = load(path)
    odr, rk, flux, ctw = _scheme_of(os.path.basename(path)[7:-4].split("_", 1)[1])
    f = d["f0"].copy()
    dn = tuple(int(x) for x in d["dn"])
    for _ in range(int(d["nsteps"])):
        O.step_3d(f, float(d["dt"]), float(d["dx"]), float(d["dy"]), float(d["dz"]), 4, float(d["gamma"]),
                  O.scheme(odr, rk, flux, ctw), dn=dn)
    assert np.array_equal(bits(f), bits(d["f"]))


def test_step3d_walls_golden_covers_every_flag():
    """The 3D wall fixture runs Neumann and Dirichlet flags on all three axes
    (oc_bc3d_state, the single-rank path of 3D/MPI/OTvortex/boundary.c:227-309)."""
    d = load("step3d_walls_o4r3_mlau.npz")
    dn = list(d["dn"])
    assert 1 in dn and -1 in dn and d["f0"].shape[1:] == (13, 15, 20)
    assert not np.array_equal(d["f0"], d["f"])


def test_cfl2d_golden():
    d = load("cfl2d.npz")
    f = np.ascontiguousarray(d["f"])
    for vfloor, dt in zip(d["vfloors"], d["dt"]):
        vm = O.cfl_vmax_2d(f, 4, float(d["gamma"]), float(vfloor))
        assert 0.4 * float(d["dr"]) / vm == dt


def test_cfl3d_golden():
    """oc_cfl_vmax_3d against the 3D driver's own cflcheck
    (3D/MPI/OTvortex/cflcheck.hpp:1-39) on seeded 58^3 states, one with a cell
    whose fast speed is NaN (skipped by the reference's "vtmp > vmax")."""
    from golden.states import cfl3d_state, sha256
    d = load("cfl3d.npz")
    n = int(d["n"])
    cache = {}
    for seed, nan_cell, vfloor, dt, sha in zip(d["seed"], d["nan_cell"], d["vfloors"], d["dt"], d["sha256"]):
        key = (int(seed), bool(nan_cell))
        if key not in cache:
            cache[key] = cfl3d_state(key[0], n, nan_cell=key[1])
        f = cache[key]
        assert sha256(f) == str(sha), "rebuilt input differs from the one the reference saw"
        vm = O.cfl_vmax_3d(f, 4, float(d["gamma"]), float(vfloor))
        assert np.isfinite(vm)
        assert 0.4 * float(d["dr"]) / vm == dt


def test_cfl3d_closed_form():
    """Uniform state with vz != 0: vmax = |v| + sqrt((gamma p + B^2)/rho)
    (3D/MPI/OTvortex/cflcheck.hpp:23 uses sqrt(v2) with vz included; the 2D
    rule sqrt(vx^2+vy^2) would give a smaller value)."""
    n, g = 12, 5.0 / 3.0
    ro, v, b, pr = 2.0, (0.5, -0.25, 3.0), (0.75, 0.5, -1.0), 4.0
    f = np.empty((8, n, n, n))
    f[0] = ro
    f[1:4] = np.array(v)[:, None, None, None] * ro
    f[5:8] = np.array(b)[:, None, None, None]
    v2, b2 = sum(x * x for x in v), sum(x * x for x in b)
    f[4] = pr / (g - 1.0) + 0.5 * ro * v2 + 0.5 * b2
    vm = O.cfl_vmax_3d(f, 4, g, 0.0)
    want = math.sqrt(v2) + math.sqrt((g * pr + b2) / ro)
    assert abs(vm - want) <= 1e-14 * want
    assert vm > math.sqrt(v[0] ** 2 + v[1] ** 2) + math.sqrt((g * pr + b2) / ro) + 1.0


def test_dt_1d_readme():
    """1D/SERIAL/README.md:37-54 lists DT=0.0100509 (float32) for the default
    Dai-Woodward run (NUM=1, XMESH=800, gamma=5/3): DT = nrec*dt with
    nrec=(int)(0.01/dt+0.5) (1D/SERIAL/shock/main.cpp:113-129)."""
    s4 = math.sqrt(4 * math.pi)
    f, p = P.riemann_1d(800, (1.08, 0.95, 1.2, 0.01, 0.5, 3.6 / s4, 2.0 / s4), (1.0, 1.0, 0, 0, 0, 4.0 / s4, 2.0 / s4),
                        2.0 / s4, 5.0 / 3.0)
    dt = 0.4 * p["dx"] / O.vmax_1d(f, p["bx"], p["gamma"], 1.0)
    nrec = int(0.01 / dt + 0.5)
    assert nrec == 59
    assert round(nrec * dt, 7) == 0.0100509  # IDL prints 6 significant digits


# -------------------------------------------------------- live reference pin
ref = pytest.mark.skipif(not O.ref_available(), reason="reference tree not built here (oracle/_ref)")


@ref
def test_live_reference_random_fluxes():
    K = O.ref_lib("ref_kat")
    rng = np.random.default_rng(7)
    for t in range(2000):
        q = np.concatenate([[10 ** rng.uniform(-2, 1)], rng.normal(0, 2, 5), [10 ** rng.uniform(-2, 2)],
                            [10 ** rng.uniform(-2, 1)], rng.normal(0, 2, 5), [10 ** rng.uniform(-2, 2)]])
        bn, dv = rng.normal(), rng.normal(0, 1, 2)
        for k in range(4):
            assert np.array_equal(bits(O.flux(k, q, bn, 5 / 3, dv)), bits(O.flux(k, q, bn, 5 / 3, dv, K)))


# ------------------------------------------------------- closed forms / math
def phys_flux(q, bn, gamma):
    ro, vn, vt, vu, bt, bu, pr = q
    pt = pr + 0.5 * (bn * bn + bt * bt + bu * bu)
    en = pr / (gamma - 1) + 0.5 * ro * (vn * vn + vt * vt + vu * vu) + 0.5 * (bn * bn + bt * bt + bu * bu)
    vb = vn * bn + vt * bt + vu * bu
    return np.array([ro * vn, ro * vn * vn + pt - bn * bn, ro * vn * vt - bn * bt, ro * vn * vu - bn * bu,
                     vn * bt - vt * bn, vn * bu - vu * bn, (en + pt) * vn - bn * vb])


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_flux_consistency(kind):
    """F(U,U) = physical flux for Roe, HLLD, LHLLD (round-off only)."""
    rng = np.random.default_rng(3)
    for t in range(300):
        s = np.array([10 ** rng.uniform(-1, 1), *rng.normal(0, 1, 3), *rng.normal(0, 1, 2), 10 ** rng.uniform(-1, 1)])
        bn = rng.normal()
        f = O.flux(kind, np.concatenate([s, s]), bn, 5 / 3, np.zeros(2))
        g = phys_flux(s, bn, 5 / 3)
        scale = 1.0 + np.abs(g).max()
        assert np.abs(f - g).max() < 1e-12 * scale, (kind, t, f - g)


def test_lhlld_reduces_to_hlld():
    """theta=1 (dv>=0) and mcc=1 (convective speed above the fast speed) make
    LHLLD equal HLLD up to one reassociation (include/mhd_flux.h:640 vs :505)."""
    rng = np.random.default_rng(5)
    n = 0
    for t in range(4000):
        ql = np.array([1.0, 8 + rng.normal(), *rng.normal(0, .3, 2), *rng.normal(0, .3, 2), 0.5 + rng.random()])
        qr = ql * (1 + 0.05 * rng.normal(size=7))
        q = np.concatenate([ql, qr])
        bn = 0.3 * rng.normal()
        a = O.flux(3, q, bn, 5 / 3, np.array([0.1, 0.2]))
        b = O.flux(2, q, bn, 5 / 3, np.array([0.1, 0.2]))
        assert np.abs(a - b).max() <= 1e-13 * (1 + np.abs(b).max())
        n += 1
    assert n > 0


def test_eigen_st_inverse():
    """L R = I for the Athena eigensystem (include/mhd_eigen.h:24-192)."""
    rng = np.random.default_rng(11)
    for t in range(200):
        p = np.array([10 ** rng.uniform(-1, 1), *rng.normal(0, 1, 3), *rng.normal(0, 1, 3), 10 ** rng.uniform(-1, 1)])
        al, L, R = np.zeros(7), np.zeros(49), np.zeros(49)
        O.lib().oc_eigen_st(O._p(p), 5 / 3, O._p(al), O._p(L), O._p(R))
        LR = L.reshape(7, 7) @ R.reshape(7, 7).T
        assert np.abs(LR - np.eye(7)).max() < 1e-10


def test_wcns4_linear_exact():
    """WCNS4 reproduces linear data (the smoothness indicators vanish and the
    linear weights sum to 1): fl at x=2.5, fr at x=1.5."""
    import ctypes as C
    a, b = C.c_double(), C.c_double()
    for s in (0.3, -2.0, 7.5):
        f = [1.0 + s * k for k in range(5)]
        O.lib().oc_wcns4(*f, C.byref(a), C.byref(b))
        assert abs(a.value - (1 + 2.5 * s)) < 1e-13 and abs(b.value - (1 + 1.5 * s)) < 1e-13
        O.lib().oc_muscl_mm(*f[1:4], C.byref(a), C.byref(b))
        assert abs(a.value - (1 + 2.5 * s)) < 1e-13 and abs(b.value - (1 + 1.5 * s)) < 1e-13


def test_uniform_state_stays_uniform():
    f, p = P.blast_2d(16, 1.0, 0.5, 0.3)
    f[1] = 0.3
    f[2] = -0.2
    f[4] += 0.5 * (0.09 + 0.04)
    g = f.copy()
    for sch in (O.scheme(4, 3, "lhlld"), O.scheme(2, 3, "mlau"), O.scheme(3, 3, "roe"), O.scheme(2, 3, "hlld")):
        h = f.copy()
        O.step_2d(h, 1e-3, p["dx"], p["dy"], 4, p["gamma"], sch)
        assert np.abs(h - g).max() < 1e-13


def test_conservation_and_divb():
    """Periodic grids conserve mass, momentum and energy to round-off, and CT
    keeps the ODR-matched div B at round-off (mhd_fd_ct_2d.c:74-82)."""
    f, p = P.vortex_2d(32, 4)
    tot0 = f[[0, 1, 2, 4], 4:-4, 4:-4].sum(axis=(1, 2))

    def divb(f):
        bx, by = f[5], f[6]
        d = lambda a, ax: (27.0 * (np.roll(a, -1, ax) - a) - (np.roll(a, -2, ax) - np.roll(a, 1, ax))) / 24.0
        return (d(bx, 1) / p["dx"] + d(by, 0) / p["dy"])[4:-4, 4:-4]

    assert np.abs(divb(f)).max() < 1e-10
    for _ in range(10):
        dt = 0.4 * p["dx"] / O.cfl_vmax_2d(f, 4, p["gamma"], 1.0)
        O.step_2d(f, dt, p["dx"], p["dy"], 4, p["gamma"], O.scheme(4, 3, "lhlld"))
    tot = f[[0, 1, 2, 4], 4:-4, 4:-4].sum(axis=(1, 2))
    assert np.abs(tot - tot0).max() < 1e-11 * np.abs(tot0).max()
    assert np.abs(divb(f)).max() < 1e-10


def test_splitmix64_known_answer():
    """splitmix64 reference vector (seed 1234567)."""
    with np.errstate(over="ignore"):
        s = np.uint64(1234567)
        out = []
        for _ in range(3):
            s = np.uint64(s + np.uint64(0x9E3779B97F4A7C15))
            z = s
            z = np.uint64((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9))
            z = np.uint64((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB))
            out.append(int(z ^ (z >> np.uint64(31))))
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423]
    u = P.splitmix64_vec(1234567, 3)
    assert np.array_equal(u, np.array(out, dtype=np.float64) // 2048 * 2.0 ** -53) or np.allclose(
        u, [(x >> 11) * 2.0 ** -53 for x in out], rtol=0, atol=0)


def test_workload_draws_are_seeded_and_in_range():
    """Ensemble and serving-stream parameters: deterministic per seed, grid
    sizes multiples of 16 within the stated ranges, every kind present."""
    from mlau_amd import problems as P
    a, b = P.blast_ensemble_draws(512, seed=3), P.blast_ensemble_draws(512, seed=3)
    assert a == b and a != P.blast_ensemble_draws(512, seed=4)
    assert all(d["n"] % 16 == 0 and 64 <= d["n"] <= 512 for d in a)
    assert all(10 <= d["pratio"] < 1000 and 0.1 <= d["bmag"] < 1.0 for d in a)
    assert {d["flux"] for d in a} == {"mlau", "hlld"}
    s = P.serve_request_draws(300, seed=4)
    assert s == P.serve_request_draws(300, seed=4)
    kinds = {d["kind"] for d in s}
    assert kinds == {1, 2, 3}
    for d in s:
        if d["kind"] == 1:
            assert 256 <= d["nx"] <= 8192
        elif d["kind"] == 2:
            assert d["n"] % 16 == 0 and 64 <= d["n"] <= 1024
        else:
            assert d["n"] % 16 == 0 and 32 <= d["n"] <= 192 and 1e-3 <= d["mach"] < 0.1
