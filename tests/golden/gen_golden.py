#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ from the reference itself.

Every expected output here comes from oracle/_ref/*.so, i.e. the reference's
own sources (include/*.h, the 1D/2D/3D solvers, boundary.c, cflcheck.hpp)
compiled where they lie by oracle/ref_build.sh -- never from this project's
code.  Inputs are seeded (numpy PCG64, seed fixed below) and stored with the
outputs, so the fixtures do not depend on any machine's libm.  Run in the
container that holds the reference tree:

    python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from oracle import oracle as O  # noqa: E402
from mlau_amd import problems as P  # noqa: E402

RNG = np.random.default_rng(20240611)


def rand_states(n):
    ro = 10 ** RNG.uniform(-2, 1, n)
    pr = 10 ** RNG.uniform(-2, 2, n)
    v = RNG.normal(0, 2, (n, 3))
    b = RNG.normal(0, 1.5, (n, 3))
    z = RNG.random((n, 6)) < 0.15  # exact zeros exercise the zero-forcing and degeneracies
    v[z[:, :3]] = 0
    b[z[:, 3:]] = 0
    return ro, v, b, pr


def kat_flux(n=512):
    K = O.ref_lib("ref_kat")
    q = np.empty((n, 14))
    bn = np.empty(n)
    gam = np.array([5 / 3, 2.0, 1.4, 5 / 3])[np.arange(n) % 4]
    dv = RNG.normal(0, 1, (n, 2))
    dv[::5] = 0.0
    for t in range(n):
        (rl, vl, bl, pl), (rr, vr, br, prr) = rand_states(1), rand_states(1)
        if t % 7 == 0:  # F(U,U)
            rr, vr, br, prr = rl, vl, bl, pl
        q[t] = [rl[0], *vl[0], *bl[0, 1:], pl[0], rr[0], *vr[0], *br[0, 1:], prr[0]]
        bn[t] = 0.0 if t % 11 == 0 else bl[0, 0]
    out = np.empty((4, n, 7))
    for k in range(4):
        for t in range(n):
            out[k, t] = O.flux(k, q[t], bn[t], gam[t], dv[t], K)
    return dict(q=q, bn=bn, gamma=gam, dv=dv, out=out)


def kat_recon(n=256):
    K = O.ref_lib("ref_kat")
    st = np.empty((n, 7, 5))
    bn = np.empty(n)
    for t in range(n):
        ro, v, b, pr = rand_states(5)
        st[t] = [ro, v[:, 0], v[:, 1], v[:, 2], b[:, 1], b[:, 2], pr]
        if t % 6 == 0:
            st[t, :, :] = st[t, :, 2:3]
        bn[t] = b[2, 0]
    ul = np.empty((2, 4, n, 7))
    ur = np.empty((2, 4, n, 7))
    for kind in range(2):
        for odr in range(1, 5):
            for t in range(n):
                ul[kind, odr - 1, t], ur[kind, odr - 1, t] = O.reconst(kind, odr, st[t], bn[t], 5 / 3, K)
    return dict(st=st, bn=bn, ul=ul, ur=ur)


def kat_interp(n=512):
    import ctypes as C
    K = O.ref_lib("ref_kat")
    f = RNG.normal(0, 1, (n, 5))
    f[::9] = f[::9, 2:3]
    f[1::9] = np.linspace(-1, 1, 5)
    out = np.empty((5, n, 2))
    a, b = C.c_double(), C.c_double()
    for t in range(n):
        x = f[t]
        K.ref_wcns4(*x, C.byref(a), C.byref(b)); out[0, t] = a.value, b.value
        K.ref_wcns3(*x[1:4], C.byref(a), C.byref(b)); out[1, t] = a.value, b.value
        for w in range(3):
            K.ref_muscl(w, *x[1:4], C.byref(a), C.byref(b)); out[2 + w, t] = a.value, b.value
    return dict(f=f, out=out)


def kat_eigen(n=128):
    K = O.ref_lib("ref_kat")
    p = np.empty((n, 8))
    ro, v, b, pr = rand_states(n)
    p[:, 0], p[:, 1:4], p[:, 4:7], p[:, 7] = ro, v, b, pr
    res = {}
    for kind in range(2):
        al = np.empty((n, 7)); L = np.empty((n, 49)); R = np.empty((n, 49))
        for t in range(n):
            K.ref_eigen(kind, O._p(np.ascontiguousarray(p[t])), 5 / 3, O._p(al[t]), O._p(L[t]), O._p(R[t]))
        res["al%d" % kind], res["L%d" % kind], res["R%d" % kind] = al, L, R
    res["p"] = p
    return res


def golden_1d(name, f0, bx, gamma, dx, nsteps, dt):
    f = f0.copy()
    for _ in range(nsteps):
        O.ref_step_1d("ref1d_" + name, f, bx, dt, dx, 4, gamma)
    return dict(f0=f0, f=f, bx=bx, gamma=gamma, dx=dx, dt=dt, nsteps=nsteps)


def golden_2d(name, f0, p, nsteps, dt, dn=(0,) * 8):
    f = f0.copy()
    for _ in range(nsteps):
        O.ref_step_2d("ref2d_" + name, f, dt, p["dx"], p["dy"], 4, p["gamma"], dn=dn)
    return dict(f0=f0, f=f, gamma=p["gamma"], dx=p["dx"], dy=p["dy"], dt=dt, nsteps=nsteps, dn=np.array(dn))


def golden_3d(name, f0, p, nsteps, dt, dn=(0,) * 8):
    f = f0.copy()
    for _ in range(nsteps):
        O.ref_step_3d("ref3d_" + name, f, dt, p["dx"], p["dy"], p["dz"], 4, p["gamma"], dn=dn)
    return dict(f0=f0, f=f, gamma=p["gamma"], dx=p["dx"], dy=p["dy"], dz=p["dz"], dt=dt, nsteps=nsteps,
                dn=np.array(dn))


def golden_cfl2d():
    """Reference cflcheck (2D/SERIAL/OTvortex/cflcheck.hpp) on a 200^2 grid."""
    import ctypes as C
    L = C.CDLL(os.path.join(O.REF_DIR, "libref_cfl2d.so"))
    L.ref_cfl2d.restype = C.c_double
    L.ref_cfl2d.argtypes = [O._dp] * 8 + [C.c_double, C.c_double]
    L.ref_cfl2d_dr.restype = C.c_double
    f, p = P.vortex_2d(200, 4)
    f[0] *= 1.0 + 0.3 * RNG.random(f[0].shape)
    out = []
    for vfloor in (0.5, 3.0):
        out.append(L.ref_cfl2d(*[O._p(f[n]) for n in range(8)], vfloor, -1.0))
    return dict(f=f, vfloors=np.array([0.5, 3.0]), dt=np.array(out), dr=L.ref_cfl2d_dr(), gamma=5 / 3)


def golden_cfl3d():
    """Reference cflcheck of the 3D driver (3D/MPI/OTvortex/cflcheck.hpp:1-39,
    speed sqrt(v2) + fmode with vz included, MPI_Allreduce over one rank) on
    the 58^3 grid its global.hpp fixes.  The inputs are rebuilt from seeds by
    tests/golden/states.py; only their SHA-256 is stored."""
    import ctypes as C
    from states import cfl3d_state, sha256
    L = C.CDLL(os.path.join(O.REF_DIR, "libref_cfl3d.so"))
    L.ref_cfl3d.restype = C.c_double
    L.ref_cfl3d.argtypes = [O._dp] * 8 + [C.c_double, C.c_double]
    L.ref_cfl3d_dr.restype = C.c_double
    n = L.ref_cfl3d_n() // 1000000
    seeds, nans, vfloors, dts, shas = [], [], [], [], []
    for seed, nan_cell in ((31, False), (32, True)):
        f = cfl3d_state(seed, n, nan_cell=nan_cell)
        for vfloor in (0.5, 9.0):
            seeds.append(seed); nans.append(nan_cell); vfloors.append(vfloor); shas.append(sha256(f))
            dts.append(L.ref_cfl3d(*[O._p(f[k]) for k in range(8)], vfloor, -1.0))
    return dict(seed=np.array(seeds), nan_cell=np.array(nans), vfloors=np.array(vfloors), dt=np.array(dts),
                sha256=np.array(shas), n=n, dr=L.ref_cfl3d_dr(), gamma=5 / 3)


def golden_vortex512_long(nsteps=2000):
    """The 2D vortex scheme (LHLLD, WCNS4, CT, RK3) at the CPU-baseline size 512^2
    for its stated 2000 steps, dt from the CFL rule after every step (the
    driver loop of 2D/SERIAL/OTvortex/main.cpp:26-43 with CFLCHECK on): the
    reference solver steps, the oracle's pinned cflcheck.  The final state is
    too large to store; its SHA-256 (after x + 0.0, which maps -0.0 to +0.0)
    and per-field L1 norms are kept."""
    from states import sha256
    f, p = P.vortex_2d(512, 4)
    dr = min(p["dx"], p["dy"])
    dt = 0.4 * dr / O.cfl_vmax_2d(f, 4, p["gamma"], p["vfloor"])
    tim = 0.0
    for _ in range(nsteps):
        O.ref_step_2d("ref2d_o4r3_lhlld", f, dt, p["dx"], p["dy"], 4, p["gamma"])
        tim += dt
        dt = 0.4 * dr / O.cfl_vmax_2d(f, 4, p["gamma"], p["vfloor"])
    inner = f[:, 4:-4, 4:-4]
    return dict(n=512, nsteps=nsteps, tim=tim, dt=dt, vfloor=p["vfloor"], sha256=sha256(f + 0.0),
                l1=np.abs(inner).reshape(8, -1).sum(axis=1))


def golden_serve1d_8192():
    """The serving stream's largest 1D request (nx=8192, Brio-Wu states, MLAU,
    MUSCL, RK3, open boundaries) to t=0.1 under the 1D driver rule: one dt =
    0.4*dx/vmax from the initial state (oracle vmax_1d, pinned by the
    reference README's dt), the last step clipped at t=0.1; steps by the
    reference solver.  SHA-256 of the final state (x + 0.0) and its L1 norms."""
    from states import sha256
    d = dict(kind=1, nx=8192, rho_l=1.0, p_l=1.0, rho_r=0.125, p_r=0.1, flux="mlau")
    f, p, _ = P.serve_request(d)
    dtc = 0.4 * p["dx"] / O.vmax_1d(f, p["bx"], p["gamma"], 1.0)
    tim, n, tend = 0.0, 0, 0.1
    while True:
        dt = dtc
        last = tim + dt >= tend
        if last:
            dt = tend - tim
        O.ref_step_1d("ref1d_shock_o2r3_mlau", f, p["bx"], dt, p["dx"], 4, p["gamma"])
        tim = tend if last else tim + dt
        n += 1
        if last:
            break
    return dict(nx=8192, nsteps=n, tim=tim, sha256=sha256(f + 0.0), l1=np.abs(f[:, 4:-4]).sum(axis=1))


def main():
    if not O.ref_available():
        sys.exit("gen_golden: oracle/_ref is not built (needs the reference tree); run oracle/ref_build.sh")
    np.savez_compressed(os.path.join(HERE, "kat_flux.npz"), **kat_flux())
    np.savez_compressed(os.path.join(HERE, "kat_recon.npz"), **kat_recon())
    np.savez_compressed(os.path.join(HERE, "kat_interp.npz"), **kat_interp())
    np.savez_compressed(os.path.join(HERE, "kat_eigen.npz"), **kat_eigen())
    np.savez_compressed(os.path.join(HERE, "cfl2d.npz"), **golden_cfl2d())
    np.savez_compressed(os.path.join(HERE, "cfl3d.npz"), **golden_cfl3d())
    # 1D Brio-Wu with MLAU, MUSCL, RK3, the reference driver's t=0.1 output
    f0, p = P.brio_wu(1024)
    dt = 0.4 * p["dx"] / O.vmax_1d(f0, p["bx"], p["gamma"])
    n = P.brio_wu_steps(dt)
    np.savez_compressed(os.path.join(HERE, "step1d_brio_wu_o2r3_mlau.npz"),
                        **golden_1d("shock_o2r3_mlau", f0, p["bx"], p["gamma"], p["dx"], n, dt))
    f0, p = P.brio_wu(128)
    dt = 0.4 * p["dx"] / O.vmax_1d(f0, p["bx"], p["gamma"])
    for name in ("shock_o4r3_lhlld", "shock_o1r2_lhlld", "shock_o3r3_hlld", "shock_o2r3_roe", "wave_o4r3_mlau"):
        np.savez_compressed(os.path.join(HERE, "step1d_%s.npz" % name),
                            **golden_1d(name, f0, p["bx"], p["gamma"], p["dx"], 40, dt))
    # 2D
    f0, p = P.vortex_2d(32, 4)
    dt = 0.4 * p["dx"] / O.cfl_vmax_2d(f0, 4, p["gamma"], 1.0)
    np.savez_compressed(os.path.join(HERE, "step2d_vortex_o4r3_lhlld.npz"), **golden_2d("o4r3_lhlld", f0, p, 12, dt))
    f0, p = P.vortex_2d(32, 2)
    for name in ("o2r3_mlau", "o2r3_roe", "o1r3_lhlld", "o2r3_lhlld_ctw0"):
        np.savez_compressed(os.path.join(HERE, "step2d_vortex_%s.npz" % name), **golden_2d(name, f0, p, 8, dt))
    fb, pb = P.blast_2d(32, 300.0, 0.7, 0.6)
    dtb = 0.4 * pb["dx"] / O.cfl_vmax_2d(fb, 4, pb["gamma"], 1.0)
    np.savez_compressed(os.path.join(HERE, "step2d_blast_o2r3_hlld.npz"), **golden_2d("o2r3_hlld", fb, pb, 10, dtb))
    np.savez_compressed(os.path.join(HERE, "step2d_blast_o2r3_mlau.npz"), **golden_2d("o2r3_mlau", fb, pb, 10, dtb))
    # reflecting walls in y (2D/MPI/KHI/global.hpp flag set applied on both axes)
    fk = fb.copy()
    dn = (1, 1, -1, 1, 1, 1, -1, 1)
    for m in range(8):
        O.lib().oc_bc2d(O._p(fk[m]), fk.shape[2], fk.shape[1], 4, 4, int(m == 5), dn[m], int(m == 6), dn[m])
    np.savez_compressed(os.path.join(HERE, "step2d_walls_o2r3_hlld.npz"),
                        **golden_2d("o2r3_hlld", fk, pb, 6, dtb, dn=dn))
    # 3D
    f3, p3 = P.shear_layer_3d(8, 0.05)
    dt3 = 0.4 * p3["dx"] / O.cfl_vmax_3d(f3, 4, p3["gamma"], 0.0)
    np.savez_compressed(os.path.join(HERE, "step3d_shear_o4r3_mlau.npz"), **golden_3d("o4r3_mlau", f3, p3, 3, dt3))
    f3, p3 = P.ot_vortex_3d(8)
    dt3 = 0.4 * p3["dx"] / O.cfl_vmax_3d(f3, 4, p3["gamma"], 1.0)
    for name in ("o2r3_lhlld", "o2r3_hlld", "o3r3_roe"):
        np.savez_compressed(os.path.join(HERE, "step3d_ot_%s.npz" % name), **golden_3d(name, f3, p3, 3, dt3))
    # walls with sign flips on every axis (3D/MPI/OTvortex/boundary.c:227-309 after the
    # single-rank mpi_sdrv3d, mhd_fd_ct_3d.c:1161-1175) on a non-cubic grid
    g3, p3 = P.ot_vortex_3d(12)
    fw = np.ascontiguousarray(g3[:, :13, :15, :])  # nz=13, ny=15, nx=20
    dn = (1, -1, 1, -1, 1, 1, -1, 1)
    O.bc3d_state(fw, 4, dn)
    np.savez_compressed(os.path.join(HERE, "step3d_walls_o4r3_mlau.npz"),
                        **golden_3d("o4r3_mlau", fw, p3, 2, 2e-3, dn=dn))
    if "--long" in sys.argv:  # minutes of CPU: only when asked
        np.savez_compressed(os.path.join(HERE, "vortex512_2000.npz"), **golden_vortex512_long())
    if "--long" in sys.argv or "--serve1d" in sys.argv:
        np.savez_compressed(os.path.join(HERE, "serve1d_8192.npz"), **golden_serve1d_8192())
    print("gen_golden: fixtures written to", HERE)


if __name__ == "__main__":
    main()
