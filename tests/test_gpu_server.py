"""Solve server on the GPU: requests of all three dimensions, submitted
while others run, come back bit-identical to the same requests solved alone
and, in FP64, to the oracle's driver loop run to the same end time; mixed
results stay within 1e-6 relative L1 of FP64; requests at the serving
stream's extreme sizes (1D nx=8192, 2D 1024^2, 3D 192^3) are served correctly;
admission keeps the admitted bytes within the budget; bad requests fail
alone."""
import hashlib
import os
import threading
import time

import numpy as np
import pytest

import mlau_amd as M
from mlau_amd import ensemble as E
from mlau_amd import problems as P
from mlau_amd.server import Server
from oracle import oracle as O

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def small_draws(count, seed):
    """Serving-stream draws shrunk to test sizes (same kinds and states)."""
    out = []
    for d in P.serve_request_draws(count, seed):
        d = dict(d)
        if d["kind"] == 1:
            d["nx"] = 64 + d["nx"] % 97
        elif d["kind"] == 2:
            d["n"] = 16 + (d["n"] % 48) // 16 * 16
        else:
            d["n"] = 12 + d["n"] % 8
            d["mach"] = max(d["mach"], 0.02)
        out.append(d)
    return out


def request(d, precision="fp64", tend=None):
    f, p, s = P.serve_request(d)
    sch = M.scheme(s["odr"], s["rk"], s["flux"], s["ctw"], bc1d=s.get("bc1d", "open"), precision=precision)
    te = tend if tend is not None else {1: 0.02, 2: 0.01, 3: 1e-3}[d["kind"]]
    m = E.member(f, sch, p["dx"], p.get("dy"), p.get("dz"), tend=te, cfl=0.4, vfloor=p.get("vfloor", 0.0),
                 gamma=p["gamma"], bx=p.get("bx", 0.0))
    return f, m


def solve_alone(f, m):
    srv = Server(budget_bytes=1 << 30)
    t = srv.submit(f, m)
    assert srv.wait(t) == E.DONE
    out, tm = srv.result(t)
    srv.close()
    return out, tm


def assert_same(a, b, what):
    bad = a != b
    if bad.any():
        idx = np.argwhere(bad)[:3]
        raise AssertionError("%s: %d mismatches, first %s" % (what, bad.sum(), idx.tolist()))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_stream_alone_equals_mix(gpu, precision):
    draws = small_draws(12, seed=21)
    reqs = [request(d, precision) for d in draws]
    alone = [solve_alone(f, m) for f, m in reqs]
    srv = Server(budget_bytes=1 << 30, steps_per_round=8)
    tickets = []
    for k, (f, m) in enumerate(reqs):  # arrivals spread over the run
        tickets.append(srv.submit(f, m))
        if k % 3 == 2:
            time.sleep(0.01)
    for k, t in enumerate(tickets):
        assert srv.wait(t) == E.DONE
        out, tm = srv.result(t)
        assert tm["nstep"] == alone[k][1]["nstep"]
        assert tm["submit_s"] <= tm["admit_s"] <= tm["finish_s"]
        assert_same(out, alone[k][0], "request %d (%s)" % (k, draws[k]))
    srv.close()


def test_budget_admission_and_concurrent_clients(gpu):
    draws = small_draws(10, seed=33)
    reqs = [request(d) for d in draws]
    alone = [solve_alone(f, m)[0] for f, m in reqs]
    biggest = max(E.Ensemble.member_bytes(m) for _, m in reqs)
    budget = 2 * biggest
    srv = Server(budget_bytes=budget, steps_per_round=8)
    results, peak = {}, [0]
    stop = threading.Event()

    def watch():
        while not stop.is_set():
            peak[0] = max(peak[0], srv.load()["bytes"])
            time.sleep(0.001)

    def client(ks):
        for k in ks:
            t = srv.submit(*reqs[k])
            assert srv.wait(t) == E.DONE
            results[k] = srv.result(t)[0]

    w = threading.Thread(target=watch)
    w.start()
    cl = [threading.Thread(target=client, args=(list(range(c, len(reqs), 3)),)) for c in range(3)]
    for c in cl:
        c.start()
    for c in cl:
        c.join()
    stop.set()
    w.join()
    srv.close()
    assert peak[0] <= budget
    for k in range(len(reqs)):
        assert_same(results[k], alone[k], "request %d" % k)


def test_failures_stay_with_their_request(gpu):
    draws = small_draws(4, seed=5)
    reqs = [request(d) for d in draws]
    alone = [solve_alone(f, m)[0] for f, m in reqs]
    srv = Server(budget_bytes=1 << 30)
    good = [srv.submit(f, m) for f, m in reqs]
    f, m = reqs[0]
    bad = f.copy()
    bad[(0,) + tuple(n // 2 for n in bad.shape[1:])] = np.nan  # NaN density mid-domain: fails at its first CFL check
    t_nan = srv.submit(bad, m)
    tiny = E.member(np.zeros((8, 10, 10)), M.scheme(2, 3, "mlau", 1), 0.1, 0.1, tend=0.01)
    t_bad = srv.submit(np.zeros((8, 10, 10)), tiny)  # grid too small: rejected at admission
    for t in (t_nan, t_bad):
        assert srv.wait(t) == E.FAILED
        with pytest.raises(M.MLAUError):
            srv.result(t)
    for k, t in enumerate(good):
        assert srv.wait(t) == E.DONE
        assert_same(srv.result(t)[0], alone[k], "request %d next to failing ones" % k)
    srv.close()


def test_request_larger_than_budget_is_refused(gpu):
    f, m = request(small_draws(1, seed=2)[0])
    srv = Server(budget_bytes=E.Ensemble.member_bytes(m) // 2)
    t = srv.submit(f, m)
    assert srv.wait(t) == E.FAILED
    with pytest.raises(M.MLAUError, match="budget"):
        srv.result(t)
    srv.close()


def oracle_solve(d, f, m):
    """The oracle's driver loop (oracle.run_to_1d/2d/3d: the reference
    cflcheck -> step -> tim += dt with the end-time clip) for one request."""
    _, p, s = P.serve_request(d)
    g = f.copy()
    sch = O.scheme(s["odr"], s["rk"], s["flux"], s["ctw"])
    if d["kind"] == 1:
        r = O.run_to_1d(g, p["bx"], p["dx"], 4, p["gamma"], sch, m.tend, m.cfl, m.vfloor)
    elif d["kind"] == 2:
        r = O.run_to_2d(g, p["dx"], p["dy"], 4, p["gamma"], sch, m.tend, m.cfl, m.vfloor)
    else:
        r = O.run_to_3d(g, p["dx"], p["dy"], p["dz"], 4, p["gamma"], sch, m.tend, m.cfl, m.vfloor)
    return g, r


def rel_l1(a, b):
    return np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300)


def errors(out, ref, d, gamma):
    """Relative L1 of rho, p and B (interior) of a mixed result against FP64."""
    nd = ref.ndim - 1
    a, b = out[(slice(None),) + (slice(4, -4),) * nd], ref[(slice(None),) + (slice(4, -4),) * nd]
    if nd == 1:  # ro,mx,my,mz,en,by,bz with constant bx
        bx = 0.75
        pr = lambda f: (gamma - 1) * (f[4] - 0.5 * (f[1] ** 2 + f[2] ** 2 + f[3] ** 2) / f[0]
                                      - 0.5 * (bx * bx + f[5] ** 2 + f[6] ** 2))
        bb = lambda f: f[5:7]
    else:
        pr = lambda f: (gamma - 1) * (f[4] - 0.5 * (f[1] ** 2 + f[2] ** 2 + f[3] ** 2) / f[0]
                                      - 0.5 * (f[5] ** 2 + f[6] ** 2 + f[7] ** 2))
        bb = lambda f: f[5:8]
    return {"ro": rel_l1(a[0], b[0]), "p": rel_l1(pr(a), pr(b)), "B": rel_l1(bb(a), bb(b))}


def test_served_results_equal_the_oracle(gpu):
    """Serving-stream requests (shrunk sizes) submitted while others run: every
    FP64 result equals the oracle's driver loop to the same end time value for
    value, with the same step count; every mixed-precision result is within
    1e-6 relative L1 (rho, p, B) of it."""
    draws = small_draws(9, seed=41)
    srv = Server(budget_bytes=1 << 30, steps_per_round=8)
    jobs = []
    for k, d in enumerate(draws):
        for prec in ("fp64", "mixed"):
            f, m = request(d, prec)
            jobs.append((k, prec, f, m, srv.submit(f, m)))
        if k % 3 == 2:
            time.sleep(0.01)
    ref = {}
    for k, d in enumerate(draws):
        f, m = request(d, "fp64")
        ref[k] = oracle_solve(d, f, m)
        assert ref[k][1][2] == 1 and np.all(np.isfinite(ref[k][0])), (d, ref[k][1])
    for k, prec, f, m, t in jobs:
        assert srv.wait(t) == E.DONE
        out, tm = srv.result(t)
        g, (tim, n, st) = ref[k]
        if prec == "fp64":
            assert tm["nstep"] == n, (draws[k], tm["nstep"], n)
            assert_same(out, g, "served FP64 request %d (%s) vs oracle" % (k, draws[k]))
        else:
            e = errors(out, g, draws[k], P.serve_request(draws[k])[1]["gamma"])
            assert max(e.values()) <= 1e-6, (draws[k], e)
    srv.close()


def test_serving_stream_extreme_sizes(gpu):
    """Requests at the serving stream's size limits, served together.
    1D nx=8192 to t=0.1 (FP64 against the reference solver's result, kept as
    the fixture serve1d_8192.npz) and the smallest 2D (64^2) and 3D (32^3)
    requests (against the oracle's driver loop) are checked value for value in
    FP64, their mixed results within 1e-6 of it.  The largest 2D (1024^2 blast to t=0.02) and
    3D (192^3 shear layer to t=2e-4) requests, too slow for the oracle here,
    are checked by size-independent properties: FP64 conserves mass and
    energy to round-off and keeps the ODR-matched div B at round-off, and the
    mixed result (mixed_recon for the strong 2D blast, as the serving bench
    runs it) is within 1e-6 relative L1 of the FP64 one."""
    base = {1: dict(kind=1, nx=8192, rho_l=1.0, p_l=1.0, rho_r=0.125, p_r=0.1, flux="mlau"),
            2: dict(kind=2, n=64, pratio=300.0, bmag=0.7, bang=1.1, flux="hlld"),
            3: dict(kind=3, n=32, mach=0.003, flux="mlau")}
    big = {2: dict(kind=2, n=1024, pratio=500.0, bmag=0.6, bang=0.4, flux="mlau"),
           3: dict(kind=3, n=192, mach=0.01, flux="mlau")}
    tend = {1: 0.1, 2: 0.02, 3: 2e-4}
    srv = Server(budget_bytes=64 << 30, steps_per_round=8)
    sub = {}
    for key, d in list(base.items()) + [(10 + k, v) for k, v in big.items()]:
        kind = d["kind"]
        precs = ("fp64", "mixed_recon" if kind == 2 else "mixed")
        for prec in precs:
            f, m = request(d, prec, tend=tend[kind])
            sub[(key, prec)] = (d, f, m, srv.submit(f, m))
    res = {}
    for k, (d, f, m, t) in sub.items():
        assert srv.wait(t, timeout=900.0) == E.DONE, (k, d)
        res[k] = srv.result(t)[0]
    srv.close()
    for key, d in base.items():
        f, m = sub[(key, "fp64")][1:3]
        out = res[(key, "fp64")]
        if d["kind"] == 1:
            # 7691 steps of 8192 cells take the CPU a minute: the reference solver's
            # result to t=0.1 is kept as a fixture (SHA-256 of x + 0.0 and L1 norms)
            gd = np.load(os.path.join(G, "serve1d_8192.npz"), allow_pickle=False)
            assert hashlib.sha256(np.ascontiguousarray(out + 0.0).tobytes()).hexdigest() == str(gd["sha256"])
            assert np.array_equal(np.abs(out[:, 4:-4]).sum(axis=1), gd["l1"])
            g = out
        else:
            g, (tim, n, st) = oracle_solve(d, f, m)
            assert st == 1
            assert_same(out, g, "extreme-size request %s vs oracle" % d)
        mp = "mixed_recon" if d["kind"] == 2 else "mixed"
        e = errors(res[(key, mp)], g, d, P.serve_request(d)[1]["gamma"])
        assert max(e.values()) <= 1e-6, (d, e)
    for key, d in big.items():
        f, m = sub[(10 + key, "fp64")][1:3]
        out = res[(10 + key, "fp64")]
        nd = d["kind"]
        inner = lambda a: a[(slice(None),) + (slice(4, -4),) * nd]
        for c in (0, 4):  # mass and energy on the periodic grid
            t0, t1 = inner(f)[c].sum(), inner(out)[c].sum()
            assert abs(t1 - t0) <= 1e-12 * np.abs(inner(f)[c]).sum(), (d, c, t0, t1)
        dx = P.serve_request(d)[1]["dx"]
        odr = 2 if nd == 2 else 4
        dd = (lambda a, ax: np.roll(a, -1, ax) - a) if odr == 2 else \
            (lambda a, ax: (27.0 * (np.roll(a, -1, ax) - a) - (np.roll(a, -2, ax) - np.roll(a, 1, ax))) / 24.0)
        axes = list(range(nd - 1, -1, -1))
        div = lambda a: np.abs(sum(dd(a[5 + k], axes[k]) / dx for k in range(nd))[(slice(4, -4),) * nd]).max()
        bscale = np.abs(f[5:5 + nd]).max() / dx
        assert div(out) <= max(10 * div(f), 1e-12 * bscale), (d, div(out), div(f))
        mp = "mixed_recon" if nd == 2 else "mixed"
        e = errors(res[(10 + key, mp)], out, d, P.serve_request(d)[1]["gamma"])
        assert max(e.values()) <= 1e-6, (d, e)
