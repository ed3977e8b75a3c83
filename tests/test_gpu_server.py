"""Solve server on the GPU: requests of all three dimensions, submitted
while others run, come back bit-identical to the same requests solved alone;
admission keeps the admitted bytes within the budget; bad requests fail
alone."""
import threading
import time

import numpy as np
import pytest

import mlau_amd as M
from mlau_amd import ensemble as E
from mlau_amd import problems as P
from mlau_amd.server import Server

pytestmark = pytest.mark.gpu


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
