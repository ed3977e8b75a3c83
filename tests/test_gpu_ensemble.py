"""Batched ensemble engine on the GPU.

Every member must give exactly the oracle's bits when run to its end time
(per-member CFL step with the end-time clip), and exactly the same bits alone,
in any mix, under any launch tuning, with or without the HIP graph, and after
members around it were removed or added (continuous batching)."""
import numpy as np
import pytest

import mlau_amd as M
from mlau_amd import ensemble as E
from mlau_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TEND = 0.02


def blast_member(n, pratio, bmag, bang, flux, odr=2, rk=3, ctw=1, precision="fp64", tend=TEND, max_steps=0):
    f, p = P.blast_2d(n, pratio, bmag, bang)
    sch = M.scheme(odr, rk, flux, ctw, precision=precision)
    m = E.member(f, sch, p["dx"], p["dy"], tend=tend, cfl=0.4, vfloor=p["vfloor"], gamma=p["gamma"],
                 max_steps=max_steps)
    return f, p, m


def oracle_run(f, p, odr, rk, flux, ctw, tend=TEND, max_steps=0):
    g = f.copy()
    tim, n, st = O.run_to_2d(g, p["dx"], p["dy"], 4, p["gamma"], O.scheme(odr, rk, flux, ctw), tend, 0.4,
                             p["vfloor"], max_steps=max_steps)
    assert np.all(np.isfinite(g)), "oracle case must stay finite"
    return g, tim, n, st


def assert_same(a, b, what):
    bad = a != b
    if bad.any():
        idx = np.argwhere(bad)[:3]
        raise AssertionError("%s: %d mismatches, first %s a=%r b=%r" % (
            what, bad.sum(), idx.tolist(), a[tuple(idx[0])], b[tuple(idx[0])]))


MIX = [  # (n, pratio, bmag, bang, flux)
    (32, 100.0, 0.5, 0.3, "mlau"), (48, 20.0, 1.0, 2.0, "hlld"), (16, 800.0, 0.1, 4.0, "roe"),
    (64, 300.0, 0.7, 5.5, "lhlld"), (40, 15.0, 0.3, 1.0, "mlau"), (24, 50.0, 0.9, 3.3, "hlld"),
]


def test_members_match_oracle(gpu):
    ens = E.Ensemble(budget_bytes=1 << 30)
    ids, ref = [], []
    for (n, pr, bm, ba, fl) in MIX:
        f, p, m = blast_member(n, pr, bm, ba, fl)
        ids.append(ens.add(f, m))
        ref.append(oracle_run(f, p, 2, 3, fl, 1))
    ens.run_all(chunk=16)
    for k, i in enumerate(ids):
        g, tim, n, st = ref[k]
        s, t, _, ns = ens.status(i)
        assert (s, ns) == (E.DONE, n) and st == 1, (MIX[k], s, ns, n)
        assert t == tim == TEND
        assert_same(ens.download(i), g, "member %s" % (MIX[k],))
    ens.close()


# blasts that every scheme below carries to t=0.02 without a negative pressure
SCHEME_CASES = [(32, 100.0, 0.5, 0.3, "mlau"), (48, 20.0, 1.0, 2.0, "hlld"), (24, 50.0, 0.9, 3.3, "roe"),
                (40, 15.0, 0.3, 1.0, "lhlld")]


@pytest.mark.parametrize("odr,rk,ctw", [(4, 3, 1), (1, 2, 0), (3, 1, 1)])
def test_other_schemes_match_oracle(gpu, odr, rk, ctw):
    ens = E.Ensemble(budget_bytes=1 << 30)
    cases = SCHEME_CASES
    ids, ref = [], []
    for (n, pr, bm, ba, fl) in cases:
        f, p, m = blast_member(n, pr, bm, ba, fl, odr=odr, rk=rk, ctw=ctw)
        ids.append(ens.add(f, m))
        ref.append(oracle_run(f, p, odr, rk, fl, ctw))
    ens.run_all(chunk=16)
    for k, i in enumerate(ids):
        assert ens.status(i)[3] == ref[k][2]
        assert_same(ens.download(i), ref[k][0], "odr%d rk%d ctw%d %s" % (odr, rk, ctw, cases[k]))
    ens.close()


def run_mix(cases, tune=None, graph=True, chunk=16, balance=False):
    ens = E.Ensemble(budget_bytes=1 << 30, graph=graph, balance=balance)
    ids = [ens.add(*blast_member(*c)[::2]) for c in cases]
    if tune is not None:
        ens.set_tune(tune)
    ens.run_all(chunk=chunk)
    out = [ens.download(i) for i in ids]
    steps = [ens.status(i)[3] for i in ids]
    ens.close()
    return out, steps


def test_alone_equals_mix_and_tunings(gpu):
    draws = P.blast_ensemble_draws(count=24, seed=11, nmin=16, nmax=96)
    cases = [(d["n"], d["pratio"], d["bmag"], d["bang"], d["flux"]) for d in draws]
    base, steps = run_mix(cases)
    for k in (0, 7, 19):
        alone, st = run_mix([cases[k]])
        assert st[0] == steps[k]
        assert_same(alone[0], base[k], "alone vs mix, member %d" % k)
    for tune in (0x11, 0x42, 0x84 | 1, 0x20 | 4 | 2, 0x100 | 0x21, 0x400 | 0x11, 0x1600 | 0x20):
        other, st = run_mix(cases, tune=tune)
        assert st == steps
        for k in range(len(cases)):
            assert_same(other[k], base[k], "tune %#x member %d" % (tune, k))
    # balanced rounds (cheap groups take more steps per launch), with the server's shape
    other, st = run_mix(cases, tune=0x1400, balance=True)
    assert st == steps
    for k in range(len(cases)):
        assert_same(other[k], base[k], "balanced, tune 0x1400, member %d" % k)


def test_graph_replays_equal_eager(gpu):
    """The captured step loop is replayed many times, with host synchronisation
    and downloads between calls; every call's state must equal the eager run."""
    cases = MIX
    eg = E.Ensemble(budget_bytes=1 << 30, graph=True)
    ee = E.Ensemble(budget_bytes=1 << 30, graph=False)
    ig = [eg.add(*blast_member(*c)[::2]) for c in cases]
    ie = [ee.add(*blast_member(*c)[::2]) for c in cases]
    for call in range(6):
        ng = eg.run(8)
        ne = ee.run(8)
        assert ng == ne
        for a, b in zip(ig, ie):
            assert_same(eg.download(a), ee.download(b), "graph call %d" % call)
            assert eg.status(a) == ee.status(b)
    eg.close()
    ee.close()


def test_continuous_batching_and_budget(gpu):
    """Members added while others run, finished ones removed and their memory
    reused: each result equals its alone run; admission respects the budget."""
    cases = MIX + [(56, 40.0, 0.6, 0.7, "hlld"), (32, 500.0, 0.2, 1.7, "roe")]
    alone = [run_mix([c])[0][0] for c in cases]
    _, _, m0 = blast_member(*cases[3])
    budget = 3 * E.Ensemble.member_bytes(m0)
    ens = E.Ensemble(budget_bytes=budget)
    pending = list(range(len(cases)))
    live, done = {}, {}
    while pending or live:
        while pending:
            f, _, m = blast_member(*cases[pending[0]])
            try:
                mid = ens.add(f, m)
            except M.MLAUError as err:
                assert "budget" in str(err)
                assert live, "a lone member must fit"
                break
            live[mid] = pending.pop(0)
            assert ens.used_bytes() <= budget
        ens.run(5)
        for mid in ens.finished():
            assert ens.status(mid)[0] == E.DONE
            done[live.pop(mid)] = ens.download(mid)
            ens.remove(mid)
    assert ens.used_bytes() == 0
    for k in range(len(cases)):
        assert_same(done[k], alone[k], "batched member %d" % k)
    ens.close()


def test_failed_member_is_isolated(gpu):
    cases = MIX[:3]
    ref = run_mix(cases)[0]
    ens = E.Ensemble(budget_bytes=1 << 30)
    ids = []
    broken = None
    for k, c in enumerate(cases):
        f, _, m = blast_member(*c)
        if k == 1:
            f = f.copy()
            f[0, 20, 20] = np.nan  # a broken member: a non-finite density fails it at its first CFL check
            broken = f
        ids.append(ens.add(f, m))
    ens.run_all(chunk=16)
    assert ens.status(ids[1])[0] == E.FAILED and ens.status(ids[1])[3] == 0
    # the failing step advanced nothing: the state at failure is the state it failed on
    assert np.array_equal(ens.download(ids[1]), broken, equal_nan=True)
    for k in (0, 2):
        assert ens.status(ids[k])[0] == E.DONE
        assert_same(ens.download(ids[k]), ref[k], "healthy member %d" % k)
    with pytest.raises(M.MLAUError):
        ens.remove(12345)
    ens.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_nan_speed_cell_is_skipped_like_the_reference(gpu, dim):
    """A finite state with one cell where gamma*p + B^2 < 0 has a NaN fast
    speed there.  The reference's cflcheck skips that cell ("if (vtmp > vmax)",
    2D/SERIAL/OTvortex/cflcheck.hpp:21-23, 3D/MPI/OTvortex/cflcheck.hpp:23-25):
    the member must take the same first dt as the single-grid device CFL check
    and the oracle, and step on instead of failing.  Away from that cell (more
    than one RK step's stencil reach) its state after the step equals the
    oracle's value for value; at the cell itself the solvers meet square roots
    of negative numbers, where NaN propagation through structurally zero
    eigenvector entries (which the GPU skips) is not the reference's."""
    if dim == 2:
        f, p = P.blast_2d(32, 100.0, 0.5, 0.3)
        cell = (20, 20)  # mid-domain: one step's reach (9 cells) touches no cell with a periodic image
        run = lambda g: O.run_to_2d(g, p["dx"], p["dy"], 4, p["gamma"], O.scheme(2, 3, "mlau", 1), TEND, 0.4,
                                    p["vfloor"], max_steps=1)
        vmax_o = O.cfl_vmax_2d
        sch = M.scheme(2, 3, "mlau", 1)
        mk = lambda g: E.member(g, sch, p["dx"], p["dy"], tend=TEND, cfl=0.4, vfloor=p["vfloor"], gamma=p["gamma"],
                                max_steps=1)
        solver = lambda g: M.Solver(g, sch, p["dx"], p["dy"], gamma=p["gamma"])
    else:
        f, p = P.ot_vortex_3d(32)
        cell = (20, 20, 20)
        run = lambda g: O.run_to_3d(g, p["dx"], p["dy"], p["dz"], 4, p["gamma"], O.scheme(2, 3, "mlau", 1), 0.01,
                                    0.4, p["vfloor"], max_steps=1)
        vmax_o = O.cfl_vmax_3d
        sch = M.scheme(2, 3, "mlau", 1)
        mk = lambda g: E.member(g, sch, p["dx"], p["dy"], p["dz"], tend=0.01, cfl=0.4, vfloor=p["vfloor"],
                                gamma=p["gamma"], max_steps=1)
        solver = lambda g: M.Solver(g, sch, p["dx"], p["dy"], p["dz"], gamma=p["gamma"])
    f = f.copy()
    f[(4,) + cell] = -40.0  # energy far below the kinetic + magnetic part: gamma*p + B^2 < 0
    assert np.all(np.isfinite(f))
    vm = vmax_o(f, 4, p["gamma"], p["vfloor"])
    assert np.isfinite(vm)
    s = solver(f)
    dt_dev, vm_dev = s.cfl(0.4, p["vfloor"])
    s.close()
    assert vm_dev == vm
    g = f.copy()
    tim, n, st = run(g)
    assert (n, st) == (1, 3)
    ens = E.Ensemble(budget_bytes=1 << 30)
    i = ens.add(f, mk(f))
    ens.run_all(chunk=4)
    es, et, _, en = ens.status(i)
    assert (es, en) == (E.CAPPED, 1) and et == tim == dt_dev  # stepped with the reference's dt
    out = ens.download(i)
    ens.close()
    far = np.ones(out.shape[1:], dtype=bool)
    far[tuple(slice(max(c - 12, 0), c + 13) for c in cell)] = False
    assert np.array_equal(out[:, far], g[:, far]), dim


def test_max_steps_caps(gpu):
    f, p, m = blast_member(*MIX[0], max_steps=3)
    g, tim, n, st = oracle_run(f, p, 2, 3, MIX[0][4], 1, max_steps=3)
    ens = E.Ensemble(budget_bytes=1 << 30)
    i = ens.add(f, m)
    ens.run_all(chunk=16)
    s, t, _, ns = ens.status(i)
    assert (s, ns, st) == (E.CAPPED, 3, 3) and t == tim
    assert_same(ens.download(i), g, "capped member")
    with pytest.raises(M.MLAUError):
        ens.add(f[:, :10, :10].copy(), E.member(f[:, :10, :10].copy(), m.scheme, 0.1, 0.1, tend=0.02))
    ens.close()


def test_autotune_keeps_state(gpu):
    draws = P.blast_ensemble_draws(count=12, seed=5, nmin=32, nmax=128)
    cases = [(d["n"], d["pratio"], d["bmag"], d["bang"], d["flux"]) for d in draws]
    base, steps = run_mix(cases)
    ens = E.Ensemble(budget_bytes=1 << 30)
    ids = [ens.add(*blast_member(*c)[::2]) for c in cases]
    ens.run(4)
    tune = ens.autotune(steps=4)
    assert 0 <= tune < 512
    ens.run_all(chunk=16)
    for k, i in enumerate(ids):
        assert ens.status(i)[3] == steps[k]
        assert_same(ens.download(i), base[k], "after autotune, member %d" % k)
    ens.close()


@pytest.mark.parametrize("precision", ["mixed", "mixed_recon"])
def test_mixed_precision_members(gpu, precision):
    """Mixed-precision members (both mixed forms, side by side with nothing
    else in their group) against the FP64 oracle: relative L1 of rho, p and B
    at or below 1e-6."""
    ens = E.Ensemble(budget_bytes=1 << 30)
    cases = MIX[:4]
    ids, ref = [], []
    for (n, pr, bm, ba, fl) in cases:
        f, p, m = blast_member(n, pr, bm, ba, fl, precision=precision)
        ids.append(ens.add(f, m))
        ref.append((oracle_run(f, p, 2, 3, fl, 1), p))
    ens.run_all(chunk=16)
    for k, i in enumerate(ids):
        (g, tim, _, _), p = ref[k]
        s, t, _, _ = ens.status(i)
        assert s == E.DONE and t == tim
        a = ens.download(i)[:, 4:-4, 4:-4]
        b = g[:, 4:-4, 4:-4]
        errs = {"rho": rel_l1(a[0], b[0]), "p": rel_l1(pressure(a, p["gamma"]), pressure(b, p["gamma"])),
                "B": rel_l1(a[5:8], b[5:8])}
        assert max(errs.values()) <= 1e-6, (cases[k], errs)
    ens.close()


def rel_l1(a, b):
    return np.abs(a - b).sum() / np.abs(b).sum()


def pressure(f, gamma):
    b2 = f[5] ** 2 + f[6] ** 2 + f[7] ** 2  # face values: adequate for a relative comparison
    return (gamma - 1) * (f[4] - 0.5 * (f[1] ** 2 + f[2] ** 2 + f[3] ** 2) / f[0] - 0.5 * b2)


# ------------------------------------------------------------ 1D and 3D members
def oned_member(nx, bc="open", flux="mlau", odr=2, rk=3, tend=0.05, rho_l=1.0):
    f, p = P.riemann_1d(nx, (rho_l, 1.0, 0, 0, 0, 1.0, 0), (0.125, 0.1, 0, 0, 0, -1.0, 0), 0.75, 2.0)
    sch = M.scheme(odr, rk, flux, 1, bc1d=bc)
    m = E.member(f, sch, p["dx"], tend=tend, cfl=0.4, vfloor=1.0, gamma=p["gamma"], bx=p["bx"])
    return f, p, m


def threed_member(kind, n, flux, odr, tend, mach=0.01):
    if kind == "shear":
        f, p = P.shear_layer_3d(n, mach=mach)
    else:
        f, p = P.ot_vortex_3d(n)
    sch = M.scheme(odr, 3, flux, 1)
    m = E.member(f, sch, p["dx"], p["dy"], p["dz"], tend=tend, cfl=0.4, vfloor=p["vfloor"], gamma=p["gamma"])
    return f, p, m


ONED = [(100, "open", "mlau", 2, 3), (256, "periodic", "hlld", 2, 3), (64, "open", "lhlld", 1, 2),
        (130, "periodic", "roe", 4, 3)]
THREED = [("shear", 16, "mlau", 4, 1e-3), ("ot", 16, "lhlld", 2, 0.03), ("ot", 12, "hlld", 3, 0.02)]


def test_1d_members_match_oracle(gpu):
    ens = E.Ensemble(budget_bytes=1 << 30)
    ids, ref = [], []
    for (nx, bc, fl, odr, rk) in ONED:
        f, p, m = oned_member(nx, bc, fl, odr, rk)
        ids.append(ens.add(f, m))
        g = f.copy()
        tim, n, st = O.run_to_1d(g, p["bx"], p["dx"], 4, p["gamma"], O.scheme(odr, rk, fl, 1), 0.05, 0.4, 1.0,
                                 periodic=(bc == "periodic"))
        assert st == 1 and np.all(np.isfinite(g))
        ref.append((g, tim, n))
    ens.run_all(chunk=64)
    for k, i in enumerate(ids):
        s, t, _, ns = ens.status(i)
        assert (s, ns, t) == (E.DONE, ref[k][2], ref[k][1]), ONED[k]
        assert_same(ens.download(i), ref[k][0], "1D member %s" % (ONED[k],))
    ens.close()


def test_3d_members_match_oracle(gpu):
    ens = E.Ensemble(budget_bytes=1 << 30)
    ids, ref = [], []
    for (kind, n, fl, odr, tend) in THREED:
        f, p, m = threed_member(kind, n, fl, odr, tend)
        ids.append(ens.add(f, m))
        g = f.copy()
        tim, ns, st = O.run_to_3d(g, p["dx"], p["dy"], p["dz"], 4, p["gamma"], O.scheme(odr, 3, fl, 1), tend, 0.4,
                                  p["vfloor"])
        assert st == 1 and np.all(np.isfinite(g))
        ref.append((g, tim, ns))
    ens.run_all(chunk=16)
    for k, i in enumerate(ids):
        s, t, _, ns = ens.status(i)
        assert (s, ns, t) == (E.DONE, ref[k][2], ref[k][1]), THREED[k]
        assert_same(ens.download(i), ref[k][0], "3D member %s" % (THREED[k],))
    ens.close()


@pytest.mark.parametrize("balance", [False, True])
def test_mixed_dimensions_alone_equals_mix(gpu, balance):
    """1D, 2D and 3D members (three launch groups, each on its own stream) in
    one engine give the bits each gives alone, also when cheap groups take
    more steps per call than expensive ones (balance)."""
    members = ([oned_member(*c)[::2] for c in ONED[:2]] + [blast_member(*c)[::2] for c in MIX[:3]] +
               [threed_member(*c)[::2] for c in THREED[:2]])
    alone = []
    for f, m in members:
        ens = E.Ensemble(budget_bytes=1 << 30)
        i = ens.add(f, m)
        ens.run_all(chunk=16)
        alone.append((ens.download(i), ens.status(i)))
        ens.close()
    ens = E.Ensemble(budget_bytes=1 << 30, balance=balance)
    ids = [ens.add(f, m) for f, m in members]
    while ens.run(8) > 0:
        pass
    for k, i in enumerate(ids):
        assert ens.status(i) == alone[k][1]
        assert_same(ens.download(i), alone[k][0], "mixed-dimension member %d" % k)
    ens.close()


@pytest.mark.parametrize("walls", ["neumann", "khi"])
def test_wall_members_match_oracle(gpu, walls):
    """Members with the reference drivers' wall flag sets (per member, next to
    periodic ones) equal the oracle."""
    dnx, dny = {"neumann": ((1,) * 8, (1,) * 8), "khi": ((0,) * 8, (1, 1, -1, 1, 1, 1, -1, 1))}[walls]
    ens = E.Ensemble(budget_bytes=1 << 30)
    f, p = P.blast_2d(40, 100.0, 0.6, 0.4)
    for n in range(8):  # ghosts consistent with the wall rule first
        O.lib().oc_bc2d(O._p(f[n]), f.shape[2], f.shape[1], 4, 4, int(n == 5), dnx[n], int(n == 6), dny[n])
    sch = M.scheme(2, 3, "hlld", 1, dnxs=dnx, dnys=dny)
    i = ens.add(f, E.member(f, sch, p["dx"], p["dy"], tend=0.01, cfl=0.4, vfloor=p["vfloor"], gamma=p["gamma"]))
    g0, _, m0 = blast_member(*MIX[0], tend=0.01)
    j = ens.add(g0, m0)
    ens.run_all(chunk=16)
    ref = f.copy()
    tim, n, st = O.run_to_2d(ref, p["dx"], p["dy"], 4, p["gamma"], O.scheme(2, 3, "hlld", 1), 0.01, 0.4,
                             p["vfloor"], dn=dnx, dny=dny)
    assert ens.status(i)[3] == n
    assert_same(ens.download(i), ref, "wall member " + walls)
    assert_same(ens.download(j), oracle_run(g0, blast_member(*MIX[0])[1], 2, 3, MIX[0][4], 1, tend=0.01)[0],
                "periodic neighbour")
    ens.close()


def test_3d_wall_member_matches_oracle(gpu):
    """A 3D member with sign-flipping wall flags on every axis, beside a
    periodic 3D member of the same group, equals the oracle run to its end
    time."""
    dn = (1, -1, 1, -1, 1, 1, -1, 1)
    g, p = P.ot_vortex_3d(24)
    f = np.ascontiguousarray(g[:, :19, :17, :])  # nz=19, ny=17, nx=32
    O.bc3d_state(f, 4, dn)
    sch = M.scheme(4, 3, "mlau", 1, dnxs=dn)
    ens = E.Ensemble(budget_bytes=1 << 30)
    i = ens.add(f, E.member(f, sch, p["dx"], p["dy"], p["dz"], tend=0.01, cfl=0.4, vfloor=p["vfloor"],
                            gamma=p["gamma"]))
    g0, p0, m0 = threed_member("ot", 16, "mlau", 4, 0.01)
    j = ens.add(g0, m0)
    ens.run_all(chunk=8)
    ref = f.copy()
    tim, n, st = O.run_to_3d(ref, p["dx"], p["dy"], p["dz"], 4, p["gamma"], O.scheme(4, 3, "mlau", 1), 0.01, 0.4,
                             p["vfloor"], dn=dn)
    assert st == 1 and np.all(np.isfinite(ref))
    assert ens.status(i)[:2] == (E.DONE, tim) and ens.status(i)[3] == n
    assert_same(ens.download(i), ref, "3D wall member")
    ref0 = g0.copy()
    O.run_to_3d(ref0, p0["dx"], p0["dy"], p0["dz"], 4, p0["gamma"], O.scheme(4, 3, "mlau", 1), 0.01, 0.4, p0["vfloor"])
    assert_same(ens.download(j), ref0, "periodic 3D neighbour")
    ens.close()


def test_launch_sync_overlaps_add_download_remove(gpu):
    """Steps queued by launch run while members are added (uploaded) and
    finished members are downloaded and removed; every member still gives the
    bits it gives alone, and a running member read during the launch shows
    the state after it."""
    members = ([oned_member(*c)[::2] for c in ONED[:2]] + [blast_member(*c)[::2] for c in MIX[:4]] +
               [threed_member(*c)[::2] for c in THREED[:2]])
    alone = []
    for f, m in members:
        ens = E.Ensemble(budget_bytes=1 << 30)
        i = ens.add(f, m)
        ens.run_all(chunk=16)
        alone.append((ens.download(i), ens.status(i)))
        ens.close()
    ens = E.Ensemble(budget_bytes=1 << 30)
    key, results = {}, {}
    order = [2, 6, 0, 3, 7, 1, 4, 5]  # mixed dimensions join while others run
    for k in order[:2]:
        key[ens.add(*members[k])] = k
    pending = order[2:]
    checked_running = False
    with pytest.raises(M.MLAUError):
        ens.launch(8)
        ens.launch(8)
    ens.sync()
    for _ in range(10000):
        done = set(ens.finished())
        live = [i for i in key if key[i] not in results and i not in done]
        ens.launch(8)
        for i in done:  # not running at the last sync: copied while the steps run
            results[key[i]] = (ens.download(i), ens.status(i))
            ens.remove(i)
        if pending:
            k = pending.pop(0)
            key[ens.add(*members[k])] = k
        if not checked_running and live:
            mid = live[0]
            during = ens.download(mid)  # a running member: waits for the launched steps
            st = ens.status(mid)
            ens.sync()
            assert_same(ens.download(mid), during, "running member read during a launch")
            assert ens.status(mid) == st
            checked_running = True
        while not ens.ready():  # polling, as a server admitting arrivals mid-launch does
            pass
        nrun = ens.sync()
        if nrun == 0 and not pending and not ens.finished():
            break
    assert checked_running and len(results) == len(members)
    for k in range(len(members)):
        assert results[k][1] == alone[k][1], k
        assert_same(results[k][0], alone[k][0], "pipelined member %d" % k)
    assert ens.used_bytes() == 0
    ens.close()


# periodic MUSCL members at rk >= 2 run each stage as one fused kernel; the
# separate stage kernels (fused=False) must give the same bits in FP64, for
# every solver, CT weighting and rk, on tiles cut at the grid edge.  The mixed
# forms let the compiler contract FP32 multiply-adds where it likes, which it
# does differently in the two kernels: there the bound is the mixed forms'
# own, relative L1 per field <= 1e-6.
FUSED_CASES = [(16, 100.0, 0.5, 0.3, "mlau"), (40, 20.0, 1.0, 2.0, "hlld"), (24, 50.0, 0.9, 3.3, "roe"),
               (72, 15.0, 0.3, 1.0, "lhlld"), (33, 300.0, 0.7, 5.5, "mlau")]


@pytest.mark.parametrize("precision,rk,ctw", [("fp64", 3, 1), ("fp64", 2, 0), ("mixed", 3, 1), ("mixed_recon", 3, 0)])
def test_fused_stage_equals_separate_kernels(gpu, precision, rk, ctw):
    out = {}
    for fused in (True, False):
        ens = E.Ensemble(budget_bytes=1 << 30, fused=fused)
        ens.kernel_timing(True)
        ids = []
        for (n, pr, bm, ba, fl) in FUSED_CASES:
            f, p, m = blast_member(n, pr, bm, ba, fl, rk=rk, ctw=ctw, precision=precision)
            ids.append(ens.add(f, m))
        ens.run_all(chunk=16)
        stages = ens.kernel_stats().get("ens_stage", (0.0, 0))[1]
        assert (stages > 0) == fused, (fused, stages)
        out[fused] = [(ens.status(i), ens.download(i)) for i in ids]
        ens.close()
    for k, case in enumerate(FUSED_CASES):
        (sa, a), (sb, b) = out[True][k], out[False][k]
        assert sa[0] == sb[0] == E.DONE and sa[1] == sb[1] == TEND, (case, sa, sb)
        if precision == "fp64":
            assert sa == sb, (case, sa, sb)
            assert_same(a, b, "fused vs separate %s" % (case,))
        else:
            for v in range(a.shape[0]):
                d, r = np.abs(a[v] - b[v]).sum(), np.abs(b[v]).sum()
                assert d <= 1e-6 * r, (case, v, d, r)


def test_fused_size_limit_matches_oracle(gpu):
    """Interiors smaller than two ghost widths (N < 8) stay on the separate
    stage kernels; N >= 8 runs fused (one tile, a tile of 8 or 9 cells, a
    29-cell interior cut into a 28-cell and a 1-cell tile).  Both kinds in one
    ensemble equal the oracle bit for bit, ghost cells included."""
    cases = [(6, 40.0, 0.5, 0.3, "mlau"), (7, 40.0, 0.5, 1.3, "hlld"), (8, 60.0, 0.8, 2.0, "mlau"),
             (9, 60.0, 0.4, 4.0, "hlld"), (29, 200.0, 0.7, 5.0, "mlau")]
    ens = E.Ensemble(budget_bytes=1 << 30)
    ens.kernel_timing(True)
    ids, ref = [], []
    for (n, pr, bm, ba, fl) in cases:
        f, p, m = blast_member(n, pr, bm, ba, fl, tend=0.01)
        ids.append(ens.add(f, m))
        ref.append(oracle_run(f, p, 2, 3, fl, 1, tend=0.01))
    ens.run_all(chunk=16)
    st = ens.kernel_stats()
    assert st.get("ens_stage", (0.0, 0))[1] > 0 and st.get("ens_xsweep", (0.0, 0))[1] > 0, st
    for k, i in enumerate(ids):
        g, tim, n, s = ref[k]
        assert ens.status(i)[0] == E.DONE and ens.status(i)[3] == n, (cases[k], ens.status(i), n)
        assert_same(ens.download(i), g, "member %s" % (cases[k],))
    ens.close()
