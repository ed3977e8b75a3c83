#!/usr/bin/env python3
"""Headline benchmark: zone-updates/s of the plain FP64 2D step at 2048^2.

Workload: periodic Orszag-Tang-type vortex on [0,1]^2, 2048x2048 interior
cells (+4 ghosts per side), LHLLD flux, 4th-order WCNS reconstruction, CT,
SSP-RK3, CFL 0.4 with dt recomputed on the device after every step.
A "step" is one full RK3 time step of the whole grid (3 stages + CFL).
Inputs are resident in HBM before the timed region starts.

Prints ONE JSON line (rank 0).  One GPU only: --gpus N>1 exits with a message.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, vendor figure
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64 vector, vendor figure
BYTES_PER_ZONE_UPDATE = 512  # 8 fields x 8 B x 8 passes per full RK3 step

# Per-launch figures of the 2048^2 FP64 headline kernels, copied from the
# committed rocprofv3 summaries named here (tests/test_bench_profiles.py checks
# that they agree): HBM bytes per launch (FETCH_SIZE / WRITE_SIZE scaled by the
# copy-probe calibration) and FP64 VALU wave-instructions per launch.
PROFILE_2048 = {
    "traffic_file": "profiles/vortex2d_2048_fp64_hbm_traffic.md",
    # HBM bytes per launch (update2d: mean of the first-stage and other-stage launches, 1:2)
    "traffic": {"xsweep2d_inl": 601.4e6, "ysweep2d_inl": 993.5e6, "update2d": (996.8e6 + 2 * 1264.7e6) / 3,
                "ez2d": 203.8e6, "cfl2d": 294.6e6},
    "fp64_file": "profiles/vortex2d_2048_fp64_valu.md",
    "fp64": {"xsweep2d_inl": 136.31e6, "ysweep2d_inl": 142.89e6, "update2d": (13.37e6 + 2 * 14.42e6) / 3,
             "ez2d": 6.11e6, "cfl2d": 4.08e6},
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=2048, help="interior cells per side")
    ap.add_argument("--cpu-seconds", type=float, default=24.0,
                    help="bounded CPU-baseline sample (split over 3 repeats)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "mixed", "mixed_recon"], default=None,
                    help="fp64: the reference's arithmetic; mixed: FP32 reconstruction and Riemann solver; "
                         "mixed_recon: FP32 reconstruction, FP64 Riemann solver.  Default: fp64, except mixed for "
                         "serve (the serving stream is specified in mixed precision; its 2D blasts run mixed_recon)")
    ap.add_argument("--workload", choices=["vortex2d", "ensemble", "serve", "shear3d", "riemann1d"],
                    default="vortex2d",
                    help="vortex2d: the headline single-grid step; ensemble: batched 2D blasts run to t=0.02; "
                         "serve: Poisson stream of 1D/2D/3D solve requests through the solve server")
    ap.add_argument("--requests", type=int, default=96, help="serve: requests per arrival rate")
    ap.add_argument("--rates", default="10,100,1000", help="serve: arrival rates (requests/s), comma separated")
    ap.add_argument("--serve-batch", action="store_true",
                    help="serve: the whole request mix as one batch through the server's ensemble engine (no server "
                         "thread, no arrivals), with per-kernel times")
    ap.add_argument("--members", type=int, default=512, help="ensemble members")
    ap.add_argument("--sequential-every", type=int, default=8,
                    help="ensemble: time every k-th member alone for the one-at-a-time comparison (0: skip)")
    ap.add_argument("--autotune", action="store_true", help="ensemble: autotune the launch shape first")
    ap.add_argument("--tune", type=int, default=0, help="ensemble: launch shape (mlau_ens_set_tune bits)")
    ap.add_argument("--no-graph", action="store_true", help="ensemble: eager launches instead of HIP graph replay")
    a = ap.parse_args()
    if a.precision is None:
        a.precision = "mixed" if a.workload == "serve" else "fp64"
    return a


def host_cpu():
    """CPU model (lscpu / /proc/cpuinfo) and the host's logical CPU count."""
    model = None
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as fh:
                for line in fh:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return model, os.cpu_count() or 1


def _cpu_threads():
    """Threads for the CPU baseline: every host CPU this process may use, up
    to the CPU share the host gives one GPU (its OMP_NUM_THREADS when set, as
    on the GPU boxes, where it is 16)."""
    from oracle import oracle as O
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    nth = min(avail, share) if share > 0 else avail
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    O.set_threads(nth)
    return nth


def cpu_baseline(seconds, repeats=3):
    """Reference CPU path (oracle/_ref build of the reference sources when it is
    present, else the bit-identical restatement) on a bounded sample of the same
    workload at 512^2 (the size named for the CPU): each repeat runs steps from
    t=0 for seconds/repeats of wall clock around the step loop only (with the
    CFL check after every step, as on the GPU); the reported rate is the median
    of the repeats."""
    from oracle import oracle as O
    from mlau_amd import problems as P
    n = 512
    nth = _cpu_threads()
    kind = "port"
    lib = None
    if os.path.exists(os.path.join(O.REF_DIR, "libref2d_o4r3_lhlld.so")):
        kind, lib = "reference", "ref2d_o4r3_lhlld"
    rates, counts = [], []
    for _ in range(repeats):
        f, p = P.vortex_2d(n, 4)
        if lib:
            step = lambda dt: O.ref_step_2d(lib, f, dt, p["dx"], p["dy"], 4, p["gamma"])
        else:
            step = lambda dt: O.step_2d(f, dt, p["dx"], p["dy"], 4, p["gamma"], O.scheme(4, 3, "lhlld", 1))
        dt = 0.4 * p["dx"] / O.cfl_vmax_2d(f, 4, p["gamma"], p["vfloor"])
        steps = 0
        t0 = time.perf_counter()
        while True:
            step(dt)
            vm = O.cfl_vmax_2d(f, 4, p["gamma"], p["vfloor"])
            if np.isfinite(vm):
                dt = 0.4 * p["dx"] / vm
            steps += 1
            el = time.perf_counter() - t0
            if el >= seconds / repeats or steps >= 2000:
                break
        rates.append(steps * n * n / el)
        counts.append(steps)
    model, ncpu = host_cpu()
    return {"value": float(np.median(rates)), "unit": "zone-updates/s", "cores": nth, "kind": kind,
            "sample": "2D vortex 512^2 LHLLD-WCNS4-CT-RK3 from t=0, %d repeats of %.0f s (%s steps), median"
                      % (repeats, seconds / repeats, "/".join(str(c) for c in counts)),
            "repeats": [float(r) for r in rates], "cpu_model": model, "host_cpus": ncpu,
            "threads_note": "OpenMP threads = the CPUs this process may use, capped at the host's per-GPU share "
                            "(OMP_NUM_THREADS); OMP_PROC_BIND=close, OMP_PLACES=cores"}


def mixed_errors(mixed, fp64, gamma, bx=None):
    """Relative L1 of rho, p and |B| components of a mixed-precision result
    against the FP64 result of the same problem, over interior cells
    (face B values for 2D/3D, as the tests compare them)."""
    nd = mixed.ndim - 1
    sl = (slice(None),) + (slice(4, -4),) * nd
    m, f = mixed[sl], fp64[sl]

    def pr(u):
        if u.shape[0] == 7:  # 1D: ro,mx,my,mz,en,by,bz with constant bx
            b2 = bx * bx + u[5] ** 2 + u[6] ** 2
        else:
            b2 = u[5] ** 2 + u[6] ** 2 + u[7] ** 2
        return (gamma - 1.0) * (u[4] - 0.5 * (u[1] ** 2 + u[2] ** 2 + u[3] ** 2) / u[0] - 0.5 * b2)

    rel = lambda x, y: float(np.abs(x - y).sum() / np.abs(y).sum())
    return {"rho": rel(m[0], f[0]), "p": rel(pr(m), pr(f)), "B": rel(m[5:], f[5:])}


def bench_ensemble(a):
    """Batched ensemble: a.members independent 2D blasts (grid N^2 with N
    log-uniform in [64,512] in multiples of 16, pressure ratio log-uniform in
    [10,1000], |B| in [0.1,1] at a random angle, MLAU or HLLD by coin flip; five
    splitmix64 draws per member, seed 3), MUSCL-minmod (ODR=2), CT, SSP-RK3,
    CFL 0.4 with vfast 1, each run to t=0.02 (last step clipped).  The timed
    region runs every member to its end time; inputs are resident before it.
    Value: interior zone-updates per second summed over members."""
    import mlau_amd as M
    from mlau_amd import ensemble as E
    from mlau_amd import problems as P
    draws = P.blast_ensemble_draws(a.members, seed=3)
    tend = 0.02

    def add_all(ens, sel):
        ids = []
        for k in sel:
            d = draws[k]
            f, p = P.blast_2d(d["n"], d["pratio"], d["bmag"], d["bang"])
            sch = M.scheme(2, 3, d["flux"], 1, precision=a.precision)
            ids.append(ens.add(f, E.member(f, sch, p["dx"], p["dy"], tend=tend, cfl=0.4, vfloor=p["vfloor"],
                                           gamma=p["gamma"])))
        return ids

    budget = 96 << 30
    ens = E.Ensemble(budget, graph=not a.no_graph)
    ids = add_all(ens, range(a.members))
    tune = ens.autotune(steps=8) if a.autotune else a.tune
    if tune and not a.autotune:
        ens.set_tune(tune)
    ens.run(0)  # build the graph outside the timed region
    t0 = time.perf_counter()
    ens.run_all(chunk=64)
    t1 = time.perf_counter()
    el = t1 - t0
    zu = ens.zone_updates()
    steps = [ens.status(i)[3] for i in ids]
    bad = [i for i in ids if ens.status(i)[0] != E.DONE]
    if bad:
        print("bench.py: %d members did not finish" % len(bad), file=sys.stderr)
        sys.exit(3)
    err_sample = list(range(a.members))
    kept = {k: ens.download(ids[k]) for k in err_sample} if a.precision != "fp64" else {}
    ens.close()
    mixed_err = None
    if kept:  # the sampled members again in FP64: relative L1 of the mixed results
        ref = E.Ensemble(budget, graph=True)
        a_prec, a.precision = a.precision, "fp64"
        rids = add_all(ref, err_sample)
        a.precision = a_prec
        ref.run_all(chunk=64)
        errs = [mixed_errors(kept[k], ref.download(i), 5.0 / 3.0) for k, i in zip(err_sample, rids)]
        ref.close()
        mixed_err = {q: max(e[q] for e in errs) for q in ("rho", "p", "B")}
        mixed_err["members"] = len(errs)
    # per-kernel times from an eager pass over the same job (not the timed run)
    ens = E.Ensemble(budget, graph=False)
    add_all(ens, range(a.members))
    ens.set_tune(tune)
    ens.kernel_timing(True)
    ens.run_all(chunk=64)
    stats = ens.kernel_stats()
    ens.close()
    kernels = {k: {"avg_us": v[0] / v[1] * 1e3, "launches": v[1], "share": v[0] / sum(x[0] for x in stats.values())}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1][0])}
    seq = None
    if a.sequential_every > 0:
        sel = list(range(0, a.members, a.sequential_every))
        sz, st = 0, 0.0
        for k in sel:
            one = E.Ensemble(1 << 30, graph=True)
            add_all(one, [k])
            one.set_tune(tune)
            one.run(0)
            s0 = time.perf_counter()
            one.run_all(chunk=64)
            st += time.perf_counter() - s0
            sz += one.zone_updates()
            one.close()
        seq = {"members": len(sel), "zone_updates_per_s": sz / st,
               "speedup": (zu / el) / (sz / st)}
    cpu = None
    if not a.no_cpu:
        from oracle import oracle as O
        nth = _cpu_threads()
        czu, ct, done = 0, 0.0, 0
        for k in range(0, a.members, 64):  # members 0, 64, 128, ... while the sample stays within the budget
            d = draws[k]
            f, p = P.blast_2d(d["n"], d["pratio"], d["bmag"], d["bang"])
            lib, kind = _ref_or_port(2, "o2r3_" + d["flux"])
            tim, nst = 0.0, 0
            t0 = time.perf_counter()
            while True:
                vm = O.cfl_vmax_2d(f, 4, p["gamma"], p["vfloor"])
                dt = 0.4 * min(p["dx"], p["dy"]) / vm
                last = tim + dt >= tend
                if last:
                    dt = tend - tim
                if lib:
                    O.ref_step_2d(lib, f, dt, p["dx"], p["dy"], 4, p["gamma"])
                else:
                    O.step_2d(f, dt, p["dx"], p["dy"], 4, p["gamma"], O.scheme(2, 3, d["flux"], 1))
                tim = tend if last else tim + dt
                nst += 1
                if last:
                    break
            ct += time.perf_counter() - t0
            czu += nst * d["n"] ** 2
            done += 1
            if ct > a.cpu_seconds:
                break
        cpu = {"value": czu / ct, "unit": "zone-updates/s", "cores": nth, "kind": kind,
               "sample": "%d members (every 64th) run to t=0.02 on the host (%.1f s wall)" % (done, ct)}
    res = {
        "metric": "zone-updates/s & %HBM peak (2D 2048\u00b2 FP64); batched solves/s, p50/p99 latency",
        "value": zu / el,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": max(steps),
        "warmup": 0,
        "ms_per_step": el / max(steps) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded blast initial states)",
        "config": {"workload": "batched ensemble: %d 2D periodic blasts, N in [64,512], MLAU/HLLD, MUSCL (ODR=2), "
                               "CT, SSP-RK3, CFL 0.4, t=0.02" % a.members, "members": a.members,
                   "zones": int(sum(d["n"] ** 2 for d in draws)), "parallelism": "single GPU", "tune": tune},
        "solves_per_s": a.members / el,
        "wall_s": el,
        "zone_updates": zu,
        "member_steps": {"min": min(steps), "max": max(steps), "mean": float(np.mean(steps))},
        "hbm_frac_step": zu / el * BYTES_PER_ZONE_UPDATE / 1e9 / HBM_PEAK_GBS,
        "one_at_a_time": seq,
        "kernels": kernels,
    }
    if mixed_err:
        res["rel_l1_mixed_vs_fp64_max"] = mixed_err
    if cpu:
        res["cpu_baseline"] = cpu
    print(json.dumps(res))


def _ref_or_port(dim, name):
    """Step function of the reference build oracle/_ref/libref<dim>d_<name>.so
    when present ("reference"), else of the bit-identical restatement ("port")."""
    from oracle import oracle as O
    lib = "ref%dd_%s" % (dim, name)
    if os.path.exists(os.path.join(O.REF_DIR, "lib%s.so" % lib)):
        return lib, "reference"
    return None, "port"


def bench_riemann1d(a):
    """1D coplanar MHD Riemann problem (Brio-Wu states, gamma=2,
    Bx=0.75) on 1024 cells with open boundaries, MLAU + MUSCL-minmod + SSP-RK3,
    one dt = 0.4*dx/vmax from the initial state (the 1D driver's rule, vfast 1)
    and the driver's output cadence to t = 0.1 (nrec*10 = 960 steps).  Times the
    whole run on the GPU (state resident) and the reference's own 1D solver on
    the host for the same run, and checks that both end in the same bits."""
    import mlau_amd as M
    from mlau_amd import problems as P
    from oracle import oracle as O
    f0, p = P.brio_wu(1024)
    s = M.Solver(f0, M.scheme(2, 3, "mlau"), p["dx"], gamma=p["gamma"], bx=p["bx"])
    dt, _ = s.cfl(0.4, 1.0)
    n = P.brio_wu_steps(dt)
    s.step(a.warmup)
    s.sync()
    s.upload(f0)
    s.set_dt(dt)
    s.sync()
    t0 = time.perf_counter()
    s.step(n)
    s.sync()
    el = time.perf_counter() - t0
    g = s.download()
    s.close()
    nth = _cpu_threads()
    lib, kind = _ref_or_port(1, "shock_o2r3_mlau")
    c = f0.copy()
    t0 = time.perf_counter()
    for _ in range(n):
        if lib:
            O.ref_step_1d(lib, c, p["bx"], dt, p["dx"], 4, p["gamma"])
        else:
            O.step_1d(c, p["bx"], dt, p["dx"], 4, p["gamma"], O.scheme(2, 3, "mlau", 1))
    cel = time.perf_counter() - t0
    zones = f0.shape[1] - 8
    res = {
        "metric": "zone-updates/s & %HBM peak (2D 2048\u00b2 FP64); batched solves/s, p50/p99 latency",
        "value": n * zones / el,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": a.warmup,
        "ms_per_step": el / n * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Brio-Wu states)",
        "config": {"workload": "1D Brio-Wu Riemann problem, 1024 cells, open BC, MLAU, MUSCL-minmod, SSP-RK3, "
                               "CFL 0.4, t=0.1 (%d steps)" % n, "grid": [1024], "parallelism": "single GPU"},
        "wall_s": el,
        "gpu_equals_cpu_bitwise": bool(np.array_equal(g, c)),
        "cpu_baseline": {"value": n * zones / cel, "unit": "zone-updates/s", "cores": nth, "kind": kind,
                         "sample": "the full run (%d steps of 1024 cells, %.2f s wall)" % (n, cel)},
    }
    print(json.dumps(res))


def bench_shear3d(a):
    """3D low-Mach double shear layer, precision comparison: 256^3 periodic
    box, gamma=5/3, rho=1, M=0.01, p=1/(gamma M^2), beta=100 (seeded noise,
    seed 2); MLAU, WCNS4 (ODR=4), CT, SSP-RK3, 500 steps with the device
    cflcheck after every step.  Runs the plain FP64 path and the mixed path
    from the same state and reports zone-updates/s of each, relative L1 of
    rho, p and |B| (mixed vs FP64) and max |div B| (4th-order stencil)."""
    import mlau_amd as M
    from mlau_amd import problems as P
    n = 256 if a.n == 2048 else a.n
    steps = 500 if a.steps == 2000 else a.steps
    f0, p = P.shear_layer_3d(n)
    res_runs = {}
    out = {}
    for prec in ("fp64", "mixed"):
        s = M.Solver(f0, M.scheme(4, 3, "mlau", 1, precision=prec), p["dx"], p["dy"], p["dz"], gamma=p["gamma"])
        s.cfl(0.4, p["vfloor"])
        s.step_cfl(2, 0.4, p["vfloor"])  # warm-up, then restart from t=0
        s.sync()
        s.upload(f0)
        s.set_time(0.0, 0)
        s.cfl(0.4, p["vfloor"])
        s.sync()
        t0 = time.perf_counter()
        s.step_cfl(steps, 0.4, p["vfloor"])
        s.sync()
        el = time.perf_counter() - t0
        out[prec] = s.download()
        res_runs[prec] = {"zone_updates_per_s": steps * n ** 3 / el, "ms_per_step": el / steps * 1e3,
                          "t_end": s.time()[0]}
        s.close()
    a64, amx = out["fp64"][:, 4:-4, 4:-4, 4:-4], out["mixed"][:, 4:-4, 4:-4, 4:-4]

    def pres(f):
        b2 = f[5] ** 2 + f[6] ** 2 + f[7] ** 2
        return (p["gamma"] - 1) * (f[4] - 0.5 * (f[1] ** 2 + f[2] ** 2 + f[3] ** 2) / f[0] - 0.5 * b2)

    def rel(x, y):
        return float(np.abs(x - y).sum() / np.abs(y).sum())

    def divb(f):
        d = lambda q, ax: (27.0 * (np.roll(q, -1, ax) - q) - (np.roll(q, -2, ax) - np.roll(q, 1, ax))) / 24.0
        return float(np.abs(sum(d(f[5 + k], 2 - k) for k in range(3)) / p["dx"])[4:-4, 4:-4, 4:-4].max())

    errs = {"rho": rel(amx[0], a64[0]), "p": rel(pres(amx), pres(a64)),
            "B": rel(np.sqrt((amx[5:8] ** 2).sum(0)), np.sqrt((a64[5:8] ** 2).sum(0)))}
    res = {
        "metric": "zone-updates/s & %HBM peak (2D 2048\u00b2 FP64); batched solves/s, p50/p99 latency",
        "value": res_runs["mixed"]["zone_updates_per_s"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 2,
        "ms_per_step": res_runs["mixed"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (analytic shear layer + seeded noise)",
        "config": {"workload": "3D low-Mach double shear layer %d^3, M=0.01, beta=100, MLAU, WCNS4, CT, SSP-RK3, "
                               "%d steps, FP64 vs mixed" % (n, steps), "grid": [n, n, n],
                   "parallelism": "single GPU"},
        "runs": res_runs,
        "rel_l1_mixed_vs_fp64": errs,
        "max_divB": {"fp64": divb(out["fp64"]), "mixed": divb(out["mixed"])},
        "tolerance": 1e-6,
    }
    if not a.no_cpu:
        from oracle import oracle as O
        nth = _cpu_threads()
        lib, kind = _ref_or_port(3, "o4r3_mlau")
        c, pc = P.shear_layer_3d(64)
        dtc = 0.4 * pc["dx"] / O.cfl_vmax_3d(c, 4, pc["gamma"], pc["vfloor"])
        t0 = time.perf_counter()
        for _ in range(50):
            if lib:
                O.ref_step_3d(lib, c, dtc, pc["dx"], pc["dy"], pc["dz"], 4, pc["gamma"])
            else:
                O.step_3d(c, dtc, pc["dx"], pc["dy"], pc["dz"], 4, pc["gamma"], O.scheme(4, 3, "mlau", 1))
        cel = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": 50 * 64 ** 3 / cel, "unit": "zone-updates/s", "cores": nth, "kind": kind,
                               "sample": "64^3 shear layer, 50 FP64 steps at the initial dt (%.1f s wall)" % cel}
    print(json.dumps(res))


def bench_serve(a):
    """Serving stream: a.requests seeded solve requests (1D Riemann nx 256-8192
    to t=0.1, 2D blasts 64^2-1024^2 to t=0.02, 3D low-Mach shear layers
    32^3-192^3 with M in [0.001,0.1] to t=2e-4; seed 4) arriving as a Poisson
    process at each rate, solved by the solve server in a.precision under a
    200 GB budget.  Per rate: p50/p99 time-to-solution (arrival to result),
    solves/s and zone-updates/s over the makespan.  Baseline: the same
    requests (every 4th) solved one at a time in plain FP64."""
    import mlau_amd as M
    from mlau_amd import ensemble as E
    from mlau_amd import problems as P
    from mlau_amd.server import Server
    import threading
    draws = P.serve_request_draws(a.requests, seed=4)

    def build(d, precision):
        f, p, sc = P.serve_request(d)
        sch = M.scheme(sc["odr"], sc["rk"], sc["flux"], sc["ctw"], bc1d=sc.get("bc1d", "open"), precision=precision)
        m = E.member(f, sch, p["dx"], p.get("dy"), p.get("dz"), tend=P.SERVE_TEND[d["kind"]], cfl=0.4,
                     vfloor=p.get("vfloor", 0.0), gamma=p["gamma"], bx=p.get("bx", 0.0))
        return f, m

    # Mixed serving: the 2D blasts run hundreds of steps through strong shocks, where FP32
    # Riemann-solver rounding accumulates past 1e-6 relative L1; they use the FP32
    # reconstruction with FP64 solvers.  1D and 3D requests keep both in FP32.
    prec_of = {k: a.precision for k in (1, 2, 3)}
    if a.precision == "mixed":
        prec_of[2] = "mixed_recon"
    reqs = [build(d, prec_of[d["kind"]]) for d in draws]
    budget = 200 * 10 ** 9
    out = []
    served = {}
    for rate in [float(x) for x in a.rates.split(",")]:
        gaps = np.random.default_rng(4).exponential(1.0 / rate, size=len(reqs))
        # rounds of 4 steps: arrivals join sooner (p50 -20 % against 8) at the same throughput;
        # the one-at-a-time baseline below keeps 8, which suits a single request
        srv = Server(budget_bytes=budget, steps_per_round=4)
        tickets = [None] * len(reqs)
        base = srv.clock() + 0.05

        def feed():
            at = base
            for k, (f, m) in enumerate(reqs):
                at += gaps[k]
                wait = at - srv.clock()
                if wait > 0:
                    time.sleep(wait)
                tickets[k] = srv.submit(f, m)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        tts, zu, fin, sub, st, kz = [], 0, [], [], [], {}
        for k, t in enumerate(tickets):
            st.append(srv.wait(t, timeout=1200.0))
            res_f, tm = srv.result(t)
            if k % 4 == 0:
                served[k] = res_f
            tts.append(tm["finish_s"] - tm["submit_s"])
            zu += tm["zone_updates"]
            fin.append(tm["finish_s"])
            sub.append(tm["submit_s"])
            kz.setdefault(draws[k]["kind"], []).append((tm["zone_updates"], tts[-1], tm["finish_s"]))
        srv.close()
        span = max(fin) - min(sub)
        # per kind: share of the work, median time to solution, when its last request finished
        per_kind = {"%dd" % kind: {"requests": len(v), "zone_updates": int(sum(x[0] for x in v)),
                                   "p50_s": float(np.percentile([x[1] for x in v], 50)),
                                   "last_finish_s": max(x[2] for x in v) - min(sub)}
                    for kind, v in sorted(kz.items())}
        out.append({"rate": rate, "p50_s": float(np.percentile(tts, 50)), "p99_s": float(np.percentile(tts, 99)),
                    "solves_per_s": len(reqs) / span, "zone_updates_per_s": zu / span, "makespan_s": span,
                    "failed": int(sum(1 for x in st if x != E.DONE)), "per_kind": per_kind})
    # one request at a time, plain FP64 (every 4th request)
    srv = Server(budget_bytes=budget, steps_per_round=8)
    seq_t, seq_zu = 0.0, 0
    seq_kind = {}
    errs = {1: [], 2: [], 3: []}
    for k in range(0, len(draws), 4):
        f, m = build(draws[k], "fp64")
        t0 = srv.clock()
        t = srv.submit(f, m)
        srv.wait(t, timeout=1200.0)
        res64, tm = srv.result(t)
        dt_solo = srv.clock() - t0
        seq_t += dt_solo
        seq_zu += tm["zone_updates"]
        sk = seq_kind.setdefault("%dd" % draws[k]["kind"], {"requests": 0, "seconds": 0.0, "zone_updates": 0})
        sk["requests"] += 1
        sk["seconds"] += dt_solo
        sk["zone_updates"] += int(tm["zone_updates"])
        if a.precision != "fp64" and k in served:
            g = m.grid
            errs[draws[k]["kind"]].append(mixed_errors(served[k], res64, g.gamma, bx=g.bx))
    srv.close()
    mixed_err = {("%dd" % kind): {q: max(e[q] for e in v) for q in ("rho", "p", "B")} for kind, v in errs.items() if v}
    seq = {"zone_updates_per_s": seq_zu / seq_t, "solves_per_s": len(range(0, len(draws), 4)) / seq_t,
           "requests": len(range(0, len(draws), 4)), "per_kind": seq_kind}
    best = max(out, key=lambda r: r["zone_updates_per_s"])
    res = {
        "metric": "zone-updates/s & %HBM peak (2D 2048\u00b2 FP64); batched solves/s, p50/p99 latency",
        "value": best["zone_updates_per_s"],
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded request stream)",
        "config": {"workload": "solve server: Poisson stream of %d requests (1D Riemann, 2D blasts, 3D low-Mach "
                               "shear layers) per rate, 200 GB budget" % len(reqs), "parallelism": "single GPU",
                   "precision_per_kind": {"%dd" % k: v for k, v in prec_of.items()},
                   "rates": [r["rate"] for r in out], "kinds": {k: sum(1 for d in draws if d["kind"] == k)
                                                                 for k in (1, 2, 3)}},
        "per_rate": out,
        "one_at_a_time_fp64": seq,
        "speedup_vs_one_at_a_time_fp64": best["zone_updates_per_s"] / seq["zone_updates_per_s"],
    }
    if mixed_err:  # every 4th request: served mixed result against its FP64 one-at-a-time result
        res["rel_l1_mixed_vs_fp64_max"] = mixed_err
    print(json.dumps(res))


def bench_serve_batch(a):
    """The serving stream's requests (bench_serve's draws and precisions) as
    one batch: every request added to one ensemble with the server engine's
    launch shape (balance mode, 16 CUs left out of the persistent stage
    kernels, cheap groups up to 256 steps per step of the costliest: ensemble
    tune (1 << 11) | (2 << 9), as mlau_server.hip sets it), run to completion
    with inputs resident.  The kernels are the served ones, launched without
    the server's thread, so the job can be traced kernel by kernel; the
    per-kernel times come from a separate eager pass over the same batch."""
    import mlau_amd as M
    from mlau_amd import ensemble as E
    from mlau_amd import problems as P
    server_tune = (1 << 11) | (2 << 9)
    draws = P.serve_request_draws(a.requests, seed=4)
    prec_of = {k: a.precision for k in (1, 2, 3)}
    if a.precision == "mixed":
        prec_of[2] = "mixed_recon"

    def add_all(ens):
        ids = []
        for d in draws:
            f, p, sc = P.serve_request(d)
            sch = M.scheme(sc["odr"], sc["rk"], sc["flux"], sc["ctw"], bc1d=sc.get("bc1d", "open"),
                           precision=prec_of[d["kind"]])
            ids.append(ens.add(f, E.member(f, sch, p["dx"], p.get("dy"), p.get("dz"), tend=P.SERVE_TEND[d["kind"]],
                                           cfl=0.4, vfloor=p.get("vfloor", 0.0), gamma=p["gamma"],
                                           bx=p.get("bx", 0.0))))
        return ids

    budget = 200 * 10 ** 9
    ens = E.Ensemble(budget, graph=True, balance=True)
    ids = add_all(ens)
    ens.set_tune(server_tune)
    ens.run(0)  # build the graphs outside the timed region
    t0 = time.perf_counter()
    ens.run_all(chunk=8)
    el = time.perf_counter() - t0
    zu = ens.zone_updates()
    bad = [i for i in ids if ens.status(i)[0] != E.DONE]
    ens.close()
    if bad:
        print("bench.py: %d requests did not finish" % len(bad), file=sys.stderr)
        sys.exit(3)
    ens = E.Ensemble(budget, graph=False, balance=True)
    add_all(ens)
    ens.set_tune(server_tune)
    ens.kernel_timing(True)
    ens.run_all(chunk=8)
    stats = ens.kernel_stats()
    ens.close()
    tot = sum(v[0] for v in stats.values())
    kernels = {k: {"avg_us": v[0] / v[1] * 1e3, "launches": v[1], "share": v[0] / tot}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1][0])}
    res = {
        "metric": "zone-updates/s & %HBM peak (2D 2048\u00b2 FP64); batched solves/s, p50/p99 latency",
        "value": zu / el,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": None,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded request stream)",
        "config": {"workload": "serving stream's %d requests as one batch (1D Riemann, 2D blasts, 3D low-Mach "
                               "shear layers), server engine shape, no server thread" % len(draws),
                   "parallelism": "single GPU", "precision_per_kind": {"%dd" % k: v for k, v in prec_of.items()},
                   "tune": server_tune},
        "solves_per_s": len(draws) / el,
        "wall_s": el,
        "zone_updates": zu,
        "kernels": kernels,
    }
    print(json.dumps(res))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.workload == "ensemble":
        bench_ensemble(a)
        return
    if a.workload == "serve":
        if a.serve_batch:
            bench_serve_batch(a)
        else:
            bench_serve(a)
        return
    if a.workload == "shear3d":
        bench_shear3d(a)
        return
    if a.workload == "riemann1d":
        bench_riemann1d(a)
        return
    import mlau_amd as M
    from mlau_amd import problems as P

    n = a.n
    f0, p = P.vortex_2d(n, 4)
    sch = M.scheme(4, 3, "lhlld", 1, precision=a.precision)
    cfl, vfloor = 0.4, p["vfloor"]

    def fresh():
        s = M.Solver(f0, sch, p["dx"], p["dy"], gamma=p["gamma"])
        s.cfl(cfl, vfloor)
        return s

    # warmup (compiles nothing; captures the step graph, settles clocks and
    # caches), then restart from t=0 for the timed steps: graph replay, no
    # per-kernel events in the timed region
    s = fresh()
    s.step_cfl(a.warmup, cfl, vfloor)
    s.sync()
    s.upload(f0)
    s.set_time(0.0, 0)
    s.cfl(cfl, vfloor)
    s.sync()
    t0 = time.perf_counter()
    s.step_cfl(a.steps, cfl, vfloor)
    s.sync()
    t1 = time.perf_counter()
    out = s.download()
    s.close()
    if not np.all(np.isfinite(out)):
        print("bench.py: non-finite state after the timed steps", file=sys.stderr)
        sys.exit(3)
    el = t1 - t0
    zones = n * n
    zu = a.steps * zones / el
    # per-kernel durations: a separate pass over the same workload from t=0
    # (eager launches with HIP events on the handle's stream around each one)
    s = fresh()
    s.kernel_timing(True)
    s.step_cfl(min(a.steps, 60), cfl, vfloor)
    s.sync()
    stats = s.kernel_stats()
    s.close()
    # algorithmic bytes of each kernel per launch: the arrays it must read and
    # write once per interior cell, 8 B each (update2d: 21 + 8 in the first RK
    # stage, which reads no step-start state, 29 + 8 in the other two)
    kern_arrays = {"ysweep2d_inl": 8 + 9, "xsweep2d_inl": 8 + 8, "ysweep2d": 11 + 8, "xsweep2d": 11 + 8,
                   "update2d": (29 + 37 + 37) / 3.0, "prim2d": 8 + 7, "aux2d": 5 + 3, "primaux2d": 8 + 10,
                   "ez2d": 5 + 1, "cfl2d": 8, "bc2d": 0}
    dom = max(stats.items(), key=lambda kv: kv[1][0])
    dname, (dms, dcnt) = dom
    davg_s = dms / dcnt / 1e3
    dbytes = 8 * kern_arrays.get(dname, 0) * zones
    achieved = dbytes / davg_s / 1e9
    prof = PROFILE_2048 if (n == 2048 and a.precision == "fp64") else {}
    traffic = prof.get("traffic", {}).get(dname)
    kernels = {k: {"avg_us": v[0] / v[1] * 1e3, "launches": v[1], "share": v[0] / sum(x[0] for x in stats.values())}
               for k, v in sorted(stats.items(), key=lambda kv: -kv[1][0])}
    res = {
        "metric": "zone-updates/s & %HBM peak (2D 2048\u00b2 FP64); batched solves/s, p50/p99 latency",
        "value": zu,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic (analytic vortex initial state)",
        "config": {"workload": "2D periodic vortex %dx%d, LHLLD flux, WCNS4 (ODR=4), CT, SSP-RK3, CFL 0.4, "
                               "dt from device cflcheck every step" % (n, n), "grid": [n, n], "ghosts": 4,
                   "parallelism": "single GPU"},
        "hbm_frac_step": zu * BYTES_PER_ZONE_UPDATE / 1e9 / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": dbytes, "avg_launch_us": davg_s * 1e6,
                     "timing": "timed steps replay the step graph; per-kernel averages from a separate eager pass "
                               "of %d steps from t=0" % min(a.steps, 60),
                     "traffic_source": prof.get("traffic_file")},
        "kernels": kernels,
    }
    fp64 = prof.get("fp64", {})
    if dname in fp64:
        rate = fp64[dname] / davg_s / 1e9
        res["valu_roofline"] = {"bound": "fp64-valu", "kernel": dname, "achieved": rate, "peak": 614.4,
                                "unit": "G FP64 wave-instructions/s", "frac": rate / 614.4,
                                "source": prof.get("fp64_file")}
    if not a.no_cpu:
        try:
            res["cpu_baseline"] = cpu_baseline(a.cpu_seconds)
        except Exception as e:  # reported, not fatal for the GPU number
            res["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
