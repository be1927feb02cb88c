"""Benchmark of the per-RK-stage pipeline on one MI355X.

Workload: 256^3 triply periodic decaying isotropic turbulence, FP64
(k0 = 8, M_t = 0.5, u_rms = 1, Re_lambda ~ 40 at t = 0, Pr = 0.7,
vis_exp = 0.76, artificial transport on, CFL 0.5).  A benchmark step is one
march-mode-1 step of the reference driver: Compute_CFL, Compute_Tstep, then
LDDRK4 = 5 + 6 RK stages (RHS, Hyper_RHS, update, compact filter,
pressure fix).  The metric is grid points advanced through one RK stage
per second (Mpts*stage/s), with the state resident in HBM.

Beside the headline it runs two bounded legs: a batched ensemble of 48
instances of 48^3..128^3 against the same instances one at a time, and a
finite Poisson trace of runs served with continuous batching against the
plain path one at a time (p50/p99 time to completion, throughput, and a
check that every run's result is the same bits alone and in the mix).

Prints one JSON line.  With --gpus N > 1 it exits: this build measures a
single GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "shock-turbulence_amd", "python"))

B_ALG = 2136.0          # algorithmic bytes per grid point per RK stage (FP64)
HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
STAGES_PER_STEP = 11

# kernel template of each pass kind (csrc/st_pass_*.hip)
PASS_STRUCT = {
    "lap2": "PassLap2", "lap2_gauss": "PassLap2Gauss", "lap_se_lap2": "PassLapSELap2", "gauss_eta": "PassGaussEta",
    "gauss_hdiv": "PassGaussHdiv", "hyper_div": "PassHyperDiv", "rk_filter": "PassFilterQ<0>",
    "filter_q": "PassFilterQ<1>", "filter_q_fix": "PassFilterQ<2>", "hdiv_rkf": "PassHdivRkf",
}


def kernel_symbol(kind, d, narrow, nt, n):
    """The kernel symbol (as rocprofv3 prints it) that runs pass kind.d at
    lines of n points (n a power of two, 16..256) with the given layout."""
    M = 8 if n >= 64 else max(2, n // 8)
    P = n // M
    if narrow and 2 * P <= 64:
        M, P = M // 2, 2 * P
    if kind == "grad" and d == "x":
        st = "PassGrad<1, true, false, 0, false>"
    elif kind == "grad":
        if nt == 256 and M == 8 and P == 32:
            st = "PassGrad<2, true, false, 2, true>"
        else:
            st = "PassGrad<1, %s, false, 2, false>" % ("true" if nt == 1024 else "false")
    elif kind == "grad_cross":
        st = "PassGrad<1, true, true, 4, false>" if nt == 1024 else "PassGrad<2, false, true, 4, false>"
    elif kind == "lap_se":
        st = "PassLapSE<%s>" % ("false" if d == "x" else "true")
    else:
        st = PASS_STRUCT.get(kind)
        if st is None:
            return kind
    return "k_pass<%d, %d, %d, false, %s>" % (M, P, nt, st)


def traffic_of(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary (profiles/traffic_256.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_256.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def copy_bandwidth(nbytes=1 << 31, reps=5):
    """Measured device copy bandwidth (GB/s of bytes read + written): the
    faster of a device-to-device copy and an elementwise scaled copy of
    nbytes, timed with events over reps launches."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    best = 0.0
    for op in (lambda: y.copy_(x), lambda: torch.mul(x, 1.0, out=y)):
        op()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            op()
        e1.record()
        torch.cuda.synchronize()
        best = max(best, 2.0 * nbytes / (e0.elapsed_time(e1) / reps * 1e-3) / 1e9)
    del x, y
    torch.cuda.empty_cache()
    return round(best, 1)


def host_cpu():
    """nproc, the CPU model name (lscpu / /proc/cpuinfo) and the threads this
    process may use (its CPU affinity and OMP_NUM_THREADS)."""
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if not model:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = nproc
    omp = os.environ.get("OMP_NUM_THREADS")
    granted = min(avail, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else avail
    return {"nproc": nproc, "affinity_cpus": avail, "model": model, "threads": granted}


def cpu_baseline(n, threads, Q, kw):
    """Time the CPU oracle (C restatement, OpenMP) over one full step of the
    same workload (CFL, dt, the 11 RK stages) from the same initial state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    host = host_cpu()
    if threads <= 0:
        threads = host["threads"]
    O.lib().or_threads(threads)
    o = O.Oracle(n, n, n, **kw)
    t0 = time.perf_counter()
    o.march(Q, 1)
    el = time.perf_counter() - t0
    pts = float(n) ** 3
    return {"value": round(pts * STAGES_PER_STEP / el / 1e6, 3), "unit": "Mpts*stage/s", "cores": threads,
            "kind": "port", "nproc": host["nproc"], "cpu_model": host["model"],
            "sample": "one full step (Compute_CFL, Compute_Tstep, the %d RK stages: RHS + Hyper_RHS + update + "
                      "filter + Fix_Pressure each) at %d^3 from the same initial state, oracle/st_oracle.c with %d "
                      "OpenMP threads (the CPUs this process may use: %d of nproc %d), %.2f s"
                      % (STAGES_PER_STEP, n, threads, host["affinity_cpus"], host["nproc"], el)}


def ensemble_leg(nsteps, autotune):
    """Batched ensemble (48 instances: 16 x 48^3, 16 x 64^3, 12 x 96^3, 4 x
    128^3) against the same instances one at a time on the plain path."""
    import st_amd
    import st_serve
    runs = st_serve.ensemble_config3(nsteps)
    for r in runs:
        r.make_ic()
    work = sum(r.n ** 3 * STAGES_PER_STEP * nsteps for r in runs)
    ens = st_serve.Ensemble(runs, autotune=autotune)
    tuned = dict(ens.tuned)
    ens.march(0)
    # each size's batch on its own for the first few steps (its batched rate
    # alone), then all batches together for the rest (the mix)
    k_alone = min(5, nsteps - 1)
    alone = {}
    for n, p in ens.pools.items():
        t0 = time.perf_counter()
        p.batch.march(k_alone)
        p.batch.sync()
        alone[n] = len(p.running) * n ** 3 * STAGES_PER_STEP * k_alone / (time.perf_counter() - t0) / 1e6
    t0 = time.perf_counter()
    ens.march(nsteps - k_alone)
    el = time.perf_counter() - t0
    ens.close()
    mix_rate = (work * (nsteps - k_alone) / nsteps) / el / 1e6
    # the same instances as one group: one launch per pass over all four
    # sizes (one step to capture the graph, then the same step count)
    grp = st_serve.Ensemble(runs, grouped=True)
    grp.march(1)
    t0 = time.perf_counter()
    grp.march(nsteps - k_alone)
    el_g = time.perf_counter() - t0
    grp.close()
    group_rate = (work * (nsteps - k_alone) / nsteps) / el_g / 1e6
    plain_t = {}
    for r in runs:
        s = st_amd.Solver(r.n, r.n, r.n, Re=r.Re, Pr=0.7, hyper_param=1, cfl_cal=0.5)
        if r.n in tuned:
            s.set_pass_configs(tuned[r.n])    # the batch's launch shapes: same rounding, same bits
        s.upload(r.Q)
        s.sync()
        t0 = time.perf_counter()
        s.march(nsteps)
        s.sync()
        plain_t[r.n] = plain_t.get(r.n, 0.0) + time.perf_counter() - t0
        s.close()
    plain = sum(plain_t.values())
    plain_rate = work / plain / 1e6
    per_size = {}
    for n in sorted(plain_t):
        w = sum(r.n ** 3 * STAGES_PER_STEP * nsteps for r in runs if r.n == n)
        per_size[str(n)] = {"runs": sum(1 for r in runs if r.n == n), "work_share": round(w / work, 4),
                            "plain_mpts_stage_per_s": round(w / plain_t[n] / 1e6, 1),
                            "batched_alone_mpts_stage_per_s": round(alone[n], 1)}
    return {"workload": "48 instances: 16x48^3, 16x64^3, 12x96^3, 4x128^3, seeds 1..48, k0 in {3,4,6}, "
                        "Mt in {0.2,0.3,0.5}, Re_lambda=40; %d steps each" % nsteps,
            "points": sum(r.n ** 3 for r in runs), "steps": nsteps,
            "batched_mpts_stage_per_s": round(mix_rate, 1),
            "grouped_mpts_stage_per_s": round(group_rate, 1),
            "plain_one_at_a_time_mpts_stage_per_s": round(plain_rate, 1),
            "speedup": round(mix_rate / plain_rate, 3),
            "speedup_grouped": round(group_rate / plain_rate, 3), "autotune_reps": autotune,
            "per_size": per_size,
            # what the one-at-a-time rate per size allows: a speedup of x
            # over it needs the mix to run at x times the plain rate
            "ceiling": {"rate_needed_for_2x_mpts_stage_per_s": round(2 * plain_rate, 1),
                        "rate_needed_for_2x_frac_of_hbm_roofline": round(2 * plain_rate * B_ALG * 1e6 / 1e9
                                                                         / HBM_PEAK_GBS, 3),
                        "batched_mix_frac_of_hbm_roofline": round(mix_rate * B_ALG * 1e6 / 1e9 / HBM_PEAK_GBS, 3)}}


def config2_leg(nsteps, warmup, autotune=3):
    """512x192x192 with a non-periodic x: a Mach-1.5 shock at x = L/4 with
    upstream turbulence and Rankine-Hugoniot sponge layers (20% of x at
    each end), advanced with the non-periodic link set."""
    import math
    import numpy as np
    import st_amd
    nx, ny, nz = 512, 192, 192
    d = 2 * math.pi / ny
    per = (d * (nx - 1), 2 * math.pi, 2 * math.pi)
    t0 = time.perf_counter()
    Q, s1, s2 = st_amd.ic_shock(nx, ny, nz, per, mach=1.5, frac=0.25, k0=4.0, urms=0.1, seed=2)
    t_ic = time.perf_counter() - t0
    s = st_amd.Solver(nx, ny, nz, period=per, nonperiodic_x=1, Re=500.0, Pr=0.7, hyper_param=1, cfl_cal=0.5)
    try:
        s.sponge(5.0, 2.0, 0.4, s1, s2)
        s.upload(Q)
        if autotune:
            s.autotune(autotune)          # set-up: launch shapes for this grid (restores the state)
        t, st, _ = s.march(warmup)
        s.sync()
        t0 = time.perf_counter()
        t, st, dts = s.march(nsteps, t, st)
        s.sync()
        el = time.perf_counter() - t0
        finite = bool(np.isfinite(s.download()).all())
    finally:
        s.close()
    pts = float(nx * ny * nz)
    return {"workload": "512x192x192, x non-periodic (closures, mirrored Gaussian, constant hyper coefficients), "
                        "Mach 1.5 shock at x=L/4, upstream turbulence k0=4 u_rms=0.1, sponge amp 5 power 2 over 20% "
                        "of x at each end toward the Rankine-Hugoniot states, Re=500, Pr=0.7, CFL 0.5",
            "steps": nsteps, "warmup": warmup, "ic_seconds": round(t_ic, 2), "autotune_reps": autotune,
            "mpts_stage_per_s": round(pts * nsteps * STAGES_PER_STEP / el / 1e6, 1),
            "ms_per_step": round(el / nsteps * 1e3, 2),
            "roofline_stage_frac": round(B_ALG * pts * nsteps * STAGES_PER_STEP / el / 1e9 / HBM_PEAK_GBS, 4),
            "finite": finite, "dt_last": float(dts[-1])}


def serving_leg(nruns, rate, seed, autotune=0):
    """A finite Poisson trace of runs (sizes 48^3..128^3, 50..400 steps)
    served with continuous batching, against the plain path one at a time
    (first come, first served on the same GPU)."""
    import st_serve
    runs = st_serve.poisson_trace(nruns, rate=rate, seed=seed)
    srv = st_serve.Server(runs, budget_bytes=200e9, autotune=autotune)
    try:
        served = srv.serve()
    finally:
        srv.close()
    plain, ref = st_serve.plain_one_at_a_time(runs, configs=srv.tuned)
    same = all(bool((r.result == ref[r.rid]).all()) for r in runs)
    pr = plain["mpts_stage_per_s"]
    return {"workload": "%d runs, Poisson arrivals at %.0f/s, N in {48,64,96,128}^3 uniform, steps U[50,400], "
                        "seed %d; HBM budget 200 GB" % (nruns, rate, seed),
            "served": {k: round(v, 4) if isinstance(v, float) else v for k, v in served.items()},
            "plain_one_at_a_time": {k: round(v, 4) if isinstance(v, float) else v for k, v in plain.items()
                                    if k != "per_size"},
            "plain_per_size": plain.get("per_size"),
            "throughput_ratio": round(served["mpts_stage_per_s"] / pr, 3),
            # a served throughput of 3x the plain path's needs this rate
            "ceiling": {"rate_needed_for_3x_mpts_stage_per_s": round(3 * pr, 1),
                        "rate_needed_for_3x_frac_of_hbm_roofline": round(3 * pr * B_ALG * 1e6 / 1e9
                                                                         / HBM_PEAK_GBS, 3)},
            "results_equal_alone_and_in_mix": same, "autotune_reps": autotune}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="oracle threads for the CPU baseline (0: every CPU this process may use)")
    ap.add_argument("--profile", type=int, default=1, help="per-pass HIP-event timing (in extra steps after the timed region)")
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--ensemble-steps", type=int, default=100, help="batched-ensemble leg, steps per instance (0: skip)")
    ap.add_argument("--autotune", type=int, default=2,
                    help="ensemble and serving legs: tune each batch's launch shapes at set-up (stages per "
                         "candidate; the plain path gets the same shapes); 0: the 256^3 defaults")
    ap.add_argument("--serve-runs", type=int, default=16, help="serving leg: runs in the trace (0: skip)")
    ap.add_argument("--config2-steps", type=int, default=3, help="non-periodic 512x192x192 leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--copy-bw", type=int, default=1, help="also measure a 2 GiB device copy (GB/s), before the timed region")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 2 GiB fill and a 2 GiB copy first (reference kernels for rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; --gpus %d is not supported in this build" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import st_amd

    # torch's HIP runtime must initialise the device before this library's
    # does, so the copy is measured first
    copy_gbs = copy_bandwidth() if args.copy_bw else None
    if args.calib_copy:
        # reference kernels of known traffic for the byte counters: a 2 GiB
        # fill (2 GiB written) and a 2 GiB scaled copy (2 GiB read + written)
        import torch
        x = torch.full((1 << 28,), 1.0, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        del x, y

    n = args.n
    k0, mt, urms = 8.0, 0.5, 1.0
    re = st_amd.re_for_taylor(40.0, k0, urms, mt)
    kw = dict(Re=re, Pr=0.7, gamma=1.4, vis_exp=0.76, bulk_ratio=0.0, hyper_param=1, cfl_cal=0.5)
    t0 = time.perf_counter()
    Q = st_amd.ic_isotropic(n, n, n, k0=k0, urms=urms, mt=mt, seed=12345)
    t_ic = time.perf_counter() - t0
    s = st_amd.Solver(n, n, n, **kw)
    s.upload(Q)
    time_, step = 0.0, 0
    if args.warmup:
        time_, step, _ = s.march(args.warmup, time_, step)
    s.sync()
    t0 = time.perf_counter()
    time_, step, dts = s.march(args.steps, time_, step)
    s.sync()
    el = time.perf_counter() - t0
    # per-pass HIP-event timing in a separate run of the same steps, so
    # that the events do not perturb the timed region
    cfgs = s.pass_configs()
    prof = {}
    if args.profile:
        s.profile(True)
        time_, step, _ = s.march(args.profile_steps, time_, step)
        s.sync()
        prof = s.profile_read()
        s.profile(False)
    diag = s.diagnostics()
    Qend = s.download()
    finite = bool(np.isfinite(Qend).all())
    s.close()

    pts = float(n) ** 3
    stages = args.steps * STAGES_PER_STEP
    value = pts * stages / el / 1e6
    ms_step = el / args.steps * 1e3

    # per-kernel table (pass kind x direction), then the dominant kernel: the
    # kernel symbol with the largest share of the timed region (the y sweep of
    # "grad" is a kernel of its own, the other kinds one kernel for all
    # directions)
    roof = None
    kernels = {}
    if prof:
        groups = {}
        for name, p in prof.items():
            avg = p["ms"] / p["count"]
            alg = p["fields"] * 8.0 * pts
            kernels[name] = {"launches": p["count"], "total_ms": round(p["ms"], 3), "avg_ms": round(avg, 4),
                             "alg_bytes_per_launch": alg,
                             "achieved_GBs": round(alg / (avg * 1e-3) / 1e9, 1) if alg else None}
            if not alg:
                continue
            kind, d = name.rsplit(".", 1)
            key = kernel_symbol(kind, d, *cfgs.get(name, (0, 256)), n)
            g = groups.setdefault(key, {"ms": 0.0, "count": 0, "bytes": 0.0, "names": []})
            g["ms"] += p["ms"]
            g["count"] += p["count"]
            g["bytes"] += alg * p["count"]
            g["names"].append(name)
        dom = max(groups, key=lambda k: groups[k]["ms"])
        g = groups[dom]
        avg = g["ms"] / g["count"]
        alg = g["bytes"] / g["count"]
        ach = alg / (avg * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic_of(dom),
                "kernel": dom, "passes": sorted(g["names"]), "avg_launch_ms": round(avg, 4),
                "alg_bytes_per_launch": alg}
    stage_roof = {"B_alg_per_pt_stage": B_ALG, "achieved_GBs": round(B_ALG * pts * stages / el / 1e9, 1),
                  "frac": round(B_ALG * pts * stages / el / 1e9 / HBM_PEAK_GBS, 4)}
    if copy_gbs:
        stage_roof["measured_copy_GBs"] = copy_gbs
        stage_roof["frac_of_copy"] = round(stage_roof["achieved_GBs"] / copy_gbs, 4)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, args.cpu_threads, Q, kw)
    del Q, Qend
    ensemble = ensemble_leg(args.ensemble_steps, args.autotune) if args.ensemble_steps > 0 else None
    serving = serving_leg(args.serve_runs, args.serve_rate, 2024, args.autotune) if args.serve_runs > 0 else None
    config2 = config2_leg(args.config2_steps, 1) if args.config2_steps > 0 else None

    out = {
        "metric": "Mpts*RK-stage/s FP64 (256^3 periodic RHS+LDDRK4 stage throughput)",
        "value": round(value, 2),
        "unit": "Mpts*stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "%d^3 periodic decaying isotropic turbulence, k0=8, Mt=0.5, u_rms=1, "
                               "Re_lambda=40 at t=0 (Re=%.1f), Pr=0.7, hyper on, CFL 0.5; "
                               "step = CFL + dt + LDDRK4 (11 stages)" % (n, re),
                   "grid": [n, n, n], "stages_per_step": STAGES_PER_STEP, "ic_seconds": round(t_ic, 2)},
        "roofline": roof,
        "roofline_stage": stage_roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "ensemble": ensemble,
        "serving": serving,
        "nonperiodic_512x192x192": config2,
        "check": {"finite": finite, "sim_time": time_, "step": step, "dt_last": float(dts[-1]) if len(dts) else None,
                  "diagnostics": [float(x) for x in diag]},
        # the library's source hash and the one of the sources in this tree
        "build": {"library": st_amd.build_id(), "sources": st_amd.source_id()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
