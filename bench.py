#!/usr/bin/env python3
"""Headline benchmark of the complex-Langevin engine on one MI355X.

Metric: site-updates/s of the FP32 batched path (128 chains of 2-D 128x128
lattices, Nt = 32, FP32 fields, FP64 dt / sqrt(dt) / reductions / observables)
against the FP64 plain path (one 2-D 64x64 x Nt=64 chain, the reference's
host-driven step).  A site-update is one (x,y,z,tau) site of one chain advanced
by one Langevin step (4 components of psi and psibar).

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.  Inputs are synthetic (seeded Philox initial fields).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "site-updates/sec FP32-batched vs FP64 plain (1 GPU) + p50/p99 run latency"
HBM_PEAK_GBS = 8000.0


def _params(seed, mu=1.0, nt=32, c1=-0.05):
    from cl_var_dt_amd import default_params
    return default_params(mu=mu, G=0.1, c1=c1, Q=0.1, P=0.0, eps=1.0 / nt, DT=1e-3, seed=seed)


def measured_traffic(kernel: str, schedule: int, lattice: str, chains: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE, calibrated against a
    1 GiB copy), or None when no summary matches."""
    path = os.path.join(ROOT, "profiles", f"traffic_batched_{lattice}_fp32.md")
    if chains != 128 or not os.path.exists(path):
        return None, None
    for line in open(path):
        parts = line.split()
        if len(parts) == 4 and parts[0] == "traffic_bytes_per_launch" and parts[1] == "k_" + kernel \
                and parts[2] == f"schedule={schedule}":
            return int(parts[3]), os.path.relpath(path, ROOT)
    return None, None


# MI355X single-issue vector rate: 1024 SIMDs x 16 lanes per cycle (a wave64
# instruction holds its SIMD for 4 cycles; packed FP32 off) x 2.4 GHz peak
# engine clock.  gfx950 also issues two simple FP32 / 32-bit integer VALU
# instructions from two waves in one slot, so the attainable rate for such
# code lies above this figure (at most twice it).
VALU_PEAK_LANE_INSTR = 1024 * 16 * 2.4e9


def measured_valu(kernel: str, x: int, lattice: str, chains: int):
    """From the committed counter summary of this workload: VALU instructions
    per site-update of `kernel` (SQ_INSTS_VALU per wave lane and tau slice),
    the share issued in pairs and the profiled VALU busy share; (None, None)
    when no summary matches."""
    path = os.path.join(ROOT, "profiles", f"valu_{kernel}_{lattice}_fp32.md")
    if chains != 128 or not os.path.exists(path):
        return None, None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 9 and cells[0].startswith(f"`k_{kernel}<float, {x},") and cells[3] not in ("", "-"):
            pct = lambda c: round(float(c.rstrip("%")) / 100, 2)
            return {"instr": float(cells[3]), "pairs": pct(cells[7]), "busy": pct(cells[8])}, \
                os.path.relpath(path, ROOT)
    return None, None


def run_batched(args, log, fp64=False, fp64_chains=16, steps=None, warmup=None):
    import torch
    from cl_var_dt_amd import CLV_FP32, CLV_FP64, Shape
    from cl_var_dt_amd.engine import Engine

    if fp64:
        X, Y, Nt, nch = args.fp64_lx, args.fp64_lx, args.fp64_nt, fp64_chains
    else:
        X, Y, Nt, nch = args.lx, args.lx, args.nt, args.chains
    shapes = [Shape(X, Y, 1, Nt)] * nch
    eng = Engine(shapes, precision=CLV_FP64 if fp64 else CLV_FP32)
    mus = [0.5 + i / max(nch - 1, 1) for i in range(nch)]
    for i in range(nch):
        eng.set_params(i, _params(1000 + i, mu=mus[i], nt=Nt))
        eng.init_fields(i, 1.0, 0.1)
    tuning = None
    if args.schedule:
        eng.set_schedule(args.schedule)
    else:
        import tempfile
        fd, path = tempfile.mkstemp(suffix=".jsonl")
        os.close(fd)
        eng.autotune(path)
        with open(path) as f:
            lines = [json.loads(ln) for ln in f if ln.strip()]
        os.unlink(path)
        tuning = lines[0] if lines else None
    log(f"batched: schedule {eng.schedule(0)}, {nch} chains of {X}x{Y}x1x{Nt} {'FP64' if fp64 else 'FP32'}, "
        f"device bytes {eng.device_bytes()/1e9:.2f} GB")
    steps = args.steps if steps is None else steps
    eng.propagate(args.warmup if warmup is None else warmup, 10)
    eng.prepare_graph(10)  # the graph capture (one-time set-up) stays out of the timed region
    eng.synchronize()
    torch.cuda.synchronize()
    # timed region: the production path (graph-replayed steps, no kernel events)
    t0 = time.perf_counter()
    eng.propagate(steps, 10)
    eng.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sites = X * Y * Nt * nch
    wall = t1 - t0
    # a separate pass of the same steps with HIP events around every kernel on
    # the context's stream (events force eager launches), for the per-kernel times
    eng.set_profiling(True)
    t2 = time.perf_counter()
    eng.propagate(args.prof_steps, 10)
    eng.synchronize()
    prof_wall = time.perf_counter() - t2
    kt = {k: eng.kernel_time(k) for k in ("drift_stats", "controller", "update", "obs", "obs_finish", "fused")}
    eng.set_profiling(False)
    st = eng.stats(0)
    bad = sum(eng.stats(i).nonfinite for i in range(nch))
    sched = eng.schedule(0)
    eng.close()
    return dict(sites=sites, wall=wall, su=sites * steps / wall, steps=steps, kt=kt, dt0=st.dt, nonfinite=bad,
                schedule=sched, tuning=tuning, prof_su=sites * args.prof_steps / prof_wall)


def run_plain(args, log):
    import torch
    from cl_var_dt_amd import CLV_FP64, Shape
    from cl_var_dt_amd.engine import Engine

    shape = Shape(64, 64, 1, 64)
    eng = Engine([shape], precision=CLV_FP64)
    eng.set_params(0, _params(7, nt=64))
    eng.init_fields(0, 1.0, 0.1)
    eng.propagate_plain(0, args.plain_warmup, 10)
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.propagate_plain(0, args.plain_steps, 10)
    eng.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    eng.close()
    sites = 64 * 64 * 64
    log(f"plain FP64 64x64x1x64: {args.plain_steps} steps in {wall:.3f} s")
    return dict(su=sites * args.plain_steps / wall, ms_per_step=1e3 * wall / args.plain_steps)


def run_uneven(args, log):
    """256 chains of 3-D lattices L^3 x Nt, L in {16, 24, 32}, Nt in {16, 32, 64}
    drawn from a seed, beta in [0.5, 2], FP32 fields, graph-captured steps."""
    import random

    import torch
    from cl_var_dt_amd import CLV_FP32, Shape
    from cl_var_dt_amd.engine import Engine

    rng = random.Random(3)
    shapes, betas = [], []
    for _ in range(args.uneven_chains):
        L = rng.choice([16, 24, 32])
        nt = rng.choice([16, 32, 64])
        shapes.append(Shape(L, L, L, nt))
        betas.append(rng.uniform(0.5, 2.0))
    eng = Engine(shapes, precision=CLV_FP32)
    for i, (s, beta) in enumerate(zip(shapes, betas)):
        p = _params(5000 + i, mu=1.0, nt=s.nt)
        p.eps = beta / s.nt
        eng.set_params(i, p)
        eng.init_fields(i, 1.0, 0.1)
    eng.autotune(args.tune_cache)
    eng.propagate(args.warmup, 10)
    eng.prepare_graph(10)  # the graph capture (one-time set-up) stays out of the timed region
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.propagate(args.uneven_steps, 10)
    eng.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sites = sum(s.sites for s in shapes)
    sched = sorted({eng.schedule(i) for i in range(len(shapes))})
    eng.close()
    log(f"uneven: {len(shapes)} chains, {sites} sites/step, {args.uneven_steps} steps in {wall:.3f} s")
    return {"value": round(sites * args.uneven_steps / wall, 1), "unit": "site-updates/s",
            "chains": len(shapes), "sites_per_step": sites, "steps": args.uneven_steps,
            "ms_per_step": round(1e3 * wall / args.uneven_steps, 4), "schedules": sched,
            "workload": "256 chains of 3-D L^3 x Nt, L in {16,24,32}, Nt in {16,32,64} (seeded), "
                        "beta in [0.5,2], FP32 fields, FP64 reductions, hipGraph-captured steps"}


def run_serving(args, log):
    """Poisson stream of Langevin runs (2-D L^2 x Nt, L in {64,128,256}, Nt in
    {32,64,128}, mu/c0/c1/q drawn from a seed) at ~70% of measured capacity,
    plus secant mu searches issuing follow-up runs toward density 1.0."""
    import random
    import threading

    import numpy as np
    from cl_var_dt_amd import CLV_FP32, Shape, default_params
    from cl_var_dt_amd import serving

    rng = random.Random(11)

    def draw():
        L = rng.choice([64, 128, 256])
        nt = rng.choice([32, 64, 128])
        p = default_params(mu=rng.uniform(0.5, 1.5), G=rng.uniform(0.05, 0.15),
                           c1=rng.uniform(-0.05, 0.05), Q=rng.uniform(0.0, 0.2), P=0.0,
                           eps=1.0 / nt, DT=1e-3, seed=rng.getrandbits(48))
        return Shape(L, L, 1, nt), p

    nsteps = args.serve_steps
    with serving.Server(precision=CLV_FP32) as srv:
        # a warm-up burst creates the pools of the shapes drawn (a one-time
        # cost per shape: device allocations and set-up); the capacity is then
        # the steady rate of a closed burst on the warm server
        warm = [draw() for _ in range(args.serve_burst)]
        for i in [srv.submit(s, p, nsteps=nsteps, eval_every=10) for s, p in warm]:
            srv.wait(i, release=True)
        # capacity: a closed burst
        burst = [draw() for _ in range(args.serve_burst)]
        t0 = time.perf_counter()
        ids = [srv.submit(s, p, nsteps=nsteps, eval_every=10) for s, p in burst]
        for i in ids:
            srv.wait(i)
        cap_wall = time.perf_counter() - t0
        su_burst = sum(s.sites for s, _ in burst) * nsteps
        capacity_su = su_burst / cap_wall
        mean_su = su_burst / len(burst)
        rate = args.serve_load * capacity_su / mean_su  # requests per second
        log(f"serving: capacity {capacity_su:.3e} SU/s, Poisson rate {rate:.2f} req/s")

        # secant clients on a serving shape, issuing follow-up runs while the
        # stream runs: each mu is a batch of independent runs (n0 = 0.25 start,
        # DT = 1e-2, densities averaged over the second half of each run) whose
        # mean density and standard error feed the reference's secant loop
        # (stop when the residual is inside the error)
        secant_out = []

        def secant_client(k):
            shape = Shape(args.secant_l, args.secant_l, 1, 32)

            def run_many(mus):
                ids = {}
                for m in mus:
                    ids[m] = [srv.submit(shape, default_params(
                        mu=m, G=0.1, c1=-0.05, Q=0.1, eps=1.0 / 32, DT=1e-2,
                        seed=10_000 * k + 100 * len(secant_calls) + j),
                        nsteps=args.secant_steps, eval_every=10, eval_from=args.secant_steps // 2,
                        n0=0.25, var=0.05)
                        for j in range(args.secant_runs)]
                secant_calls.append(list(mus))
                out = []
                for m in mus:
                    d = np.array([srv.wait(i, release=True).density for i in ids[m]])
                    out.append((float(d.mean()), float(d.std(ddof=1) / np.sqrt(len(d)))))
                return out

            secant_calls = []
            mu, hist, ok = serving.secant_method(run_many, args.secant_guess, 0.5, target=1.0,
                                                 max_i=8)
            secant_out.append({"mu": mu, "converged": ok, "iterations": len(hist),
                               "density": hist[-1][1], "err": hist[-1][2],
                               "runs": args.secant_runs * len(hist),
                               "lattice": f"{args.secant_l}x{args.secant_l}x1x32"})

        clients = [threading.Thread(target=secant_client, args=(k,)) for k in range(2)]
        for c in clients:
            c.start()
        # open-loop Poisson arrivals on the server's clock; latency counts from
        # the intended arrival time, so time a client spends in submit counts
        stream = []
        t_start = srv.clock()
        t_next = t_start
        while srv.clock() - t_start < args.serve_seconds:
            t_next += rng.expovariate(rate)
            while srv.clock() < t_next:
                time.sleep(min(0.001, max(0.0, t_next - srv.clock())))
            s, p = draw()
            stream.append((s, t_next, srv.submit(s, p, nsteps=nsteps, eval_every=10)))
        results = [(s, t_arr, srv.wait(i)) for s, t_arr, i in stream]
        for c in clients:
            c.join()
        wall = srv.clock() - t_start
        info = srv.info()
    lat = np.array([r.raw.t_done - t_arr for _, t_arr, r in results])
    lag = np.array([r.raw.t_submit - t_arr for _, t_arr, r in results])
    adm = np.array([r.raw.t_admit - r.raw.t_submit for _, _, r in results])
    su = sum(s.sites for s, _, _ in results) * nsteps
    return {"throughput": round(su / wall, 1), "unit": "site-updates/s",
            "p50_s": round(float(np.percentile(lat, 50)), 4),
            "p99_s": round(float(np.percentile(lat, 99)), 4),
            "latency_from": "intended Poisson arrival time to result on the host",
            "admission_wait_p50_s": round(float(np.percentile(adm, 50)), 4),
            "admission_wait_p99_s": round(float(np.percentile(adm, 99)), 4),
            "submit_lag_max_s": round(float(lag.max()), 5),
            "requests": len(results), "rate_req_per_s": round(rate, 3),
            "load_fraction": args.serve_load, "capacity_su_per_s": round(capacity_su, 1),
            "capacity_from": f"closed burst of {args.serve_burst} runs after a warm-up burst of as many",
            "steps_per_request": nsteps, "pools": int(info.pools),
            "pool_setup_s": round(info.pool_setup_seconds, 3),
            "hbm_bytes_in_use": int(info.bytes_in_use), "secant": secant_out,
            "workload": "Poisson stream of 2-D runs, L in {64,128,256}, Nt in {32,64,128}, "
                        "mu/c0/c1/q from a seed, FP32, HBM-budget admission, continuous batching"}


def run_cpu_baseline(args, log):
    """FP64 oracle (C++/OpenMP), time-bounded sample of the batched workload:
    its chain 0 (same lattice, parameters and seed), observables every 10."""
    from cl_var_dt_amd import Shape
    from oracle import pyoracle

    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        ncores = os.cpu_count() or 1
    X, Nt = args.lx, args.nt
    shape = Shape(X, X, 1, Nt)
    orc = pyoracle.Oracle(shape, _params(1000, mu=0.5, nt=Nt), nthreads=ncores)
    orc.init_fields(1.0, 0.1)
    orc.propagate(2, 0)
    steps, t0 = 0, time.perf_counter()
    while True:
        orc.propagate(1, 10)
        steps += 1
        el = time.perf_counter() - t0
        if el >= args.cpu_seconds or steps >= 2000:
            break
    orc.close()
    sites = X * X * Nt
    log(f"cpu baseline: {steps} steps in {el:.2f} s on {ncores} threads")
    return {"value": sites * steps / el, "unit": "site-updates/s", "cores": ncores, "kind": "port",
            "sample": f"FP64 oracle (project Stockham FFT, OpenMP), chain 0 of the batched workload "
                      f"(2-D {X}x{X}x1x{Nt}, mu=0.5), {steps} adaptive steps with observables every 10, "
                      f"{el:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=300)
    ap.add_argument("--plain-warmup", type=int, default=200,
                    help="untimed steps of the plain FP64 leg before its timed window")
    ap.add_argument("--chains", type=int, default=128)
    ap.add_argument("--lx", type=int, default=128)
    ap.add_argument("--nt", type=int, default=32)
    ap.add_argument("--schedule", type=int, default=0)
    ap.add_argument("--fp64-chains", default="16,32",
                    help="FP64 batched leg: comma-separated chain counts (0: skip the leg)")
    ap.add_argument("--fp64-lx", type=int, default=64)
    ap.add_argument("--fp64-nt", type=int, default=64)
    ap.add_argument("--fp64-warmup", type=int, default=400,
                    help="FP64 batched legs: at least this many untimed steps before the timed window")
    ap.add_argument("--fp64-steps", type=int, default=300,
                    help="FP64 batched legs: at least this many timed steps")
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32",
                    help="--mode batched: precision of the batched leg (fp64 uses --fp64-*)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--mode", choices=["all", "batched", "uneven", "copy"], default="all",
                    help="batched / uneven: only that leg (for profiling runs); copy: a 1 GiB "
                         "device-to-device copy, the calibration kernel of the traffic counters")
    ap.add_argument("--uneven-chains", type=int, default=256)
    ap.add_argument("--uneven-steps", type=int, default=20)
    ap.add_argument("--serve-steps", type=int, default=100)
    ap.add_argument("--serve-burst", type=int, default=48)
    ap.add_argument("--serve-seconds", type=float, default=15.0)
    ap.add_argument("--serve-load", type=float, default=0.7)
    ap.add_argument("--secant-steps", type=int, default=4000)
    ap.add_argument("--secant-runs", type=int, default=4, help="independent runs per secant mu")
    ap.add_argument("--secant-l", type=int, default=64, help="secant client lattice L (L x L x 1 x 32)")
    ap.add_argument("--secant-guess", type=float, default=0.1)
    ap.add_argument("--prof-steps", type=int, default=10,
                    help="steps of the separate pass with per-kernel HIP events")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--tune-cache", default=None, metavar="PATH",
                    help="JSON-lines autotuner cache for the uneven batch (default: tune without a cache)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: chains are independent, so a multi-GPU run is "
              "N copies of this one (no collectives); run with --gpus 1.", file=sys.stderr)
        return 2

    def log(msg):
        if not args.quiet:
            print(msg, file=sys.stderr, flush=True)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1

    if args.mode == "copy":
        x = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4, "copies": 3}), flush=True)
        return 0
    if args.mode == "uneven":
        print(json.dumps(run_uneven(args, log)), flush=True)
        return 0
    fp64_counts = [int(c) for c in str(args.fp64_chains).split(",") if int(c) > 0]
    b = run_batched(args, log, fp64=(args.mode == "batched" and args.precision == "fp64"),
                    fp64_chains=fp64_counts[0] if fp64_counts else 16)
    if args.mode == "batched":
        print(json.dumps({"batched_su_per_s": b["su"], "kernel_ms": b["kt"]}), flush=True)
        return 0
    # the FP64 batched path: chains of the plain path's lattice advanced together
    # (autotuned schedule), against 256 algorithmic bytes per site-update.  Its
    # steps are short (about 0.3 ms at 16 chains) and the GPU's clocks take
    # about 100 ms of steady load after the autotuner to settle (the first 30
    # steps run about 12% slower than the 300th), so it warms up and times
    # windows of at least --fp64-warmup and --fp64-steps steps.
    b64 = []
    for nc in fp64_counts:
        r = run_batched(args, log, fp64=True, fp64_chains=nc, steps=max(args.steps, args.fp64_steps),
                        warmup=max(args.warmup, args.fp64_warmup))
        step_s = r["wall"] / r["steps"]
        t = r["tuning"] or {}
        b64.append({"value": round(r["su"], 1), "unit": "site-updates/s", "chains": nc,
                    "lattice": f"{args.fp64_lx}x{args.fp64_lx}x1x{args.fp64_nt}", "schedule": r["schedule"],
                    "tau_blocks": t.get("fblocks"), "ms_per_step": round(1e3 * step_s, 4),
                    "steps": r["steps"], "warmup": max(args.warmup, args.fp64_warmup),
                    "hbm_frac": round(256 * r["sites"] / step_s / 1e9 / HBM_PEAK_GBS, 4),
                    "workload": "FP64 fields and reductions, per-chain adaptive dt, observables every 10 steps, "
                                "autotuned, graph-replayed steps; hbm_frac = 256 algorithmic bytes per "
                                "site-update / step time / 8 TB/s"})
    p = run_plain(args, log)
    # roofline of the dominant kernel: algorithmic bytes per launch / average launch time
    esz = 8  # complex64
    alg_per_site = {"update": 4 * 4 * esz, "fused": 4 * 4 * esz, "drift_stats": 2 * 4 * esz,
                    "obs": 2 * 4 * esz}
    dom = max((k for k in b["kt"] if b["kt"][k][1] > 0), key=lambda k: b["kt"][k][0])
    ms, n = b["kt"][dom]
    avg_s = ms / 1e3 / n
    alg_bytes = alg_per_site.get(dom, 0) * b["sites"]
    achieved = alg_bytes / avg_s / 1e9 if alg_bytes else 0.0
    traffic, tsrc = measured_traffic(dom, b["schedule"], f"{args.lx}x{args.lx}x1x{args.nt}", args.chains)
    step_ms = 1e3 * b["wall"] / args.steps
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
            "avg_launch_ms": round(ms / n, 4), "alg_bytes_per_launch": alg_bytes,
            "traffic_source": tsrc, "schedule": b["schedule"],
            "launch_times": f"HIP events on the context's stream, a separate pass of {args.prof_steps} "
                            "eager steps right after the timed (graph-replayed) region",
            # the whole production step (observables included) against the same bytes
            "step_achieved": round(alg_per_site["fused"] * b["sites"] / (step_ms / 1e3) / 1e9, 1),
            "step_frac": round(alg_per_site["fused"] * b["sites"] / (step_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    # the issue-rate bound the fused step actually meets: its VALU instructions
    # per site-update (counter summary) at the live launch time
    vm, vsrc = measured_valu(dom, args.lx, f"{args.lx}x{args.lx}x1x{args.nt}", args.chains)
    valu = None
    if vm:
        ach = vm["instr"] * b["sites"] / avg_s
        valu = {"instr_per_site_update": vm["instr"], "achieved": round(ach / 1e12, 2),
                "single_issue_peak": round(VALU_PEAK_LANE_INSTR / 1e12, 2), "unit": "T lane-instr/s",
                "frac_of_single_issue": round(ach / VALU_PEAK_LANE_INSTR, 4),
                "issued_in_pairs_profiled": vm["pairs"], "busy_profiled": vm["busy"],
                "kernel": dom, "source": vsrc}
    unev = None if args.no_uneven else run_uneven(args, log)
    serv = None if args.no_serving else run_serving(args, log)
    cpu = None if args.no_cpu_baseline else run_cpu_baseline(args, log)
    out = {
        "metric": METRIC,
        "value": round(b["su"], 1),
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched FP32: {args.chains} chains of 2-D {args.lx}x{args.lx} x Nt={args.nt}, "
                        "mu on a grid in [0.5, 1.5], c0=0.1, c1=-0.05, q=0.1, per-chain adaptive dt, "
                        "observables every 10 steps",
            "chains": args.chains, "lattice": f"{args.lx}x{args.lx}x1x{args.nt}",
            "sites_per_step": b["sites"], "parallelism": "chains on one GPU",
        },
        "fp64_plain": {"value": round(p["su"], 1), "unit": "site-updates/s",
                       "ms_per_step": round(p["ms_per_step"], 4),
                       "steps": args.plain_steps, "warmup": args.plain_warmup,
                       "workload": "one 2-D 64x64 x Nt=64 chain, FP64, host-driven dt per step"},
        "batched_over_plain": round(b["su"] / p["su"], 2),
        "fp64_batched": b64 or None,
        "kernel_ms": {k: round(v[0], 3) for k, v in b["kt"].items() if v[1]},
        "events_pass_su_per_s": round(b["prof_su"], 1),
        "roofline": roof,
        "valu_issue": valu,
        "autotune": b["tuning"],
        "uneven_batch": unev,
        "serving": serv,
        "p50_s": serv["p50_s"] if serv else None,
        "p99_s": serv["p99_s"] if serv else None,
        "cpu_baseline": cpu,
        "nonfinite_chains": b["nonfinite"],
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
