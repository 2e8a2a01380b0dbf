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


def run_batched(args, log):
    import torch
    from cl_var_dt_amd import CLV_FP32, Shape
    from cl_var_dt_amd.engine import Engine

    X, Y, Nt, nch = args.lx, args.lx, args.nt, args.chains
    shapes = [Shape(X, Y, 1, Nt)] * nch
    eng = Engine(shapes, precision=CLV_FP32)
    mus = [0.5 + i / max(nch - 1, 1) for i in range(nch)]
    for i in range(nch):
        eng.set_params(i, _params(1000 + i, mu=mus[i], nt=Nt))
        eng.init_fields(i, 1.0, 0.1)
    if args.schedule:
        eng.set_schedule(args.schedule)
    log(f"batched: {nch} chains of {X}x{Y}x1x{Nt} FP32, device bytes {eng.device_bytes()/1e9:.2f} GB")
    eng.propagate(args.warmup, 10)
    eng.synchronize()
    torch.cuda.synchronize()
    eng.set_profiling(True)
    t0 = time.perf_counter()
    eng.propagate(args.steps, 10)
    eng.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sites = X * Y * Nt * nch
    wall = t1 - t0
    kt = {k: eng.kernel_time(k) for k in ("drift_stats", "controller", "update", "obs", "obs_finish", "fused")}
    eng.set_profiling(False)
    st = eng.stats(0)
    bad = sum(eng.stats(i).nonfinite for i in range(nch))
    eng.close()
    return dict(sites=sites, wall=wall, su=sites * args.steps / wall, kt=kt, dt0=st.dt, nonfinite=bad)


def run_plain(args, log):
    import torch
    from cl_var_dt_amd import CLV_FP64, Shape
    from cl_var_dt_amd.engine import Engine

    shape = Shape(64, 64, 1, 64)
    eng = Engine([shape], precision=CLV_FP64)
    eng.set_params(0, _params(7, nt=64))
    eng.init_fields(0, 1.0, 0.1)
    eng.propagate_plain(0, 10, 10)
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


def run_cpu_baseline(args, log):
    """FP64 oracle (C++/OpenMP) on the plain workload's shape, time-bounded sample."""
    from cl_var_dt_amd import Shape
    from oracle import pyoracle

    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        ncores = os.cpu_count() or 1
    shape = Shape(64, 64, 1, 64)
    orc = pyoracle.Oracle(shape, _params(7, nt=64), nthreads=ncores)
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
    sites = 64 * 64 * 64
    log(f"cpu baseline: {steps} steps in {el:.2f} s on {ncores} threads")
    return {"value": sites * steps / el, "unit": "site-updates/s", "cores": ncores, "kind": "port",
            "sample": f"FP64 oracle (project Stockham FFT, OpenMP), 2-D 64x64x1x64 chain, "
                      f"{steps} adaptive steps with observables every 10, {el:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--plain-steps", type=int, default=300)
    ap.add_argument("--chains", type=int, default=128)
    ap.add_argument("--lx", type=int, default=128)
    ap.add_argument("--nt", type=int, default=32)
    ap.add_argument("--schedule", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--mode", choices=["all", "batched"], default="all",
                    help="batched: only the batched leg (for profiling runs)")
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

    b = run_batched(args, log)
    if args.mode == "batched":
        print(json.dumps({"batched_su_per_s": b["su"], "kernel_ms": b["kt"]}), flush=True)
        return 0
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
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
            "avg_launch_ms": round(ms / n, 4),
            "alg_bytes_per_launch": alg_bytes}
    cpu = None if args.no_cpu_baseline else run_cpu_baseline(args, log)
    step_ms = 1e3 * b["wall"] / args.steps
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
                       "workload": "one 2-D 64x64 x Nt=64 chain, FP64, host-driven dt per step"},
        "batched_over_plain": round(b["su"] / p["su"], 2),
        "kernel_ms": {k: round(v[0], 3) for k, v in b["kt"].items() if v[1]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "nonfinite_chains": b["nonfinite"],
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
