"""Benchmark: energy steps/s of the DLRA CSD transport step on one MI355X.

Workload (the configuration the headline metric is quoted on): the 200x200
heterogeneous ellipse phantom over 10x10 cm (seed 0: 1.0 g/cc background,
0.25 g/cc lung and 1.85 g/cc bone ellipses), 15 MeV electron pencil beam,
P_21 (253 moments), rank-adaptive augmented BUG with truncation tolerance 1e-2
(relative) and rank cap 40, energy grid by the reference's rule (823 points).
The plain FP64 form runs here.

A step is one iteration of the energy loop (boundary source, dose, uncollided
sweep, streaming K/L/S, truncation, out- and in-scattering, dose).  The
context (all operands in HBM) is created before the timed region; W warmup
steps run first, then K steps are timed between stream synchronisations.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds T]

Prints one JSON line.  `roofline` is computed from the library's own HIP-event
timing of its dominant kernel during the timed region and the algorithmic
bytes/flops of those launches; `cpu_baseline` times the FP64 oracle (numpy,
all BLAS threads) on a bounded prefix of the same solve.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "energy steps/s & beamlets/s, mixed vs FP64 GPU at dose rel-L2 error ≤1e-4"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
RIDGE = FP64_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9)

# HBM bytes per launch of the dominant kernel measured with rocprofv3 --pmc
# (FETCH_SIZE + WRITE_SIZE, scaled by a calibration copy); None until measured.
MEASURED_TRAFFIC = {}


def build_workload(n=200, N=21, cap=40, seed=0):
    from titus_amd.problem import INTEGRATOR_BUG, build_problem
    from titus_amd.settings import heterogeneous_phantom
    s = heterogeneous_phantom(n=n, N=N, seed=seed, r=cap)
    return build_problem(s, "subset", INTEGRATOR_BUG)


def cpu_baseline(P, seconds):
    """Time the FP64 oracle on the first steps of the same solve (bounded)."""
    from oracle import dlra
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    steps, elapsed = 0, 0.0
    k = 4
    while elapsed < seconds and steps < P["nE"] - 2:
        k = min(k, P["nE"] - 2)
        t0 = time.perf_counter()
        dlra.solve_bug(P, steps=k)
        elapsed = time.perf_counter() - t0
        steps = k
        if elapsed < seconds:
            k = min(P["nE"] - 2, max(k + 1, int(k * max(1.5, 0.9 * seconds / max(elapsed, 1e-3)))))
    return {"value": steps / elapsed, "unit": "energy steps/s", "cores": int(threads),
            "kind": "port",
            "sample": f"FP64 numpy oracle, steps 1..{steps} of the same solve, {elapsed:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(beamlet sharding across GPUs is a design note, not code)", file=sys.stderr)
        sys.exit(2)

    from titus_amd import _lib
    P = build_workload()
    last = P["nE"] - 2
    W = min(args.warmup, last - 1)
    K = min(args.steps, last - W)
    ctx = _lib.Context(P, _lib.INTEGRATOR_BUG)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None
    ctx.solve(W)
    sync()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    ctx.solve(W + K)
    sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    stats = {name: ctx.kernel_stats(name) for name in ctx.KERNELS}
    done, rank = ctx.progress()
    ranks, _ = ctx.history(done)
    ctx.close()

    dev_ms = sum(v[0] for v in stats.values())
    dom = max(stats, key=lambda k: stats[k][0])
    ms, launches, by, fl = stats[dom]
    t_avg = ms / max(launches, 1) * 1e-3
    intensity = fl / by if by > 0 else 0.0
    if intensity > RIDGE:
        roof = {"bound": "mfma", "achieved": fl / max(launches, 1) / t_avg / 1e12,
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": by / max(launches, 1) / t_avg / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = MEASURED_TRAFFIC.get(dom)
    roof["kernel"] = dom
    roof["avg_launch_us"] = t_avg * 1e6
    roof["launches"] = launches

    out = {
        "metric": METRIC,
        "value": K / elapsed,
        "unit": "energy steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded ellipse phantom, analytic beam source)",
        "config": {
            "workload": "2D heterogeneous phantom 200x200 over 10x10 cm, 15 MeV e-, P_21 "
                        "(253 moments), rank-adaptive BUG theta=1e-2, rank cap 40, "
                        f"{P['nE']}-point energy grid, plain FP64",
            "cells": P["C"], "moments": P["Nm"], "ordinates": P["Q"],
            "steps_timed": f"{W + 1}..{W + K}", "rank_range": [int(ranks[W:].min()), int(ranks[W:].max())],
            "parallelism": "single GPU",
        },
        "device_ms_per_step": dev_ms / K,
        "kernel_ms": {k: round(v[0], 3) for k, v in stats.items()},
        "roofline": roof,
        "cpu_baseline": None if args.no_cpu else cpu_baseline(P, args.cpu_seconds),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
