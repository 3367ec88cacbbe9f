"""Benchmark: energy steps/s of the DLRA CSD transport step on one MI355X.

Workload (the configuration the headline metric is quoted on): the 200x200
heterogeneous ellipse phantom over 10x10 cm (seed 0: 1.0 g/cc background,
0.25 g/cc lung and 1.85 g/cc bone ellipses), 15 MeV electron pencil beam,
P_21 (253 moments), rank-adaptive augmented BUG with truncation tolerance 1e-2
(relative) and rank cap 40, energy grid by the reference's rule (823 points).

Both forms run on the same workload: the plain FP64 form and the mixed form
(FP32 psi storage and uncollided sweep; FP64 streaming stencil, Galerkin and
psi products, Cholesky-QR, truncation SVD and dose sum).  The form with the
streaming stencil, psi and Galerkin products in FP32 as well is run untimed
beside them for its error.  `value` is the mixed form's rate;
the plain form's rate, their ratio and the dose difference after the timed
steps are reported beside it.

A step is one iteration of the energy loop (boundary source, dose, uncollided
sweep, streaming K/L/S, truncation, out- and in-scattering, dose).  Each
context (all operands in HBM) is created before its timed region.  Steps
1..A (A = --advance, 300) run untimed first, so that the timed window sits
at the rank the solve keeps for most of its grid (ranks grow from 1-4 over
the first steps); then W warmup steps, then K steps are timed between stream
synchronisations.  The rate of each whole solve (every step of the grid) is
reported beside the window's.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--advance A] [--cpu-seconds T]

Prints one JSON line.  `roofline` is computed from the library's own HIP-event
timing of the dominant kernel during the timed region and the algorithmic
bytes/flops of those launches; `cpu_baseline` times the FP64 oracle (numpy,
all BLAS threads) on a bounded prefix of the same solve.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "energy steps/s & beamlets/s, mixed vs FP64 GPU at dose rel-L2 error ≤1e-4"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}

# HBM bytes per call of the headline's dominant kernel over the timed window,
# measured with two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) of the
# mixed form's headline solve (`bench.py --profile-form mixed`, the same
# steps and window as the timed run), scaled by
# what the counters read for a 256 MiB device copy (x2 fetch, x1 write).  The
# per-dispatch figures, the command and the window are in
# profiles/dlra_step_phantom200_p21_pmc.md; the entry names them, and the
# bench line repeats them beside the number.
MEASURED_TRAFFIC = {"gram": {"bytes": 5.89e6, "kernels": "TnBody<1,...> + TnReduceBody",
                             "window": "steps 306..325, mixed form",
                             "command": "rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) -- "
                                        "python3 bench.py --profile-form mixed --steps 20 --warmup 5",
                             "profile": "profiles/dlra_step_phantom200_p21_pmc.md"}}
# HBM bytes per launch of the full-rank solver's collided kernel on the
# phantom (FETCH_SIZE x 2 + WRITE_SIZE x 1, profiles/fcs_collide_phantom200_p21.md)
MEASURED_TRAFFIC_FULL_PN = {"collide": 2.698e8}


def build_workload(n=200, N=21, cap=40, seed=0):
    from titus_amd.problem import INTEGRATOR_BUG, build_problem
    from titus_amd.settings import heterogeneous_phantom
    s = heterogeneous_phantom(n=n, N=N, seed=seed, r=cap)
    return build_problem(s, "subset", INTEGRATOR_BUG)


def cpu_info():
    """Host CPU: model name, logical CPUs, and the BLAS threads numpy uses."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    return model, os.cpu_count() or 1, int(threads)


def cpu_baseline(P, seconds, start):
    """The FP64 oracle (numpy, all BLAS threads) on the same solve: steps
    1..start untimed, then the steps from start + 1 that fit in `seconds`,
    timed.  Returns the JSON record and (last step, oracle dose there)."""
    from oracle import dlra
    model, nproc, threads = cpu_info()
    marks = {"t0": None}

    def on_step(s, dose):
        now = time.perf_counter()
        if s == start:
            marks["t0"] = now
        elif s > start:
            marks["t1"], marks["s1"] = now, s
            if now - marks["t0"] >= seconds:
                return False
        return True

    t = time.perf_counter()
    out = dlra.solve_bug(P, steps=P["nE"] - 2, on_step=on_step)
    if marks["t0"] is None:
        marks["t0"], marks["t1"], marks["s1"] = t, time.perf_counter(), len(out["ranks"])
    n = marks["s1"] - start
    dt = marks["t1"] - marks["t0"]
    rec = {"value": n / dt, "unit": "energy steps/s", "cores": threads, "nproc": nproc,
           "cpu_model": model, "kind": "port",
           "sample": f"FP64 numpy oracle (a restatement, not the reference), steps {start + 1}.."
                     f"{marks['s1']} of the same solve after untimed steps 1..{start}, {dt:.1f} s"}
    return rec, marks["s1"], out["dose"]


def device_dose(P, integrator, precision, last, rank_tol):
    from titus_amd import _lib
    ctx = _lib.Context(P, integrator, precision=precision, rank_tol=rank_tol)
    try:
        ctx.solve(last)
        return ctx.state()["dose"]
    finally:
        ctx.close()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def whole_solve(P, integrator, precision, rank_tol, sync):
    """Rate of a whole solve (every step of the grid), context created first."""
    from titus_amd import _lib
    ctx = _lib.Context(P, integrator, precision=precision, rank_tol=rank_tol)
    try:
        sync()
        t0 = time.perf_counter()
        ctx.solve(0)
        sync()
        dt = time.perf_counter() - t0
        done, _ = ctx.progress()
        return {"energy_steps_per_s": done / dt, "steps": done, "seconds": dt}, ctx.state()["dose"]
    finally:
        ctx.close()


def run_form(P, precision, W, K, period, sync, rank_tol):
    """Solve steps 1..W untimed, then time steps W+1..W+K.

    `period` > 0: HIP events and host-phase clocks on every period-th step of
    the timed region (0: none); `sampled` counts those steps.
    """
    from titus_amd import _lib
    ctx = _lib.Context(P, _lib.INTEGRATOR_BUG, precision=precision, rank_tol=rank_tol)
    try:
        ctx.solve(W)
        sync()
        ctx.set_timing(period)
        t0 = time.perf_counter()
        ctx.solve(W + K)
        sync()
        elapsed = time.perf_counter() - t0
        stats = {name: ctx.kernel_stats(name) for name in ctx.KERNELS}
        host = {name: ctx.kernel_stats(name)[0] for name in ctx.HOST_PHASES}
        done, _ = ctx.progress()
        ranks, _ = ctx.history(done)
        dose = ctx.state()["dose"]
    finally:
        ctx.close()
    sampled = sum(1 for s in range(W + 1, W + K + 1) if period and s % period == 0)
    return dict(elapsed=elapsed, rate=K / elapsed, stats=stats, host=host, sampled=sampled,
                ranks=ranks, dose=dose)


def beamlet_batch(n_beamlets, nthreads, sync, profile_only=False, oracle_steps=40):
    """Config-4 plan: the plain FP64 GPU path (one beamlet after another)
    against the batched forms.

    Batched forms: shared launches (titus_solve_batch_lockstep: all beamlets
    in LOCKSTEP_GROUPS groups, every kernel one launch per group), mixed and
    FP64; and one stream per beamlet over `nthreads` host workers
    (titus_solve_batch), mixed.  Every set of contexts is created (operands
    in HBM) before its timed region.  Returns throughput of each, the ratios
    and per-beamlet dose differences (mixed vs FP64) of the same plan.
    """
    from titus_amd import _lib
    from titus_amd.beamlets import LOCKSTEP_GROUPS, beamlet_plan
    from titus_amd.solver import MIXED_RANK_TOL
    plan = beamlet_plan(n_beamlets=n_beamlets)
    total_steps = sum(P["nE"] - 2 for _, P in plan)

    def rate(t):
        return {"beamlets_per_s": n_beamlets / t, "energy_steps_per_s": total_steps / t, "seconds": t}

    def run(precision, mode):
        ctxs = [_lib.Context(P, _lib.INTEGRATOR_BUG, precision=precision, rank_tol=MIXED_RANK_TOL)
                for _, P in plan]
        stats = None
        try:
            sync()
            t0 = time.perf_counter()
            if mode == "lockstep":
                _, stats = _lib.solve_batch_lockstep(ctxs, 0, LOCKSTEP_GROUPS)
            elif mode == "streams":
                _lib.solve_batch(ctxs, nthreads)
            else:
                for c in ctxs:
                    c.solve(0)
            sync()
            dt = time.perf_counter() - t0
            doses = np.stack([c.state()["dose"] for c in ctxs], axis=1)
        finally:
            for c in ctxs:
                c.close()
        return dt, doses, stats

    if profile_only:
        tlmx, _, stmx = run(_lib.PRECISION_MIXED, "lockstep")
        return {"mixed_batched": dict(rate(tlmx), **stmx)}
    t64, d64, _ = run(_lib.PRECISION_FP64, "one_at_a_time")
    tl64, dl64, st64 = run(_lib.PRECISION_FP64, "lockstep")
    tlmx, dlmx, stmx = run(_lib.PRECISION_MIXED, "lockstep")
    tsmx, dsmx, _ = run(_lib.PRECISION_MIXED, "streams")
    # the form with the streaming stencil, psi and Galerkin products in FP32
    # as well (FP32 accumulation), batched the same way
    tlf32, dlf32, stf32 = run(_lib.PRECISION_MIXED_FULL_FP32, "lockstep")
    errs = np.linalg.norm(dlmx - d64, axis=0) / np.linalg.norm(d64, axis=0)
    errs32 = np.linalg.norm(dlf32 - d64, axis=0) / np.linalg.norm(d64, axis=0)
    w = np.ones(len(plan))
    plan_err = float(np.linalg.norm((dlmx - d64) @ w) / np.linalg.norm(d64 @ w))
    def shared(st):
        return {"groups": LOCKSTEP_GROUPS, "host_syncs": st["flushes"], "launches": st["launches"],
                "recorded_launches": st["ops"],
                "beamlets_per_launch": st["ops"] / max(st["launches"], 1),
                "launches_per_beamlet_step": st["launches"] / total_steps,
                "tuned_tiles_per_group": st["tiles"]}
    return {
        "workload": f"{n_beamlets} pencil beamlets on a 128x128 ellipse phantom (10x10 cm), "
                    "P_15 (136 moments), 6-20 MeV, +-15 deg, rank caps 5-40, seed 0",
        "beamlet_steps": int(total_steps),
        "mixed_batched": dict(rate(tlmx), form="shared launches", **shared(stmx)),
        "fp64_batched": dict(rate(tl64), form="shared launches", **shared(st64)),
        "mixed_one_stream_per_beamlet": dict(rate(tsmx), host_threads=nthreads),
        "mixed_full_fp32_batched": dict(
            rate(tlf32), form="shared launches",
            speedup_vs_fp64_one_at_a_time=t64 / tlf32,
            dose_rel_l2_per_beamlet={"max": float(errs32.max()), "median": float(np.median(errs32)),
                                     "above_1e-4": int((errs32 > 1e-4).sum())},
            plan_dose_rel_l2=float(np.linalg.norm((dlf32 - d64) @ np.ones(len(plan)))
                                   / np.linalg.norm(d64 @ np.ones(len(plan))))),
        "fp64_one_at_a_time": rate(t64),
        "speedup_mixed_batched_vs_fp64_one_at_a_time": t64 / tlmx,
        "speedup_from_batching_fp64": t64 / tl64,
        "speedup_from_precision_batched": tl64 / tlmx,
        "speedup_shared_launches_vs_streams_mixed": tsmx / tlmx,
        "batched_equals_one_at_a_time_fp64": bool(np.array_equal(d64, dl64)),
        "shared_launches_equal_streams_mixed": bool(np.array_equal(dlmx, dsmx)),
        "dose_rel_l2_per_beamlet": {"max": float(errs.max()), "median": float(np.median(errs)),
                                    "above_1e-4": int((errs > 1e-4).sum()),
                                    "worst": [[int(b), float(errs[b])] for b in np.argsort(errs)[::-1][:3]]},
        "plan_dose_rel_l2": plan_err,
        "vs_oracle": beamlet_vs_oracle(plan[0][1], oracle_steps),
        "note": "per-beamlet errors are mixed vs FP64 at the same rank floor; the FP64 solves "
                "match the oracle to ~5e-12 (GPU tests, two beamlets over 120 steps)",
    }


def beamlet_vs_oracle(P, steps):
    """Beamlet 0 of the plan over its first `steps` steps: FP64 and mixed
    device doses against the FP64 oracle's (the oracle at the plain floor;
    the mixed form at its own floor, and the oracle rerun at that floor)."""
    from oracle import dlra
    from titus_amd import _lib
    from titus_amd.problem import INTEGRATOR_BUG
    from titus_amd.solver import MIXED_RANK_TOL, RANK_TOL
    ref = dlra.solve_bug(P, steps=steps)["dose"]
    saved = dlra.RANK_TOL
    try:
        dlra.RANK_TOL = MIXED_RANK_TOL
        ref4 = dlra.solve_bug(P, steps=steps)["dose"]
    finally:
        dlra.RANK_TOL = saved
    return {"beamlet": 0, "steps": steps,
            "dose_rel_l2_fp64": rel(device_dose(P, INTEGRATOR_BUG, _lib.PRECISION_FP64, steps, RANK_TOL), ref),
            "dose_rel_l2_mixed_vs_oracle_same_floor": rel(
                device_dose(P, INTEGRATOR_BUG, _lib.PRECISION_MIXED, steps, MIXED_RANK_TOL), ref4),
            "dose_rel_l2_mixed_vs_oracle_plain_floor": rel(
                device_dose(P, INTEGRATOR_BUG, _lib.PRECISION_MIXED, steps, MIXED_RANK_TOL), ref)}


def slab_bench(sync, cpu):
    """The CPU-runnable configuration: 80x80 water slab (5x5 cm), 21 MeV
    electron pencil beam, P_9 (55 moments), fixed rank 20, FP64 and mixed
    whole solves, each dose against the FP64 oracle's."""
    from titus_amd import _lib
    from titus_amd.problem import INTEGRATOR_FIXED
    from titus_amd.solver import MIXED_RANK_TOL, RANK_TOL
    from titus_amd.problem import build_problem
    from titus_amd.settings import water_slab_pencil_beam
    P = build_problem(water_slab_pencil_beam(), "subset", INTEGRATOR_FIXED)
    out = {"workload": f"80x80 water slab 5x5 cm, 21 MeV e- pencil beam, P_9 ({P['Nm']} moments, "
                       f"{P['Q']} ordinates), fixed rank 20, {P['nE'] - 1} steps"}
    doses = {}
    for name, prec, tol in (("fp64", _lib.PRECISION_FP64, RANK_TOL),
                            ("mixed", _lib.PRECISION_MIXED, MIXED_RANK_TOL),
                            ("fp64_floor_1e-4", _lib.PRECISION_FP64, MIXED_RANK_TOL)):
        out[name], doses[name] = whole_solve(P, INTEGRATOR_FIXED, prec, tol, sync)
    # the precision's share (same rank floor) and the floor's share
    out["dose_rel_l2_mixed_vs_fp64_same_floor"] = rel(doses["mixed"], doses["fp64_floor_1e-4"])
    out["dose_rel_l2_fp64_floor_1e-4_vs_1e-7"] = rel(doses["fp64_floor_1e-4"], doses["fp64"])
    out["dose_rel_l2_mixed_vs_fp64"] = rel(doses["mixed"], doses["fp64"])
    if cpu:
        from oracle import dlra
        t0 = time.perf_counter()
        ref = dlra.solve_fixed(P)["dose"]
        out["cpu_oracle_energy_steps_per_s"] = (P["nE"] - 1) / (time.perf_counter() - t0)
        out["dose_rel_l2_fp64_vs_oracle"] = rel(doses["fp64"], ref)
        out["dose_rel_l2_mixed_vs_oracle"] = rel(doses["mixed"], ref)
    return out



def stencil_scale(sync, steps=4, n=1024, calibrate=False):
    """HBM efficiency of the sweep and streaming-stencil kernels at scale.

    A 1024x1024 homogeneous slab, P_5 with a broad beam (72 ordinates), fixed
    rank 16: the sweep moves 2 x 72 x 8 MB per launch (FP64), large enough to
    fill the chip.  Per kernel: achieved GB/s = algorithmic bytes / HIP-event
    time, and the fraction of the 8 TB/s peak.  `calibrate` also runs a plain
    256 MiB device copy (the reference kernel for the PMC traffic scaling).
    """
    from titus_amd import _lib
    from titus_amd.problem import INTEGRATOR_FIXED, build_problem
    from titus_amd.settings import Settings
    s = Settings(n + 1, n + 1, 16, "waterBeamElectrons", "Electrons", a=0.0, b=20.0, c=0.0,
                 d=20.0, x0=10.0, y0=0.0, nPN=5, eMin=1.0, eMax=10.5, sigmaX=2.0, sigmaY=2.0,
                 sigmaO1Inv=2.0, sigmaO3Inv=2.0)
    P = build_problem(s, "subset", INTEGRATOR_FIXED)
    out = {"workload": f"{n}x{n} slab, P_5, {P['Q']} ordinates, rank 16, {steps} steps"}
    for name, prec in (("fp64", _lib.PRECISION_FP64), ("mixed", _lib.PRECISION_MIXED)):
        ctx = _lib.Context(P, INTEGRATOR_FIXED, precision=prec)
        try:
            ctx.solve(2)
            sync()
            ctx.set_timing(True)
            ctx.solve(2 + steps)
            res = {}
            for k in ("sweep", "stream_K", "galerkin", "gemm"):
                ms, launches, by, fl = ctx.kernel_stats(k)
                if launches:
                    gbs = by / (ms * 1e-3) / 1e9
                    res[k] = {"GB/s": round(gbs, 1), "frac_of_8TB/s": round(gbs / HBM_PEAK_GBS, 3),
                              "avg_launch_us": round(ms / launches * 1e3, 2),
                              "bytes_per_launch": by / launches}
            res["sweep_plan"] = ctx.tuning("sweep")
            out[name] = res
        finally:
            ctx.close()
    # a plain 256 MiB device copy: the practical HBM ceiling on this GPU, and
    # the reference kernel for scaling the PMC traffic counters
    import torch
    if torch.cuda.is_available():
        a = torch.empty(256 * 2 ** 20 // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        reps = 1 if calibrate else 20
        for _ in range(reps):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        gbs = 2 * a.numel() * 8 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
        out["copy_256MiB"] = {"GB/s": round(gbs, 1), "frac_of_8TB/s": round(gbs / HBM_PEAK_GBS, 3)}
        if not calibrate:
            out["hbm_ceiling"] = hbm_ceiling(n + 1, n + 1, P["Q"], out)
    return out


def hbm_ceiling(nx, ny, nq, scale):
    """The measured ceilings (titus_amd/hbm_ceiling, a child process): best
    read, fill and copy of 1 GiB, and the best copy with the sweep's access
    pattern on the same psi planes; each sweep's rate as a share of them."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "titus_amd", "hbm_ceiling")
    if not os.path.exists(exe):
        return {"error": "titus_amd/hbm_ceiling is not built (python -m titus_amd.build)"}
    res = subprocess.run([exe, str(nx), str(ny), str(nq)], capture_output=True, text=True, timeout=120)
    if res.returncode != 0:
        return {"error": res.stderr.strip()[-300:]}
    raw = json.loads(res.stdout)
    best = lambda runs: max(runs, key=lambda e: e["GB/s"])
    out = {"read": best(raw["read"]), "fill": best(raw["fill"]), "copy": best(raw["copy"]),
           "copy_hipMemcpy": next(e for e in raw["copy"] if e["shape"] == "hipMemcpyAsync")}
    for form, prefix in (("fp64", "fp64"), ("mixed", "fp32")):
        strip = best([e for e in raw["strip_copy"]["runs"] if e["shape"].startswith(prefix)])
        out["sweep_pattern_copy_" + prefix] = strip
        sw = scale.get(form, {}).get("sweep")
        if sw:
            out["sweep_" + form + "_share"] = {
                "of_pattern_copy": round(sw["GB/s"] / strip["GB/s"], 3),
                "of_best_copy": round(sw["GB/s"] / out["copy"]["GB/s"], 3)}
    return out


def proton_cube_bench(sync, cpu_steps=2):
    """3D configuration: 90 MeV protons into a 64^3 grid of 0.1 cm voxels
    (seeded smooth density in [0.3, 1.6] g/cc), P_13 with all 196 moments,
    fixed rank 30, 300 energy steps, FP64 and mixed forms.  Whole solves are
    timed (contexts created before the timed region); kernels are timed with
    events on every 10th step of a second, identical FP64 solve.  The FP64
    oracle's per-step time on a few steps of the same problem is the CPU
    figure."""
    from titus_amd import _lib
    from titus_amd.problem import INTEGRATOR_FIXED
    from titus_amd.problem3d import build_problem3d, proton_cube
    from titus_amd.solver import MIXED_RANK_TOL
    P = build_problem3d(proton_cube())
    out = {"workload": f"3D proton beam 90 MeV, 64^3 x 0.1 cm, P_13 ({P['Nm']} moments, "
                       f"{P['Q']} ordinates), fixed rank 30, {P['nE'] - 1} steps, seed 0"}
    doses = {}
    for name, prec in (("fp64", _lib.PRECISION_FP64), ("mixed", _lib.PRECISION_MIXED)):
        ctx = _lib.Context(P, INTEGRATOR_FIXED, precision=prec, rank_tol=MIXED_RANK_TOL)
        try:
            sync()
            t0 = time.perf_counter()
            ctx.solve(0)
            sync()
            dt = time.perf_counter() - t0
            done, _ = ctx.progress()
            doses[name] = ctx.state()["dose"]
            out[name] = {"energy_steps_per_s": done / dt, "ms_per_step": dt / done * 1e3,
                         "sweep_plan": ctx.tuning("sweep")}
        finally:
            ctx.close()
    out["dose_rel_l2_mixed_vs_fp64"] = float(np.linalg.norm(doses["mixed"] - doses["fp64"])
                                             / np.linalg.norm(doses["fp64"]))
    ctx = _lib.Context(P, INTEGRATOR_FIXED, rank_tol=MIXED_RANK_TOL)
    try:
        ctx.set_timing(10)
        ctx.solve(0)
        sync()
        sampled = (P["nE"] - 1) // 10
        k = {}
        for name in ctx.KERNELS:
            ms, n, by, fl = ctx.kernel_stats(name)
            if n:
                k[name] = {"ms_per_step": round(ms / sampled, 4),
                           "GB/s": round(by / (ms * 1e-3) / 1e9, 1),
                           "TFLOP/s": round(fl / (ms * 1e-3) / 1e12, 2)}
        out["fp64_kernels"] = k
        out["fp64_host_ms_per_step"] = {n.split(":")[1]: round(ctx.kernel_stats(n)[0] / sampled, 4)
                                        for n in ctx.HOST_PHASES}
    finally:
        ctx.close()
    if cpu_steps:
        from oracle import dlra
        t0 = time.perf_counter()
        ref = dlra.solve_fixed(P, steps=cpu_steps)["dose"]
        out["cpu_oracle_s_per_step"] = (time.perf_counter() - t0) / cpu_steps
        out[f"dose_rel_l2_vs_oracle_after_{cpu_steps}_steps"] = {
            "fp64": rel(device_dose(P, INTEGRATOR_FIXED, _lib.PRECISION_FP64, cpu_steps, 1e-7), ref),
            "mixed": rel(device_dose(P, INTEGRATOR_FIXED, _lib.PRECISION_MIXED, cpu_steps,
                                     MIXED_RANK_TOL), ref)}
    return out


def full_pn_bench(sync, cpu_steps=2, warmup=10, steps=200):
    """The full-rank P_N solve (SolveFirstCollisionSource) on the headline
    phantom: 200x200 cells, P_21 (253 moments, 24 ordinates), FP64.  The
    context is created (and its block-sparse plan chosen) before the timed
    region; `steps` energy steps after `warmup` are timed between stream
    synchronisations, then the same steps of a second context with HIP events
    around every kernel.  The collided kernel's roofline: algorithmic flops =
    2 C (nnz(|Ax|) + nnz(|Az|) + nnz(Ax) + nnz(Az) + nm nq) per launch,
    algorithmic bytes = 8 (2 C nm + C nq); the matrix-core work executed is
    larger (whole 4 x 16 operator blocks) and reported beside it."""
    from titus_amd import _lib
    from titus_amd.problem import build_fcs_problem
    from titus_amd.settings import heterogeneous_phantom
    P = build_fcs_problem(heterogeneous_phantom(n=200, N=21, seed=0, r=2))
    last = warmup + steps
    out = {"workload": f"full-rank P_N, 2D phantom 200x200 (seed 0), P_21 ({P['Nm']} moments, "
                       f"{P['Q']} ordinates), FP64, steps {warmup + 1}..{last} of {P['nE'] - 1}"}
    ctx = _lib.FcsContext(P)
    try:
        ctx.solve(warmup)
        sync()
        t0 = time.perf_counter()
        ctx.solve(last)
        sync()
        dt = time.perf_counter() - t0
        out["energy_steps_per_s"] = steps / dt
        out["ms_per_step"] = dt / steps * 1e3
        out["plan"] = ctx.plan()
    finally:
        ctx.close()
    ctx = _lib.FcsContext(P)
    try:
        ctx.solve(warmup)
        ctx.set_timing(True)
        ctx.solve(last)
        sync()
        k = {}
        for name in ctx.KERNELS:
            ms, n, by, fl = ctx.kernel_stats(name)
            if n:
                k[name] = {"us_per_launch": round(ms / n * 1e3, 2),
                           "GB/s": round(by / (ms * 1e-3) / 1e9, 1),
                           "TFLOP/s": round(fl / (ms * 1e-3) / 1e12, 2)}
        out["kernels"] = k
        ms, n, by, fl = ctx.kernel_stats("collide")
        t_avg = ms / n * 1e-3
        plan = out["plan"]
        tiles = (P["C"] + 15) // 16
        executed = 2048.0 * tiles * (plan["nonzero_blocks"]
                                              + plan["groups"] * plan["column_blocks_per_wave"]
                                              * ((P["Q"] + 3) // 4))
        out["roofline_collide"] = {
            "bound": "mfma", "achieved": fl / n / t_avg / 1e12, "peak": PEAK_TFLOPS["f64"],
            "unit": "TFLOP/s", "frac": fl / n / t_avg / 1e12 / PEAK_TFLOPS["f64"],
            "traffic": MEASURED_TRAFFIC_FULL_PN["collide"], "kernel": "collide",
            "avg_launch_us": t_avg * 1e6,
            "hbm_GB/s": by / n / t_avg / 1e9,
            "executed_matrix_core_TFLOP/s": executed / t_avg / 1e12}
    finally:
        ctx.close()
    if cpu_steps:
        from oracle import fullpn
        t0 = time.perf_counter()
        fullpn.solve_fcs(P, steps=cpu_steps)
        out["cpu_oracle_s_per_step"] = (time.perf_counter() - t0) / cpu_steps
    return out


def adjoint_bench(sync, nthreads, n_beamlets=64, every=16, fd_dirs=8, fd_beamlets=8, seed=0):
    """Gradient of the plan objective J = ||sum_b w_b D_b(rho0 o s) - D_t||^2
    over 64 fixed-rank beamlets (128x128 phantom, P_15, ranks 5-30): forward
    solves with checkpoints every 16 steps, then the device adjoint of every
    beamlet, both batched with shared launches (titus_solve_batch_lockstep,
    titus_adjoint_batch_lockstep).

    dJ/ds is checked against central differences of J along `fd_dirs`
    random unit directions v (s +- h v, FP64 forward solves), beamlet by
    beamlet (J is a sum of beamlet terms), each direction on `fd_beamlets`
    beamlets drawn at random (the difference solves of all 64 would take six
    forward batches per direction).  The forward map is smooth only
    between its discrete decisions (kept ranks, completion vectors), and its
    rounding is amplified by near-null singular directions, so a difference
    quotient is trusted only where it is stable: for each beamlet, differences
    at h = 1e-5, 1e-6 and 1e-7 (relative) whose perturbed solves took the base
    solve's decisions (equal decision hashes); the beamlet counts when two
    consecutive step sizes agree to 1e-6, and the smaller one is compared.
    The reported error is that of the sum over the counted beamlets.  dJ/dw
    is exact (J is quadratic in w).
    """
    from titus_amd import _lib
    from titus_amd.adjoint import piecewise_target, plan_objective_rrule
    from titus_amd.beamlets import beamlet_plan
    from titus_amd.problem import INTEGRATOR_FIXED
    log = lambda m: print(f"[adjoint] {m}", file=sys.stderr, flush=True)
    plan = beamlet_plan(n_beamlets=n_beamlets, rank_range=(5, 30), integrator=INTEGRATOR_FIXED)
    problems = [P for _, P in plan]
    C = problems[0]["C"]
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.5, 1.5, n_beamlets)
    s = np.ones(C)
    rho0 = np.asarray(problems[0]["density"])
    steps = sum(P["nE"] - 1 for P in problems)
    log(f"{n_beamlets} beamlets built, {steps} beamlet steps")

    def solve_all(scale, which=None):
        which = range(n_beamlets) if which is None else which
        ctxs = [_lib.Context(dict(problems[b], density=rho0 * scale), INTEGRATOR_FIXED)
                for b in which]
        try:
            _lib.solve_batch(ctxs, nthreads)
            return (np.stack([c.state()["dose"] for c in ctxs], axis=1),
                    [c.decisions()[0] for c in ctxs])
        finally:
            for c in ctxs:
                c.close()

    D0, _ = solve_all(s)
    target = piecewise_target(C, problems[0]["ny"], seed=seed,
                              scale=2.0 * float(np.abs(D0 @ w).mean()))
    sync()
    t0 = time.perf_counter()
    # forward solves and adjoints with shared launches, all beamlets at once
    J, pullback = plan_objective_rrule(problems, w, s, target, every=every, nthreads=0)
    sync()
    t1 = time.perf_counter()
    wbar, sbar, per_b, hash0 = pullback(1.0, per_beamlet=True)
    sync()
    t2 = time.perf_counter()
    log(f"forward {t1 - t0:.2f} s, adjoint {t2 - t1:.2f} s")
    bars = 2.0 * pullback.residual[:, None] * w[None, :]            # dJ/dD_b
    steps_h = (1e-5, 1e-6, 1e-7)
    rows = []
    for k in range(fd_dirs):
        v = rng.standard_normal(C)
        v /= np.linalg.norm(v)
        sel = np.sort(rng.choice(n_beamlets, size=min(fd_beamlets, n_beamlets), replace=False))
        fds = []
        for hr in steps_h:
            h = hr * np.sqrt(C)
            Dp, hp = solve_all(s + h * v, sel)
            Dm, hm = solve_all(s - h * v, sel)
            bs = bars[:, sel]
            d = (np.einsum("cb,cb->b", bs, Dp) - np.einsum("cb,cb->b", bs, Dm)) / (2 * h)
            ok = np.array([hp[i] == hash0[b] and hm[i] == hash0[b] for i, b in enumerate(sel)])
            fds.append(np.where(ok, d, np.nan))
        adj_b = (per_b.T @ v)[sel]
        fd_sum, ad_sum, counted = 0.0, 0.0, 0
        for b in range(len(sel)):
            best = None
            for a in range(len(steps_h) - 1):
                x1, x2 = fds[a][b], fds[a + 1][b]
                if np.isfinite(x1) and np.isfinite(x2):
                    gap = abs(x1 - x2) / max(abs(x1), abs(x2), 1e-300)
                    if gap <= 1e-6 and (best is None or gap < best[0]):
                        best = (gap, x2)
            if best is None:
                continue
            fd_sum += best[1]
            ad_sum += adj_b[b]
            counted += 1
        rows.append(dict(rel_err=abs(fd_sum - ad_sum) / abs(fd_sum) if counted else None,
                         counted=counted))
        log(f"direction {k + 1}: rel err {rows[-1]['rel_err']} over {counted} of {len(sel)} beamlets")
    return {
        "workload": f"{n_beamlets} fixed-rank pencil beamlets on a 128x128 ellipse phantom, P_15, "
                    "6-20 MeV, +-15 deg, ranks 5-30, objective ||D(w, rho0 s) - D_t||^2 with a "
                    "seeded 4x4 piecewise-constant D_t, checkpoints every 16 steps, seed 0",
        "beamlet_steps": int(steps),
        "forward_with_checkpoints_s": t1 - t0,
        "adjoint_s": t2 - t1,
        "gradients_per_s": 1.0 / (t2 - t0),
        "adjoint_over_forward": (t2 - t1) / (t1 - t0),
        "grad_density_vs_differences": {
            "directions": fd_dirs, "h_relative": list(steps_h),
            "beamlets_per_direction": fd_beamlets,
            "max_rel_err": max([r["rel_err"] for r in rows if r["rel_err"] is not None] or [None]),
            "beamlets_counted_per_direction": [r["counted"] for r in rows],
        },
        "form": "forward and adjoint with shared launches (8 groups); difference solves one "
                f"stream per beamlet over {nthreads} host threads",
    }


def per_step(run, ms):
    """Milliseconds per sampled step (0 when nothing was sampled)."""
    return ms / run["sampled"] if run["sampled"] else 0.0


def roofline(stats, fp32_kernels=()):
    """Roofline of the dominant kernel (largest device time); `fp32_kernels`
    compute in FP32 (their flop peak is the FP32 one)."""
    dom = max(stats, key=lambda k: stats[k][0])
    ms, launches, by, fl = stats[dom]
    t_avg = ms / max(launches, 1) * 1e-3
    peak_f = PEAK_TFLOPS["f32" if dom in fp32_kernels else "f64"]
    ridge = peak_f * 1e12 / (HBM_PEAK_GBS * 1e9)
    if by > 0 and fl / by > ridge:
        roof = {"bound": "mfma", "achieved": fl / max(launches, 1) / t_avg / 1e12,
                "peak": peak_f, "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": by / max(launches, 1) / t_avg / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    meas = MEASURED_TRAFFIC.get(dom)
    roof["traffic"] = meas["bytes"] if meas else None
    if meas:
        roof["traffic_source"] = {k: v for k, v in meas.items() if k != "bytes"}
    roof["algorithmic_bytes_per_launch"] = by / max(launches, 1)
    roof["kernel"] = dom
    roof["avg_launch_us"] = t_avg * 1e6
    roof["launches"] = launches
    return roof


def pcie_inclusive(P, precision, W, K, sync, rank_tol):
    """The same solve from host buffers to host results: context creation
    (problem upload and the sweep tuner), steps 1..W+K, and the download of
    the state (dose, psi, factors).  Reported beside the headline, never as
    its value (the headline starts with the inputs resident in HBM)."""
    from titus_amd import _lib
    sync()
    t0 = time.perf_counter()
    ctx = _lib.Context(P, _lib.INTEGRATOR_BUG, precision=precision, rank_tol=rank_tol)
    try:
        sync()
        t1 = time.perf_counter()
        ctx.solve(W + K)
        sync()
        t2 = time.perf_counter()
        ctx.state()
        t3 = time.perf_counter()
    finally:
        ctx.close()
    return {"energy_steps_per_s": (W + K) / (t3 - t0), "steps": W + K,
            "create_s": t1 - t0, "solve_s": t2 - t1, "download_s": t3 - t2}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--advance", type=int, default=300,
                    help="untimed steps before the warmup (the solve's ranks grow over the "
                         "first steps; the window then sits at the rank the solve keeps)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-events", action="store_true",
                    help="time without per-kernel HIP events (no roofline)")
    ap.add_argument("--event-period", type=int, default=20,
                    help="HIP events on every n-th step of the timed region (a sampled step "
                         "brackets every launch with events and replays no captured graph; "
                         "one in 20 gives the roofline kernel 16 timed launches per sample)")
    ap.add_argument("--beamlets", type=int, default=64,
                    help="beamlets of the batched plan (0: skip)")
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host workers of the beamlet batch")
    ap.add_argument("--no-stencil", action="store_true",
                    help="skip the large-grid stencil measurement")
    ap.add_argument("--no-slab", action="store_true",
                    help="skip the 80x80 water-slab configuration")
    ap.add_argument("--no-cube", action="store_true",
                    help="skip the 3D proton-cube measurement")
    ap.add_argument("--cube-only", action="store_true",
                    help="only the 3D proton-cube measurement")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the plan-gradient (adjoint) measurement")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the plan-gradient (adjoint) measurement")
    ap.add_argument("--fd-dirs", type=int, default=8,
                    help="random directions of the plan-gradient difference check")
    ap.add_argument("--no-full-pn", action="store_true",
                    help="skip the full-rank P_N solver measurement")
    ap.add_argument("--full-pn-only", action="store_true",
                    help="only the full-rank P_N solver measurement")
    ap.add_argument("--stencil-only", action="store_true",
                    help="only the large-grid stencil measurement (+ calibration copy), for PMC runs")
    ap.add_argument("--profile-form", choices=("fp64", "mixed"), default=None,
                    help="only this form's headline solve (steps 1..advance+warmup+steps, the "
                         "window timed), for profiler passes; prints its rate")
    ap.add_argument("--beamlets-only", action="store_true",
                    help="only the batched beamlet plan")
    ap.add_argument("--profile-batch", action="store_true",
                    help="only the mixed shared-launch batch of the beamlet plan, for profiler runs")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the timed FP64 and mixed phantom windows (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(beamlet sharding across GPUs is a design note, not code)", file=sys.stderr)
        sys.exit(2)

    from titus_amd import _lib
    from titus_amd.problem import INTEGRATOR_BUG
    from titus_amd.solver import MIXED_RANK_TOL, RANK_TOL
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None
    if args.cube_only:
        print(json.dumps({"proton_cube": proton_cube_bench(sync, 0 if args.no_cpu else 2)}), flush=True)
        return
    if args.stencil_only:
        print(json.dumps({"stencil_scale": stencil_scale(sync, calibrate=True)}), flush=True)
        return
    if args.full_pn_only:
        print(json.dumps({"full_pn": full_pn_bench(sync, 0 if args.no_cpu else 2)}), flush=True)
        return
    if args.profile_batch:
        print(json.dumps({"beamlets": beamlet_batch(args.beamlets, args.threads, sync, profile_only=True)}),
              flush=True)
        return
    if args.beamlets_only:
        print(json.dumps({"beamlets": beamlet_batch(args.beamlets, args.threads, sync)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_bench(sync, args.threads, fd_dirs=args.fd_dirs)}), flush=True)
        return
    P = build_workload()
    last = P["nE"] - 2
    A = max(0, min(args.advance, last - 2))
    W = max(0, min(args.warmup, last - A - 1))
    K = max(1, min(args.steps, last - A - W))
    period = 0 if args.no_events else args.event_period
    if args.profile_form:
        prec = _lib.PRECISION_MIXED if args.profile_form == "mixed" else _lib.PRECISION_FP64
        one = run_form(P, prec, A + W, K, 0, sync, MIXED_RANK_TOL)
        print(json.dumps({"form": args.profile_form, "energy_steps_per_s": one["rate"],
                          "steps_timed": f"{A + W + 1}..{A + W + K}",
                          "rank_range": [int(one["ranks"][A + W:].min()),
                                         int(one["ranks"][A + W:].max())]}), flush=True)
        return
    # both timed forms at the mixed form's rank floor (FP32 noise must not
    # decide which near-null directions enter the basis)
    f64 = run_form(P, _lib.PRECISION_FP64, A + W, K, period, sync, MIXED_RANK_TOL)
    mix = run_form(P, _lib.PRECISION_MIXED, A + W, K, period, sync, MIXED_RANK_TOL)
    derr = rel(mix["dose"], f64["dose"])
    n_r = min(len(mix["ranks"]), len(f64["ranks"]))
    diff = np.nonzero(np.asarray(mix["ranks"][:n_r]) != np.asarray(f64["ranks"][:n_r]))[0]
    window = f"{A + W + 1}..{A + W + K}"
    out = {
        "metric": METRIC,
        "value": mix["rate"],
        "unit": "energy steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": mix["elapsed"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed (f32 psi storage and uncollided sweep; f64 stencils, Galerkin, QR/SVD, dose)",
        "data": "synthetic (seeded ellipse phantom, analytic beam source)",
        "config": {
            "workload": f"200x200 phantom P_21 BUG, steps {window} of {P['nE'] - 1}; 2D "
                        "heterogeneous phantom over 10x10 cm, 15 MeV e-, 253 moments, "
                        "rank-adaptive BUG theta=1e-2, rank cap 40, single beamlet",
            "cells": P["C"], "moments": P["Nm"], "ordinates": P["Q"],
            "steps_timed": window,
            "untimed_advance": A,
            "rank_range": [int(mix["ranks"][A + W:].min()), int(mix["ranks"][A + W:].max())],
            "parallelism": "single GPU",
        },
        "fp64": {"value": f64["rate"], "ms_per_step": f64["elapsed"] / K * 1e3,
                 "device_ms_per_step": per_step(f64, sum(v[0] for v in f64["stats"].values()))},
        "mixed_over_fp64": mix["rate"] / f64["rate"],
        "dose_rel_l2_mixed_vs_fp64": derr,
        "rank_history_mixed_vs_fp64": {"equal": bool(len(diff) == 0),
                                       "first_difference_step": int(diff[0]) + 1 if len(diff) else None},
        "rank_tol": MIXED_RANK_TOL,
        "device_ms_per_step": per_step(mix, sum(v[0] for v in mix["stats"].values())),
        "kernel_ms_per_step": {k: round(per_step(mix, v[0]), 4) for k, v in mix["stats"].items()},
        "host_ms_per_step": {k.split(":")[1]: round(per_step(mix, v), 4) for k, v in mix["host"].items()},
        "events_in_timed_region": {"period": period, "sampled_steps": mix["sampled"]},
    }
    if period:
        out["roofline"] = roofline(mix["stats"], fp32_kernels=("sweep",))
        out["roofline_fp64"] = roofline(f64["stats"])
    if args.headline_only:
        print(json.dumps(out), flush=True)
        return
    # whole solves (every step of the grid) beside the window
    whole = {}
    for name, prec, tol in (("fp64", _lib.PRECISION_FP64, MIXED_RANK_TOL),
                            ("mixed", _lib.PRECISION_MIXED, MIXED_RANK_TOL),
                            ("fp64_plain_floor", _lib.PRECISION_FP64, RANK_TOL),
                            ("mixed_full_fp32", _lib.PRECISION_MIXED_FULL_FP32, MIXED_RANK_TOL)):
        whole[name], whole[name + "_dose"] = whole_solve(P, INTEGRATOR_BUG, prec, tol, sync)
    out["whole_solve"] = {k: v for k, v in whole.items() if not k.endswith("_dose")}
    out["whole_solve"]["dose_rel_l2"] = {
        "mixed_vs_fp64_same_floor": rel(whole["mixed_dose"], whole["fp64_dose"]),
        "mixed_vs_fp64_plain_floor": rel(whole["mixed_dose"], whole["fp64_plain_floor_dose"]),
        "fp64_floor_1e-4_vs_1e-7": rel(whole["fp64_dose"], whole["fp64_plain_floor_dose"]),
        "mixed_full_fp32_vs_fp64_same_floor": rel(whole["mixed_full_fp32_dose"], whole["fp64_dose"]),
    }
    out["pcie_inclusive_mixed"] = pcie_inclusive(P, _lib.PRECISION_MIXED, A + W, K, sync,
                                                 MIXED_RANK_TOL)
    if not args.no_slab:
        out["slab"] = slab_bench(sync, not args.no_cpu)
    if args.beamlets > 0:
        out["beamlets"] = beamlet_batch(args.beamlets, args.threads, sync)
    if not args.no_stencil:
        out["stencil_scale"] = stencil_scale(sync)
    if not args.no_cube:
        out["proton_cube"] = proton_cube_bench(sync, 0 if args.no_cpu else 2)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench(sync, args.threads, fd_dirs=args.fd_dirs)
    if not args.no_full_pn:
        out["full_pn"] = full_pn_bench(sync, 0 if args.no_cpu else 2)
    if args.no_cpu:
        out["cpu_baseline"] = None
    else:
        rec, s1, ref = cpu_baseline(P, args.cpu_seconds, A + W)
        out["cpu_baseline"] = rec
        from oracle import dlra
        saved = dlra.RANK_TOL
        try:
            dlra.RANK_TOL = MIXED_RANK_TOL
            ref4 = dlra.solve_bug(P, steps=s1)["dose"]
        finally:
            dlra.RANK_TOL = saved
        mixed_dose = device_dose(P, INTEGRATOR_BUG, _lib.PRECISION_MIXED, s1, MIXED_RANK_TOL)
        out["phantom_vs_oracle"] = {
            "step": s1,
            "dose_rel_l2_fp64": rel(device_dose(P, INTEGRATOR_BUG, _lib.PRECISION_FP64, s1, RANK_TOL), ref),
            "dose_rel_l2_mixed_vs_oracle_same_floor": rel(mixed_dose, ref4),
            "dose_rel_l2_mixed_vs_oracle_plain_floor": rel(mixed_dose, ref),
            "note": "the mixed form runs at the rank floor 1e-4; against the FP64 oracle at that "
                    "floor its difference is the precision's, against the oracle at the plain "
                    "floor 1e-7 it is mostly the floor's (whole_solve.dose_rel_l2)",
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
