"""Batched beamlet solves on the GPU: one stream per beamlet (titus_solve_batch)
and, for the mixed-form check, shared launches (titus_solve_batch_lockstep).

Pins: the concurrent batch equals independent sequential solves bit for bit
in FP64 (each solve's arithmetic does not depend on what runs beside it);
the mixed-precision batch stays within 1e-4 (dose relative L2) of the FP64
solve of every beamlet at the same rank cut; a small batch matches the FP64 oracle to 1e-10.
"""
import numpy as np
import pytest

from oracle import dlra
from titus_amd import _lib
from titus_amd.beamlets import beamlet_plan, plan_dose, solve_plan
from titus_amd.solver import MIXED_RANK_TOL

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def small_plan():
    return beamlet_plan(n_beamlets=6, n=40, N=7, seed=3, rank_range=(4, 12), e_range=(3.0, 6.0))


def test_batch_equals_sequential_bitwise(small_plan):
    print(f"\n[case] batch of {len(small_plan)} beamlets, 40x40, P_7", flush=True)
    batched, steps = solve_plan(small_plan, precision=_lib.PRECISION_FP64, nthreads=6, lockstep=False)
    seq = []
    for _, P in small_plan:
        ctx = _lib.Context(P, _lib.INTEGRATOR_BUG)
        try:
            ctx.solve(0)
            seq.append(ctx.state()["dose"])
        finally:
            ctx.close()
    seq = np.stack(seq, axis=1)
    assert np.array_equal(batched, seq)
    assert np.all(steps == np.array([P["nE"] - 2 for _, P in small_plan]))


def test_batch_matches_oracle(small_plan):
    batched, _ = solve_plan(small_plan, precision=_lib.PRECISION_FP64, nthreads=3, lockstep=False)
    for b in (0, 3):
        ref = dlra.solve_bug(small_plan[b][1])["dose"]
        e = rel(batched[:, b], ref)
        print(f"[batch oracle] beamlet {b}: dose rel {e:.3e}", flush=True)
        assert e <= 1e-10


def test_mixed_batch_per_beamlet_error(small_plan):
    # both forms at the mixed form's rank cut (see test_gpu_parity)
    f64, _ = solve_plan(small_plan, precision=_lib.PRECISION_FP64, nthreads=6, rank_tol=MIXED_RANK_TOL)
    mix, _ = solve_plan(small_plan, precision=_lib.PRECISION_MIXED, nthreads=6, rank_tol=MIXED_RANK_TOL)
    errs = [rel(mix[:, b], f64[:, b]) for b in range(f64.shape[1])]
    print("[mixed batch] per-beamlet dose rel:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= 1e-4
    w = np.linspace(0.5, 1.5, f64.shape[1])
    assert rel(plan_dose(mix, w), plan_dose(f64, w)) <= 1e-4
