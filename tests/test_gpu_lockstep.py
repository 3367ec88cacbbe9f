"""Lockstep beamlet batches (titus_solve_batch_lockstep) on the GPU.

Every solve of a lockstep batch computes what it computes alone: doses,
factors and rank decisions equal independent titus_solve runs bit for bit.
The batches mix beamlets of different ranks, ordinate counts and energy
grids (solves end at different steps, so the group shrinks and, with a
width below the batch size, takes the next context), FP64 and mixed forms,
and the rank-adaptive, fixed-rank and parallel integrators side by side.
"""
import ctypes

import numpy as np
import pytest

import cases
from oracle import dlra
from titus_amd import _lib
from titus_amd.beamlets import beamlet_plan
from titus_amd.problem import INTEGRATOR_BUG, INTEGRATOR_FIXED, INTEGRATOR_PARALLEL

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def small_plan():
    return beamlet_plan(n_beamlets=6, n=40, N=7, seed=3, rank_range=(4, 12), e_range=(3.0, 6.0))


def solve_alone(P, integrator, precision):
    ctx = _lib.Context(P, integrator, precision=precision)
    try:
        ctx.solve(0)
        st = ctx.state()
        done, _ = ctx.progress()
        st["ranks"] = ctx.history(done)[0]
        st["decisions"] = ctx.decisions()
    finally:
        ctx.close()
    return st


def solve_lockstep(problems, width, precision):
    ctxs = [_lib.Context(P, integ, precision=precision) for P, integ in problems]
    try:
        status, stats = _lib.solve_batch_lockstep(ctxs, width)
        out = []
        for c in ctxs:
            st = c.state()
            done, _ = c.progress()
            st["ranks"] = c.history(done)[0]
            st["decisions"] = c.decisions()
            out.append(st)
    finally:
        for c in ctxs:
            c.close()
    return status, stats, out


def same(a, b):
    return (np.array_equal(a["dose"], b["dose"]) and np.array_equal(a["X"], b["X"])
            and np.array_equal(a["S"], b["S"]) and np.array_equal(a["W"], b["W"])
            and np.array_equal(a["ranks"], b["ranks"]) and a["decisions"] == b["decisions"])


@pytest.mark.parametrize("precision", [_lib.PRECISION_FP64, _lib.PRECISION_MIXED])
@pytest.mark.parametrize("width", [6, 4])
def test_lockstep_equals_alone_bitwise(small_plan, precision, width):
    print(f"\n[case] lockstep batch of {len(small_plan)} beamlets 40x40 P_7, width {width}, "
          f"precision {precision}", flush=True)
    problems = [(P, INTEGRATOR_BUG) for _, P in small_plan]
    status, stats, got = solve_lockstep(problems, width, precision)
    assert np.all(status == 0)
    steps = sum(P["nE"] - 2 for P, _ in problems)
    print(f"[lockstep] {steps} beamlet steps, {stats['flushes']} flushes, {stats['launches']} launches "
          f"carrying {stats['ops']} recorded launches", flush=True)
    assert stats["launches"] < stats["ops"]
    for b, (P, integ) in enumerate(problems):
        ref = solve_alone(P, integ, precision)
        assert same(got[b], ref), b


def test_lockstep_mixed_integrators_bitwise():
    """Fixed-rank, rank-adaptive and parallel solves in one group."""
    problems = [(cases.small_fixed(n=24, N=5, r=6), INTEGRATOR_FIXED),
                (cases.small_bug(n=30, N=7, cap=12, seed=3), INTEGRATOR_BUG),
                (cases.small_parallel(n=30, N=7, cap=12, seed=3), INTEGRATOR_PARALLEL),
                (cases.small_fixed(n=32, N=3, r=4), INTEGRATOR_FIXED)]
    print("\n[case] lockstep group of fixed, BUG and parallel solves", flush=True)
    status, stats, got = solve_lockstep(problems, 0, _lib.PRECISION_FP64)
    assert np.all(status == 0)
    for b, (P, integ) in enumerate(problems):
        assert same(got[b], solve_alone(P, integ, _lib.PRECISION_FP64)), b


def test_lockstep_matches_oracle(small_plan):
    problems = [(P, INTEGRATOR_BUG) for _, P in small_plan[:3]]
    _, _, got = solve_lockstep(problems, 0, _lib.PRECISION_FP64)
    for b, (P, _) in enumerate(problems):
        e = rel(got[b]["dose"], dlra.solve_bug(P)["dose"])
        print(f"[lockstep oracle] beamlet {b}: dose rel {e:.3e}", flush=True)
        assert e <= 1e-10


def test_lockstep_refuses_unbatched_contexts(small_plan):
    """3D grids and the limited flux are refused before any solve runs."""
    from titus_amd.problem3d import build_problem3d, proton_cube
    P3 = build_problem3d(proton_cube(n=8, N=3, r=4))
    P = small_plan[0][1]
    a = _lib.Context(P, INTEGRATOR_BUG)
    b = _lib.Context(P3, INTEGRATOR_FIXED)
    c = _lib.Context(dict(cases.small_fixed(n=24, N=5, r=6), flux_order=2), INTEGRATOR_FIXED)
    try:
        for bad in (b, c):
            with pytest.raises(RuntimeError):
                _lib.solve_batch_lockstep([a, bad], 0)
            assert a.progress()[0] == 0 and bad.progress()[0] == 0
    finally:
        for x in (a, b, c):
            x.close()


def test_lockstep_forward_and_adjoint_equal_alone():
    """Checkpointed fixed-rank solves and their density gradients with shared
    launches equal the same solves and adjoints run alone, bit for bit; the
    forward state is back afterwards."""
    problems = [cases.small_fixed(n=24, N=5, r=6), cases.small_fixed(n=32, N=3, r=4),
                cases.small_fixed(n=24, N=5, r=8, angular=1.0), cases.small_fixed(n=28, N=5, r=5)]
    rng = np.random.default_rng(5)
    bars = [rng.standard_normal(P["C"]) for P in problems]
    print("\n[case] lockstep forward with checkpoints and adjoint, 4 fixed-rank solves", flush=True)

    def alone(P, bar):
        c = _lib.Context(P, INTEGRATOR_FIXED)
        try:
            c.set_checkpoints(7)
            c.solve(0)
            d = c.state()["dose"]
            return d, c.adjoint(bar)
        finally:
            c.close()

    ctxs = [_lib.Context(P, INTEGRATOR_FIXED) for P in problems]
    try:
        for c in ctxs:
            c.set_checkpoints(7)
        _lib.solve_batch_lockstep(ctxs, 0, 2)
        doses = [c.state()["dose"] for c in ctxs]
        # the grids differ in size: each context takes its own seed buffer
        lib = _lib.load()
        n = len(ctxs)
        bufs = [np.ascontiguousarray(b) for b in bars]
        outs = [np.zeros(P["C"]) for P in problems]
        arr = (ctypes.c_void_p * n)(*[c.h for c in ctxs])
        pb = (_lib._P * n)(*[_lib._dp(b) for b in bufs])
        pg = (_lib._P * n)(*[_lib._dp(g) for g in outs])
        status = np.zeros(n, dtype=np.int32)
        _lib.check(lib.titus_adjoint_batch_lockstep(arr, n, 0, 2, pb, pg, _lib._ip(status)))
        after = [c.state()["dose"] for c in ctxs]
    finally:
        for c in ctxs:
            c.close()
    for k, P in enumerate(problems):
        d, g = alone(P, bars[k])
        assert np.array_equal(doses[k], d), k
        assert np.array_equal(outs[k], g), k
        assert np.array_equal(after[k], d), k


@pytest.mark.parametrize("precision", [_lib.PRECISION_MIXED, _lib.PRECISION_FP64])
def test_lockstep_tuned_tiles_equal_streams(precision):
    """Eight beamlets of the 128 x 128 P_15 plan, large enough for the sweep
    plan tuner at creation and for the batch's own tile tuner (strip heights,
    non-temporal streaming, tiles per wavefront): the tuned batch equals the
    one-stream-per-beamlet batch bit for bit."""
    plan = beamlet_plan(n_beamlets=8)
    print(f"\n[case] lockstep vs streams, 8 beamlets 128x128 P_15, precision {precision}", flush=True)
    problems = [(P, INTEGRATOR_BUG) for _, P in plan]
    status, stats, got = solve_lockstep(problems, 0, precision)
    assert np.all(status == 0)
    print(f"[lockstep tuned] tiles {stats['tiles']}", flush=True)
    ctxs = [_lib.Context(P, INTEGRATOR_BUG, precision=precision) for P, _ in problems]
    try:
        _lib.solve_batch(ctxs, 8)
        for b, c in enumerate(ctxs):
            assert np.array_equal(got[b]["dose"], c.state()["dose"]), b
    finally:
        for c in ctxs:
            c.close()
