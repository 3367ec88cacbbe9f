"""Captured range jobs on the GPU.

A lone context captures the device half of each range basis into a HIP graph
the first time a (K block, basis block, widths, arena position) occurs and
replays it afterwards.  Timing every step sends every range job through
direct launches instead (events bracket each launch), so a solve with timing
on is the eager form of the same solve.  Each solve below replays its graphs
many times, with the step's own eager launches, copies and stream syncs
between the replays; its doses, factors, rank histories and decision hashes
must equal the eager solve's bit for bit, at every checked step.
"""
import numpy as np
import pytest

import cases
from titus_amd import _lib
from titus_amd.problem import INTEGRATOR_BUG, INTEGRATOR_FIXED, INTEGRATOR_PARALLEL

pytestmark = pytest.mark.gpu


def states(P, integrator, marks, eager, precision=_lib.PRECISION_FP64):
    ctx = _lib.Context(P, integrator, precision=precision)
    out = []
    try:
        if eager:
            ctx.set_timing(1)
        for s in marks:
            ctx.solve(s)
            st = ctx.state()
            done, _ = ctx.progress()
            st["ranks"] = ctx.history(done)[0]
            st["decisions"] = ctx.decisions()
            out.append(st)
    finally:
        ctx.close()
    return out


def same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("dose", "X", "S", "W", "psi", "ranks")) \
        and a["decisions"] == b["decisions"]


@pytest.mark.parametrize("name, make, integrator, precision", [
    ("BUG phantom 30^2 P_7", lambda: cases.small_bug(n=30, N=7, cap=12, seed=3), INTEGRATOR_BUG,
     _lib.PRECISION_FP64),
    ("BUG phantom 30^2 P_7 mixed", lambda: cases.small_bug(n=30, N=7, cap=12, seed=3), INTEGRATOR_BUG,
     _lib.PRECISION_MIXED),
    ("fixed rank 24^2 P_5 r6", lambda: cases.small_fixed(n=24, N=5, r=6), INTEGRATOR_FIXED,
     _lib.PRECISION_FP64),
    ("parallel BUG 30^2 P_7", lambda: cases.small_parallel(n=30, N=7, cap=12, seed=3), INTEGRATOR_PARALLEL,
     _lib.PRECISION_FP64),
])
def test_replayed_range_jobs_equal_direct_launches(name, make, integrator, precision):
    P = make()
    last = P["nE"] - 2
    marks = sorted({min(last, m) for m in (5, 20, 60, last)})
    print(f"\n[case] graphs vs direct launches: {name}, checks at steps {marks}", flush=True)
    replayed = states(P, integrator, marks, eager=False, precision=precision)
    direct = states(P, integrator, marks, eager=True, precision=precision)
    for s, a, b in zip(marks, replayed, direct):
        assert same(a, b), (name, s)
    print(f"[graphs] {name}: equal at {len(marks)} steps, ranks {int(min(a['ranks']))}.."
          f"{int(max(a['ranks']))}", flush=True)


def test_replayed_range_jobs_on_the_headline_phantom():
    """The 200 x 200 P_21 phantom over 120 steps (its rank grows from 1 to 9,
    so graphs are captured, replaced and replayed)."""
    P = cases.phantom_bug()
    print("\n[case] graphs vs direct launches: phantom 200x200 P_21, steps 40 and 120", flush=True)
    a = states(P, INTEGRATOR_BUG, [40, 120], eager=False)
    b = states(P, INTEGRATOR_BUG, [40, 120], eager=True)
    for x, y in zip(a, b):
        assert same(x, y)
