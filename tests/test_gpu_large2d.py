"""2D solves on a grid of 384 x 384 cells (147 456 cells, at least 2^17, so the
Galerkin products walk strips of 16 cells row by row) against the oracle.

* fixed rank 20 (two 16-column tiles per side, the second one partial), the
  broad beam of the small CPU cases on a 3 x 3 cm grid, P_5, the first 4
  steps of its energy grid;
* the rank-adaptive (cap 10) and the parallel BUG integrators on the 10 x 10 cm
  heterogeneous phantom, P_5, the first 6 steps (equal rank histories).

Every energy grid follows the reference's step rule (CFL 0.67-0.75 here).

Bars as for the benchmark configurations: psi 1e-13, dose 1e-10 and the
solution u = X S W^T 1e-8, relative L2, all FP64.
"""
import sys
import time

import numpy as np
import pytest

import cases
from oracle import dlra
from titus_amd import _lib
from titus_amd.problem import INTEGRATOR_BUG, INTEGRATOR_FIXED, INTEGRATOR_PARALLEL

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def run(P, integrator, last):
    ctx = _lib.Context(P, integrator, precision=_lib.PRECISION_FP64)
    try:
        ctx.solve(last)
        st = ctx.state()
        done, _ = ctx.progress()
        st["ranks"] = ctx.history(done)[0]
    finally:
        ctx.close()
    return st


@pytest.mark.parametrize("name, make, integrator, solve, steps", [
    ("fixed rank 20", lambda: cases.small_fixed(n=384, N=5, r=20), INTEGRATOR_FIXED,
     dlra.solve_fixed, 4),
    ("BUG", lambda: cases.small_bug(n=384, N=5, cap=10, seed=3), INTEGRATOR_BUG,
     dlra.solve_bug, 6),
    ("parallel BUG", lambda: cases.small_parallel(n=384, N=5, cap=10, seed=3),
     INTEGRATOR_PARALLEL, dlra.solve_parallel, 6),
])
def test_large_2d_grid_vs_oracle(name, make, integrator, solve, steps):
    P = make()
    assert P["nx"] * P["ny"] >= 1 << 17
    print(f"\n[case] 384x384 P_5 {name}, {steps} steps", flush=True)
    sys.stdout.flush()
    g = run(P, integrator, steps)
    t = time.perf_counter()
    o = solve(P, steps=steps)
    ed = rel(g["dose"], o["dose"])
    eu = rel(g["X"] @ g["S"] @ g["W"].T, o["X"] @ o["S"] @ o["W"].T)
    ep = rel(g["psi"], o["psi"])
    print(f"[384^2 {name}] oracle {time.perf_counter() - t:.1f} s dose rel {ed:.3e} u rel {eu:.3e} "
          f"psi rel {ep:.3e} ranks {g['ranks'].tolist()}", flush=True)
    if "ranks" in o:
        assert list(g["ranks"]) == list(o["ranks"])
    assert ep <= 1e-13
    assert ed <= 1e-10
    assert eu <= 1e-8
