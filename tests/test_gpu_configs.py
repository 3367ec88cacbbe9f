"""The benchmark configurations at their stated sizes on the GPU.

* 2D phantom, 200 x 200 cells, P_21 (253 moments), rank-adaptive BUG
  (theta = 1e-2, cap 40): FP64 device vs the FP64 oracle at energy steps
  200 and 558 (dose relative L2 <= 1e-10, equal rank histories); the mixed form
  against the FP64 solve at the mixed rank floor and at the plain one.
* 3D proton cube, 64^3 cells of 0.1 cm, P_13 (196 moments), fixed rank 30:
  FP64 device vs the oracle over the first 8 steps; the whole 300-step solve
  bitwise reproducible, finite, and the mixed form within 1e-4 of it.
* 64 pencil beamlets on a 128 x 128 phantom, P_15, ranks 5-40: the batch
  equals independent solves bit for bit (FP64); two beamlets against the
  oracle over their first 120 steps; the mixed plan dose within 1e-4 of the
  FP64 one at the same rank floor, and the beamlets' own errors (see the
  test for the conditioning bound).
* The gradient configuration (fixed rank, ranks 5-30): one full-size beamlet's
  density gradient against the oracle adjoint over its first 32 steps with
  checkpoints every 16 (the oracle adjoint of a 128 x 128 beamlet takes
  about a minute for them).

Each case prints its name before it starts and its errors when it ends.
"""
import sys
import time

import numpy as np
import pytest

import cases
from oracle import adjoint as oadj
from oracle import dlra
from titus_amd import _lib
from titus_amd.beamlets import beamlet_plan, solve_plan
from titus_amd.problem import INTEGRATOR_BUG, INTEGRATOR_FIXED
from titus_amd.problem3d import build_problem3d, proton_cube
from titus_amd.solver import MIXED_RANK_TOL, RANK_TOL

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def announce(name):
    print(f"\n[case] {name}", flush=True)
    sys.stdout.flush()


def run(P, integrator, last=0, precision=_lib.PRECISION_FP64, rank_tol=RANK_TOL):
    ctx = _lib.Context(P, integrator, precision=precision, rank_tol=rank_tol)
    try:
        ctx.solve(last)
        st = ctx.state()
        done, _ = ctx.progress()
        st["ranks"] = ctx.history(done)[0]
        st["decisions"] = ctx.decisions()
    finally:
        ctx.close()
    return st


# ---------------------------------------------------------------------------
# 200 x 200 phantom, P_21, BUG
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def phantom():
    return cases.phantom_bug()


def test_phantom_200_p21_fp64_vs_oracle(phantom):
    """FP64 device vs the oracle at step 200 and at step 558 (the end of the
    bench's CPU-baseline window): dose <= 1e-10 at both, equal rank histories.
    The truncation's SVD follows LAPACK's algorithm family (bidiagonalization,
    then QR sweeps); with a one-sided Jacobi SVD there the device drifted from
    the oracle to 1.2e-10 by step 558 while staying at 1e-12 of an oracle run
    with the same Jacobi SVD."""
    P = phantom
    announce(f"phantom 200x200 P_21 BUG, 558 steps: C={P['C']} Nm={P['Nm']} Q={P['Q']}")
    assert P["nx"] == 200 and P["ny"] == 200 and P["Nm"] == 253 and P["rank_cap"] == 40
    t = time.perf_counter()
    ctx = _lib.Context(P, INTEGRATOR_BUG)
    try:
        ctx.solve(200)
        g200 = ctx.state()
        ctx.solve(558)
        g = ctx.state()
        g["ranks"] = ctx.history(ctx.progress()[0])[0]
    finally:
        ctx.close()
    tg = time.perf_counter() - t
    marks = {}

    def keep(s, dose):
        if s == 200:
            marks[s] = dose.copy()

    t = time.perf_counter()
    o = dlra.solve_bug(P, steps=558, on_step=keep)
    to = time.perf_counter() - t
    e200 = rel(g200["dose"], marks[200])
    ed = rel(g["dose"], o["dose"])
    eu = rel(g["X"] @ g["S"] @ g["W"].T, o["X"] @ o["S"] @ o["W"].T)
    print(f"[phantom fp64] device {tg:.2f} s oracle {to:.1f} s dose rel at 200 {e200:.3e}, at 558 "
          f"{ed:.3e}; u rel {eu:.3e} ranks {int(min(o['ranks']))}..{int(max(o['ranks']))}", flush=True)
    assert np.array_equal(g["ranks"], np.asarray(o["ranks"]))
    assert e200 <= 1e-10
    assert ed <= 1e-10
    assert eu <= 1e-8


def test_phantom_200_p21_mixed_form(phantom):
    """Mixed (FP32 psi and sweep) against FP64 at the same rank floor over
    320 steps; the FP64 solve at the plain floor is reported beside it."""
    P = phantom
    announce("phantom 200x200 P_21 BUG, mixed vs FP64, 320 steps")
    f64 = run(P, INTEGRATOR_BUG, 320, rank_tol=MIXED_RANK_TOL)
    mix = run(P, INTEGRATOR_BUG, 320, _lib.PRECISION_MIXED, rank_tol=MIXED_RANK_TOL)
    plain = run(P, INTEGRATOR_BUG, 320)
    e = rel(mix["dose"], f64["dose"])
    print(f"[phantom mixed] vs FP64 at floor {MIXED_RANK_TOL:g}: {e:.3e}; vs FP64 at "
          f"{RANK_TOL:g}: {rel(mix['dose'], plain['dose']):.3e}; ranks equal "
          f"{np.array_equal(mix['ranks'], f64['ranks'])}", flush=True)
    assert np.all(np.isfinite(mix["dose"]))
    assert e <= 1e-4


# ---------------------------------------------------------------------------
# 64^3 proton cube, P_13, rank 30
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def cube():
    return build_problem3d(proton_cube())


def test_proton_cube_64_p13_fp64_vs_oracle(cube):
    P = cube
    announce(f"proton cube 64^3 P_13 r=30, 8 steps: Nm={P['Nm']} Q={P['Q']} nE={P['nE']}")
    assert (P["nx"], P["ny"], P["nz"]) == (64, 64, 64) and P["Nm"] == 196 and P["r"] == 30
    g = run(P, INTEGRATOR_FIXED, 8)
    t = time.perf_counter()
    o = dlra.solve_fixed(P, steps=8)
    ed = rel(g["dose"], o["dose"])
    eu = rel(g["X"] @ g["S"] @ g["W"].T, o["X"] @ o["S"] @ o["W"].T)
    print(f"[cube fp64] oracle {time.perf_counter() - t:.1f} s dose rel {ed:.3e} u rel {eu:.3e} "
          f"psi rel {rel(g['psi'], o['psi']):.3e}", flush=True)
    assert rel(g["psi"], o["psi"]) <= 1e-13
    assert ed <= 1e-10
    assert eu <= 1e-8


def test_proton_cube_64_p13_whole_solve(cube):
    """All 300 steps: bitwise reproducible in FP64, finite, nonnegative
    dose where the beam deposits, mixed within 1e-4 of FP64."""
    P = cube
    announce("proton cube 64^3 P_13 r=30, 300 steps, FP64 twice and mixed")
    a = run(P, INTEGRATOR_FIXED, rank_tol=MIXED_RANK_TOL)
    b = run(P, INTEGRATOR_FIXED, rank_tol=MIXED_RANK_TOL)
    m = run(P, INTEGRATOR_FIXED, precision=_lib.PRECISION_MIXED, rank_tol=MIXED_RANK_TOL)
    e = rel(m["dose"], a["dose"])
    print(f"[cube 300] mixed vs FP64 {e:.3e}; max dose {a['dose'].max():.3e}", flush=True)
    assert np.array_equal(a["dose"], b["dose"]) and a["decisions"] == b["decisions"]
    assert np.all(np.isfinite(a["dose"])) and a["dose"].max() > 0
    assert e <= 1e-4


# ---------------------------------------------------------------------------
# 64 beamlets, 128 x 128, P_15
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def plan():
    return beamlet_plan(n_beamlets=64)


def test_beamlet_plan_batch_equals_independent_solves(plan):
    announce("64-beamlet plan 128x128 P_15: batch vs independent solves, FP64")
    assert len(plan) == 64 and plan[0][1]["Nm"] == 136 and plan[0][1]["nx"] == 128
    batched, steps = solve_plan(plan, precision=_lib.PRECISION_FP64)
    # each solve's arithmetic does not depend on what runs beside it or on
    # its launches being shared: compare eight of them, solved alone, bit for bit
    for b in range(0, 64, 8):
        g = run(plan[b][1], INTEGRATOR_BUG)
        assert np.array_equal(g["dose"], batched[:, b]), b
    print(f"[plan batch] {int(steps.sum())} beamlet steps, 8 of 64 checked bitwise", flush=True)


@pytest.mark.parametrize("b", [0, 59])
def test_beamlet_vs_oracle(plan, b):
    P = plan[b][1]
    announce(f"beamlet {b} of the 128x128 P_15 plan vs oracle, 120 steps (Q={P['Q']})")
    g = run(P, INTEGRATOR_BUG, 120)
    o = dlra.solve_bug(P, steps=120)
    ed = rel(g["dose"], o["dose"])
    print(f"[beamlet {b}] dose rel {ed:.3e} ranks {int(max(o['ranks']))} max", flush=True)
    assert np.array_equal(g["ranks"], np.asarray(o["ranks"]))
    assert ed <= 1e-10


def test_beamlet_plan_mixed_error_per_beamlet(plan):
    """Per-beamlet mixed vs FP64 error, and the weighted plan dose.

    Bar: the plan dose and at least 60 of the 64 beamlets within 1e-4, every
    beamlet within 1e-3.  The few beamlets above 1e-4 are conditioning, not
    arithmetic: the oracle itself, run with psi rounded to FP32 after every
    write, drifts to 9.6e-5 on beamlet 59, where the FP32 and FP64 runs first
    split at a truncation whose L1 tail sits within 0.27 % of its threshold
    (step 156); once the kept rank differs for a step the doses part by
    1e-4 to 5e-4."""
    announce("64-beamlet plan: mixed batch vs FP64 batch, per beamlet")
    f64, _ = solve_plan(plan, precision=_lib.PRECISION_FP64, rank_tol=MIXED_RANK_TOL)
    mix, _ = solve_plan(plan, precision=_lib.PRECISION_MIXED, rank_tol=MIXED_RANK_TOL)
    errs = np.array([rel(mix[:, b], f64[:, b]) for b in range(64)])
    worst = np.argsort(errs)[::-1][:4]
    w = np.random.default_rng(0).uniform(0.5, 1.5, 64)
    ep = rel(mix @ w, f64 @ w)
    print("[plan mixed] worst beamlets " + " ".join(f"{b}:{errs[b]:.2e}" for b in worst)
          + f"; median {np.median(errs):.2e}; above 1e-4: {int((errs > 1e-4).sum())}; "
          f"plan dose {ep:.2e}", flush=True)
    assert ep <= 1e-4
    assert (errs > 1e-4).sum() <= 4
    assert errs.max() <= 1e-3


# ---------------------------------------------------------------------------
# gradient configuration: one full-size beamlet
# ---------------------------------------------------------------------------
def test_full_size_beamlet_gradient_vs_oracle_adjoint():
    plan5 = beamlet_plan(n_beamlets=2, rank_range=(5, 30), integrator=INTEGRATOR_FIXED)
    P = plan5[1][1]
    steps = 32
    announce(f"gradient beamlet 128x128 P_15 r={P['r']}, {steps} steps, checkpoints every 16")
    rng = np.random.default_rng(7)
    Db = rng.standard_normal(P["C"])
    c = _lib.Context(P, INTEGRATOR_FIXED)
    try:
        c.set_checkpoints(16)
        c.solve(steps)
        d = c.state()["dose"]
        g = c.adjoint(Db)
    finally:
        c.close()
    t = time.perf_counter()
    dose_o, tape = oadj.forward(P, steps)
    g_o = oadj.adjoint(P, tape, Db)
    ed, eg = rel(d, dose_o), rel(g, g_o)
    print(f"[gradient] oracle {time.perf_counter() - t:.1f} s dose rel {ed:.3e} grad rel {eg:.3e}",
          flush=True)
    assert ed <= 1e-10
    assert eg <= 1e-6
