"""Edge sizes on the GPU against the FP64 oracle: the smallest rank, moment
set, grid and energy grid the library accepts, and the largest rank cap of
the rank-adaptive integrator (2 r = 96, the device Cholesky's limit).

Bar: dose relative L2 <= 1e-10 (FP64), ranks equal step by step.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import dlra, fullpn
from titus_amd import _lib
from titus_amd.problem import INTEGRATOR_BUG, INTEGRATOR_FIXED, build_problem
from titus_amd.settings import Settings, heterogeneous_phantom

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def announce(name, P):
    print(f"\n[case] {name}: C={P['C']} Nm={P['Nm']} Q={P['Q']} nE={P['nE']} r={P['r']}",
          flush=True)
    sys.stdout.flush()


def run(P, integrator, last_step=0):
    ctx = _lib.Context(P, integrator)
    try:
        ctx.solve(last_step)
        st = ctx.state()
        done, _ = ctx.progress()
        st["ranks"] = ctx.history(done)[0]
    finally:
        ctx.close()
    return st


def beam(n, N, r, steps=None, integrator=INTEGRATOR_FIXED):
    s = Settings(n + 1, n + 1, r, "waterBeamElectrons", "Electrons", a=0.0, b=3.0, c=0.0, d=3.0,
                 x0=1.5, y0=0.0, nPN=N, eMin=0.5, sigmaX=0.4, sigmaY=0.4, eMax=12.5,
                 sigmaO1Inv=2.0, sigmaO3Inv=2.0, nEnergies=None if steps is None else steps + 1)
    return build_problem(s, "subset", integrator)


@pytest.mark.parametrize("N", [1, 5])
def test_rank_one_fixed(N):
    P = beam(24, N, 1)
    announce(f"rank 1, P_{N}", P)
    g = run(P, INTEGRATOR_FIXED)
    o = dlra.solve_fixed(P)
    ed = rel(g["dose"], o["dose"])
    print(f"[rank 1 P_{N}] dose rel {ed:.3e}", flush=True)
    assert np.abs(o["dose"]).max() > 0 and ed <= TOL


def test_bug_smallest_cap():
    """Rank cap 2 (initial rank 1) on P_3: the augmented bases have 2 and 4
    columns; the rank moves between 1 and 2."""
    P = beam(24, 3, 2, integrator=INTEGRATOR_BUG)
    announce("BUG P_3, cap 2", P)
    g = run(P, INTEGRATOR_BUG)
    o = dlra.solve_bug(P)
    ed = rel(g["dose"], o["dose"])
    print(f"[BUG cap 2] dose rel {ed:.3e}", flush=True)
    assert np.array_equal(g["ranks"], np.asarray(o["ranks"]))
    assert ed <= TOL


@pytest.mark.parametrize("flux_order", [1, 2])
def test_smallest_grid_two_steps(flux_order):
    """6 x 6 cells, 4 x 4 interior (the limited flux acts on cells 2..n-3)
    and the shortest energy grid (three points, two steps)."""
    P = dict(beam(6, 3, 2, steps=2))
    P["flux_order"] = flux_order
    rng = np.random.default_rng(1)
    psi0 = np.zeros((P["C"], P["Q"]), order="F")
    psi0[P["interior"], :] = rng.uniform(0.0, 10.0, size=(len(P["interior"]), P["Q"]))
    P["psi0"] = psi0
    announce(f"6x6 cells, 2 steps, flux order {flux_order}", P)
    assert P["nE"] == 3
    g = run(P, INTEGRATOR_FIXED)
    o = dlra.solve_fixed(P)
    ed, ep = rel(g["dose"], o["dose"]), rel(g["psi"], o["psi"])
    print(f"[smallest grid] dose rel {ed:.3e} psi rel {ep:.3e}", flush=True)
    assert np.abs(o["dose"]).max() > 0 and ed <= TOL and ep <= 1e-12


def test_largest_rank_cap_bug():
    """Rank cap 48 with a tight truncation (theta = 1e-9): the rank reaches
    the cap after six steps, so the augmented bases have 96 columns, the
    device Cholesky's limit."""
    s = heterogeneous_phantom(n=24, N=13, seed=5, r=48, eKin=8.0)
    s.sigmaO1Inv = s.sigmaO3Inv = 2.0
    s.nEnergies = 13
    s.theta = 1e-9
    P = build_problem(s, "subset", INTEGRATOR_BUG)
    announce("BUG rank cap 48", P)
    assert P["Nm"] >= 96
    g = run(P, INTEGRATOR_BUG)
    o = dlra.solve_bug(P)
    ed = rel(g["dose"], o["dose"])
    print(f"[cap 48] ranks {list(g['ranks'])} dose rel {ed:.3e}", flush=True)
    assert np.array_equal(g["ranks"], np.asarray(o["ranks"]))
    assert max(g["ranks"]) == 48
    assert ed <= TOL


def test_rank_cap_above_limit_is_refused():
    s = heterogeneous_phantom(n=24, N=13, seed=5, r=49, eKin=8.0)
    s.nEnergies = 5
    P = build_problem(s, "subset", INTEGRATOR_BUG)
    with pytest.raises(_lib.TitusError, match="96"):
        _lib.Context(P, INTEGRATOR_BUG)


def test_full_solver_smallest_grid():
    P = cases.small_fcs(n=4, N=3, steps=2)
    announce("full-rank 4x4 cells, 2 steps", P)
    ctx = _lib.FcsContext(P)
    try:
        ctx.solve(0)
        g = ctx.state()
    finally:
        ctx.close()
    o = fullpn.solve_fcs(P)
    ed = rel(g["dose"], o["dose"])
    print(f"[fcs smallest] dose rel {ed:.3e}", flush=True)
    assert ed <= TOL and rel(g["u"], o["u"]) <= TOL


def test_nonfinite_state_is_reported():
    """A beam source scaled to 1e300 overflows the state: the solve stops with
    TITUS_E_NONFINITE instead of carrying inf/NaN on (every loop of the step
    is bounded, so the overflow only propagates)."""
    P = dict(beam(24, 3, 2, steps=6))
    P["bc_scale"] = np.asarray(P["bc_scale"]) * 1e300
    announce("overflowing beam source", P)
    ctx = _lib.Context(P, INTEGRATOR_FIXED)
    try:
        with pytest.raises(_lib.TitusError) as e:
            ctx.solve(0)
        assert e.value.code == -3, e.value
    finally:
        ctx.close()


def test_cfl_number_matches_the_reference_formula():
    P = beam(24, 3, 2, steps=4)
    ctx = _lib.Context(P, INTEGRATOR_FIXED)
    try:
        cfl = ctx.cfl()
    finally:
        ctx.close()
    ref = P["dE"] / min(P["dx"], P["dy"]) * np.max(1.0 / P["density"])
    assert abs(cfl - ref) <= 1e-14 * ref


def test_failed_adjoint_leaves_the_forward_state():
    """An adjoint whose reverse pass overflows (seed 1e308) fails with
    TITUS_E_NONFINITE and leaves dose, factors, psi and the step count as
    they were; a NaN seed is refused up front; a later adjoint on the same
    context equals one on a fresh context bit for bit."""
    P = beam(24, 5, 6, steps=14)
    announce("adjoint failure and restore", P)
    seed = np.random.default_rng(2).standard_normal(P["C"])
    ctx = _lib.Context(P, INTEGRATOR_FIXED)
    ref = _lib.Context(P, INTEGRATOR_FIXED)
    try:
        for c in (ctx, ref):
            c.set_checkpoints(4)
            c.solve(12)
        before = ctx.state()
        with pytest.raises(_lib.TitusError) as e:
            ctx.adjoint(np.full(P["C"], 1e308))
        assert e.value.code == -3, e.value
        after = ctx.state()
        for k in ("dose", "X", "S", "W", "psi"):
            assert np.array_equal(before[k], after[k]), k
        assert ctx.progress() == ref.progress()
        bad = seed.copy()
        bad[3] = np.nan
        with pytest.raises(_lib.TitusError) as e:
            ctx.adjoint(bad)
        assert e.value.code == -1
        g1, g2 = ctx.adjoint(seed), ref.adjoint(seed)
        print(f"[adjoint restore] |g| {np.linalg.norm(g2):.3e}", flush=True)
        assert np.array_equal(g1, g2)
        ctx.step(13)
        ref.step(13)
        assert np.array_equal(ctx.state()["dose"], ref.state()["dose"])
    finally:
        ctx.close()
        ref.close()
