"""FP64 GPU path (libtitus_hip.so via the C-ABI) against the FP64 oracle.

Bar: dose relative L2 <= 1e-10 (the plain form's tolerance); the low-rank
solution u = X S W^T and the uncollided flux psi are compared too.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import dlra
from titus_amd import _lib
from titus_amd.solver import MIXED_RANK_TOL

pytestmark = pytest.mark.gpu

DOSE_TOL = 1e-10


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def run_gpu(P, integrator, last_step=0, precision=_lib.PRECISION_FP64, rank_tol=1e-7):
    ctx = _lib.Context(P, integrator, precision=precision, rank_tol=rank_tol)
    try:
        ctx.solve(last_step)
        st = ctx.state()
        done, rank = ctx.progress()
        ranks, etas = ctx.history(done)
    finally:
        ctx.close()
    st.update(done=done, rank=rank, ranks=ranks, etas=etas)
    return st


def announce(name, P):
    print(f"\n[case] {name}: C={P['C']} Nm={P['Nm']} Q={P['Q']} nE={P['nE']} r={P['r']}",
          flush=True)
    sys.stdout.flush()


def check_close(g, o, label):
    ed = rel(g["dose"], o["dose"])
    ug = g["X"] @ g["S"] @ g["W"].T
    uo = o["X"] @ o["S"] @ o["W"].T
    eu = rel(ug, uo)
    ep = rel(g["psi"], o["psi"])
    print(f"[{label}] dose rel {ed:.3e}  u rel {eu:.3e}  psi rel {ep:.3e}", flush=True)
    assert np.all(np.isfinite(g["dose"]))
    assert ep <= 1e-13
    assert ed <= DOSE_TOL
    assert eu <= 1e-8


@pytest.mark.parametrize("steps", [1, 2, 5])
def test_fixed_first_steps(steps):
    P = cases.small_fixed()
    announce(f"small_fixed steps={steps}", P)
    g = run_gpu(P, _lib.INTEGRATOR_FIXED, steps)
    o = dlra.solve_fixed(P, steps=steps)
    assert g["done"] == steps
    check_close(g, o, f"fixed {steps}")


def test_fixed_small_full():
    P = cases.small_fixed()
    announce("small_fixed full", P)
    g = run_gpu(P, _lib.INTEGRATOR_FIXED)
    o = dlra.solve_fixed(P)
    check_close(g, o, "fixed full")


def test_fixed_heterogeneous():
    from titus_amd.settings import ellipse_phantom
    rho = ellipse_phantom(24, 24, 0.0, 3.0, 0.0, 3.0, seed=5, rho_lung=0.5)
    P = cases.small_fixed(density=rho)
    announce("small_fixed ellipse phantom", P)
    g = run_gpu(P, _lib.INTEGRATOR_FIXED)
    o = dlra.solve_fixed(P)
    check_close(g, o, "fixed hetero")


def test_fixed_rough_density_conditioning_limited():
    """Cell-wise random density: the fixed-rank solve itself moves by ~1e-9
    under 1e-14 input noise, so the GPU is held to the oracle's own spread:
    within 10x the largest dose change of three 1e-14 density perturbations."""
    rng = np.random.default_rng(7)
    rho = rng.uniform(0.6, 1.8, size=(24, 24))
    P = cases.small_fixed(density=rho)
    announce("small_fixed random density", P)
    g = run_gpu(P, _lib.INTEGRATOR_FIXED)
    o = dlra.solve_fixed(P)
    spread = 0.0
    for k in range(3):
        P2 = dict(P)
        P2["density"] = P["density"] * (1 + 1e-14 * rng.standard_normal(P["density"].shape))
        spread = max(spread, rel(dlra.solve_fixed(P2)["dose"], o["dose"]))
    ed = rel(g["dose"], o["dose"])
    print(f"[rough] dose rel {ed:.3e}, oracle spread under 1e-14 noise {spread:.3e}", flush=True)
    assert ed <= max(DOSE_TOL, 10 * spread)


def test_bug_small_full():
    P = cases.small_bug()
    announce("small_bug full", P)
    g = run_gpu(P, _lib.INTEGRATOR_BUG)
    o = dlra.solve_bug(P)
    assert list(g["ranks"]) == list(o["ranks"])
    check_close(g, o, "bug full")
    np.testing.assert_allclose(g["etas"][: len(o["etas"])], o["etas"], rtol=1e-6, atol=1e-12)


def test_parallel_small_full():
    """The parallel integrator (SolveFirstCollisionSourceDLRParallel) over the
    whole grid: dose, factors, psi and the rank after every step."""
    P = cases.small_parallel()
    announce("small_parallel full", P)
    g = run_gpu(P, _lib.INTEGRATOR_PARALLEL)
    o = dlra.solve_parallel(P)
    assert g["done"] == P["nE"] - 1
    assert list(g["ranks"]) == [int(v) for v in o["rvec"][1, 1:]]
    check_close(g, o, "parallel full")


def test_parallel_narrow_beam_conditioning_limited():
    """Narrow beam (4 ordinates), 48 x 48, P_9: here the parallel integrator
    amplifies rounding by ~1e9 (the oracle's own dose moves by 1e-6..1e-5
    under 1e-14 density noise within 35 steps, and its truncation ranks flip
    from step 39 on), so the bar is the oracle's spread: identical ranks over
    the first 35 steps and a dose within 10x the spread."""
    P = cases.small_parallel(n=48, N=9, cap=16, seed=1, angular=1e4)
    announce("parallel narrow beam 48x48 P9, 35 steps", P)
    steps = 35
    g = run_gpu(P, _lib.INTEGRATOR_PARALLEL, last_step=steps)
    o = dlra.solve_parallel(P, steps)
    assert list(g["ranks"]) == [int(v) for v in o["rvec"][1, 1:steps + 1]]
    rng = np.random.default_rng(0)
    spread = 0.0
    for _ in range(3):
        P2 = dict(P, density=np.asarray(P["density"]) * (1 + 1e-14 * rng.standard_normal(P["C"])))
        spread = max(spread, rel(dlra.solve_parallel(P2, steps)["dose"], o["dose"]))
    ed = rel(g["dose"], o["dose"])
    print(f"[parallel narrow] dose rel {ed:.3e}, oracle spread under 1e-14 noise {spread:.3e}",
          flush=True)
    assert ed <= max(DOSE_TOL, 10 * spread)


def test_slab_beam_full():
    P = cases.slab_beam()
    announce("slab_beam (80x80, P9, r=20) full", P)
    g = run_gpu(P, _lib.INTEGRATOR_FIXED)
    o = dlra.solve_fixed(P)
    check_close(g, o, "slab beam")


def test_bug_narrow_beam_phantom():
    P = cases.small_bug(n=48, N=9, cap=16, seed=1, angular=1e4)
    announce("bug narrow beam 48x48 P9", P)
    g = run_gpu(P, _lib.INTEGRATOR_BUG)
    o = dlra.solve_bug(P)
    assert list(g["ranks"]) == list(o["ranks"])
    check_close(g, o, "bug narrow")


MIXED_TOL = 1e-4   # mixed form: psi stored and swept in FP32, the rest FP64
# Both integrators keep near-null basis directions (the fixed-rank one all r,
# the rank-adaptive one at least two), completed canonically below the rank
# cut; the FP64 dose moves by 2e-3 (slab, fixed rank) and 1.4e-2 (narrow beam,
# BUG) when the cut goes from 1e-7 to 1e-4.  FP32 arithmetic perturbs singular
# values by ~1e-7 relative, so mixed solves use the cut MIXED_RANK_TOL and are
# compared with the FP64 oracle at that same cut.


@pytest.mark.parametrize("name", ["slab_beam", "small_bug", "bug_narrow", "phantom"])
def test_mixed_form_dose_error(name, monkeypatch):
    P = {"slab_beam": cases.slab_beam,
         "small_bug": cases.small_bug,
         "bug_narrow": lambda: cases.small_bug(n=48, N=9, cap=16, seed=1, angular=1e4),
         "phantom": lambda: cases.phantom_bug(n=48, N=9, cap=16, seed=2)}[name]()
    announce(f"mixed {name}", P)
    integ = _lib.INTEGRATOR_FIXED if name == "slab_beam" else _lib.INTEGRATOR_BUG
    tol = MIXED_RANK_TOL
    monkeypatch.setattr(dlra, "RANK_TOL", tol)
    g = run_gpu(P, integ, precision=_lib.PRECISION_MIXED, rank_tol=tol)
    o = dlra.solve_fixed(P) if integ == _lib.INTEGRATOR_FIXED else dlra.solve_bug(P)
    ed = rel(g["dose"], o["dose"])
    print(f"[mixed {name}] dose rel vs FP64 oracle {ed:.3e}", flush=True)
    assert np.all(np.isfinite(g["dose"]))
    assert ed <= MIXED_TOL
    assert ed > 0.0          # FP32 arithmetic is really in use


# Streaming stencil, psi and Galerkin products in FP32 as well.  Within the bar
# on these two cases.  On the narrow beam and the heterogeneous phantom the
# rank-adaptive solve amplifies perturbations: the oracle moves by 3e-3 under
# 6e-8 relative noise on psi and by 7e-4 with an FP32 streaming stencil, and
# by 1.4e-7 even under 1e-12 noise -- so this form is not the default.
@pytest.mark.parametrize("name", ["slab_beam", "small_bug"])
def test_mixed_full_fp32_dose_error(name, monkeypatch):
    P = {"slab_beam": cases.slab_beam, "small_bug": cases.small_bug}[name]()
    announce(f"mixed full fp32 {name}", P)
    integ = _lib.INTEGRATOR_FIXED if name == "slab_beam" else _lib.INTEGRATOR_BUG
    monkeypatch.setattr(dlra, "RANK_TOL", MIXED_RANK_TOL)
    g = run_gpu(P, integ, precision=_lib.PRECISION_MIXED_FULL_FP32, rank_tol=MIXED_RANK_TOL)
    o = dlra.solve_fixed(P) if integ == _lib.INTEGRATOR_FIXED else dlra.solve_bug(P)
    ed = rel(g["dose"], o["dose"])
    print(f"[mixed full fp32 {name}] dose rel vs FP64 oracle {ed:.3e}", flush=True)
    assert ed <= MIXED_TOL


def test_tuned_sweep_plan():
    """A context large enough to tune its sweep plan still matches the oracle."""
    from titus_amd.problem import INTEGRATOR_FIXED, build_problem
    from titus_amd.settings import Settings
    s = Settings(161, 161, 8, "waterBeamElectrons", "Electrons", a=0.0, b=8.0, c=0.0, d=8.0,
                 x0=4.0, y0=0.0, nPN=5, eMin=1.0, eMax=10.5, sigmaX=1.0, sigmaY=1.0,
                 sigmaO1Inv=2.0, sigmaO3Inv=2.0)
    P = build_problem(s, "subset", INTEGRATOR_FIXED)
    announce("tuned sweep", P)
    assert P["C"] * P["Q"] >= 2 ** 18
    ctx = _lib.Context(P, _lib.INTEGRATOR_FIXED)
    try:
        plan = ctx.tuning("sweep")
        print(f"[tuned sweep] plan {plan}", flush=True)
        assert plan["tuned"] and plan["rows"] in (2, 4, 8, 16) and 1 <= plan["groups"] <= 8
        ctx.solve(4)
        g = ctx.state()
    finally:
        ctx.close()
    o = dlra.solve_fixed(P, steps=4)
    assert rel(g["dose"], o["dose"]) <= 1e-10
    assert rel(g["psi"], o["psi"]) <= 1e-12


def test_step_order_errors():
    P = cases.small_fixed()
    ctx = _lib.Context(P, _lib.INTEGRATOR_FIXED)
    try:
        with pytest.raises(_lib.TitusError) as e:
            ctx.step(2)
        assert e.value.code == -5
        ctx.step(1)
        ctx.step(2)
        assert ctx.progress()[0] == 2
    finally:
        ctx.close()
