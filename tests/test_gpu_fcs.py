"""Full-rank P_N solver (SolveFirstCollisionSource) on the GPU against the
FP64 oracle (oracle/fullpn.py), and the initial uncollided flux of the
"waterBeam" problem on the DLR path.

Bar: dose, u and psi relative L2 <= 1e-10 (FP64; the collided kernel sums
the streaming terms in a different order than the oracle's products).
"""
import sys

import numpy as np
import pytest

import cases
from oracle import dlra, fullpn
from titus_amd import _lib
from titus_amd.problem import INTEGRATOR_FIXED, build_problem
from titus_amd.settings import Settings

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def announce(name, P):
    print(f"\n[case] {name}: C={P['C']} Nm={P['Nm']} Q={P['Q']} nE={P['nE']}", flush=True)
    sys.stdout.flush()


def run_fcs(P, last_step=0):
    ctx = _lib.FcsContext(P)
    try:
        ctx.solve(last_step)
        st = ctx.state()
        st["plan"] = ctx.plan()
    finally:
        ctx.close()
    return st


def check(g, o, label):
    ed, eu, ep = rel(g["dose"], o["dose"]), rel(g["u"], o["u"]), rel(g["psi"], o["psi"])
    print(f"[{label}] plan {g['plan']}  dose rel {ed:.3e}  u rel {eu:.3e}  psi rel {ep:.3e}",
          flush=True)
    assert np.all(np.isfinite(g["dose"])) and np.abs(o["dose"]).max() > 0
    assert ed <= TOL and eu <= TOL and ep <= TOL


@pytest.mark.parametrize("steps", [1, 3, 0])
def test_fcs_small(steps):
    P = cases.small_fcs()
    announce(f"fcs small, {steps or 'all'} steps", P)
    g = run_fcs(P, steps)
    o = fullpn.solve_fcs(P, steps=steps or None)
    check(g, o, "fcs small")


def test_fcs_phantom_p9():
    P = cases.phantom_fcs(n=48, N=9, seed=3)
    announce("fcs phantom 48x48 P9", P)
    check(run_fcs(P), fullpn.solve_fcs(P), "fcs phantom")


def test_fcs_p21_several_groups():
    """P_21 (253 moments): 16 column blocks, several block groups; 20 steps."""
    P = cases.small_fcs(n=24, N=21)
    announce("fcs 24x24 P21", P)
    g = run_fcs(P, 20)
    assert g["plan"]["nonzero_blocks"] < g["plan"]["dense_blocks"]
    check(g, fullpn.solve_fcs(P, steps=20), "fcs P21")


def test_fcs_tuned_plan():
    """A grid large enough (cells x moments >= 2^20) for the creation-time
    tuner: whichever kernel plan it picks is checked against the oracle."""
    P = cases.phantom_fcs(n=72, N=21, seed=2)
    announce("fcs phantom 72x72 P21, tuned", P)
    assert P["C"] * P["Nm"] >= 2 ** 20
    g = run_fcs(P, 6)
    assert g["plan"]["tuned"]
    check(g, fullpn.solve_fcs(P, steps=6), "fcs tuned")


def test_fcs_full_moment_set():
    """(N+1)^2 moments: sine harmonics as separate blocks of the plan."""
    P = cases.small_fcs(n=16, N=5, moments="full")
    announce("fcs full moments 16x16 P5", P)
    check(run_fcs(P), fullpn.solve_fcs(P), "fcs full moments")


def test_fcs_host_mirror_and_order():
    from titus_amd.solver import SolveFirstCollisionSource, SolverCSD
    P = cases.small_fcs()
    announce("fcs host mirror", P)
    s = Settings(25, 25, 2, "waterBeamElectrons", "Electrons", a=0.0, b=3.0, c=0.0, d=3.0,
                 x0=1.5, y0=0.0, nPN=5, eMin=0.5, sigmaX=0.4, sigmaY=0.4, eMax=12.5,
                 sigmaO1Inv=2.0, sigmaO3Inv=2.0)
    obj = SolverCSD(s)
    u, dose, psi = SolveFirstCollisionSource(obj)
    o = fullpn.solve_fcs(P)
    assert rel(u, o["u_out"]) <= TOL and rel(dose, o["dose"]) <= TOL
    assert psi.shape == (24, 24, P["Q"])
    ctx = _lib.FcsContext(P)
    try:
        ctx.step(1)
        with pytest.raises(_lib.TitusError, match="TITUS_E_STATE"):
            ctx.step(3)
    finally:
        ctx.close()


def test_initial_flux_dlr_and_fcs():
    """An initial uncollided flux (SetupIC; titus_problem_v1 version 3 and
    titus_fcs_problem_v1) enters both solvers as the oracle takes it."""
    rng = np.random.default_rng(7)
    P = dict(cases.small_fixed(r=6))
    psi0 = np.zeros((P["C"], P["Q"]), order="F")
    psi0[P["interior"], :] = rng.uniform(0.0, 50.0, size=(len(P["interior"]), P["Q"]))
    P["psi0"] = psi0
    announce("DLR fixed with initial flux", P)
    ctx = _lib.Context(P, _lib.INTEGRATOR_FIXED)
    try:
        ctx.solve(12)
        g = ctx.state()
    finally:
        ctx.close()
    o = dlra.solve_fixed(P, steps=12)
    o0 = dlra.solve_fixed(cases.small_fixed(r=6), steps=12)
    ed = rel(g["dose"], o["dose"])
    print(f"[DLR initial flux] dose rel {ed:.3e} (change by the flux {rel(o['dose'], o0['dose']):.2e})",
          flush=True)
    assert rel(o["dose"], o0["dose"]) > 1e-3
    assert ed <= TOL
    F = dict(cases.small_fcs())
    F["psi0"] = psi0
    announce("full solver with initial flux", F)
    check(run_fcs(F, 12), fullpn.solve_fcs(F, steps=12), "fcs initial flux")


def test_waterbeam_builds_initial_flux():
    s = Settings(25, 25, 6, "waterBeam", "Protons", nPN=5, sigmaO1Inv=2.0, sigmaO3Inv=2.0)
    P = build_problem(s, "subset", INTEGRATOR_FIXED)
    assert P["psi0"] is not None and P["psi0"].shape == (P["C"], P["Q"])
    announce("waterBeam protons DLR", P)
    ctx = _lib.Context(P, _lib.INTEGRATOR_FIXED)
    try:
        ctx.solve(20)
        g = ctx.state()
    finally:
        ctx.close()
    ed = rel(g["dose"], dlra.solve_fixed(P, steps=20)["dose"])
    print(f"[waterBeam DLR] dose rel {ed:.3e}", flush=True)
    assert ed <= TOL
