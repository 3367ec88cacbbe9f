"""Second-order limited uncollided flux (solveFlux!, the sweep of
CSolveFirstCollisionSourceDLR) on the GPU against the FP64 oracle.

Bar: dose relative L2 <= 1e-10 and psi <= 1e-12 (FP64).  The boundary beam
does not cross the first interior ring under this flux (see
tests/test_limited_flux.py), so most cases start from a smooth initial flux
inside the grid.
"""
import sys

import numpy as np
import pytest

import cases
from oracle import dlra
from test_limited_flux import smooth_initial_flux
from titus_amd import _lib
from titus_amd.solver import MIXED_RANK_TOL

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def announce(name, P):
    print(f"\n[case] {name}: C={P['C']} Nm={P['Nm']} Q={P['Q']} nE={P['nE']} r={P['r']}",
          flush=True)
    sys.stdout.flush()


def run(P, integrator, last_step=0, precision=_lib.PRECISION_FP64, rank_tol=1e-7):
    ctx = _lib.Context(P, integrator, precision=precision, rank_tol=rank_tol)
    try:
        ctx.solve(last_step)
        st = ctx.state()
        st["plan"] = ctx.tuning("sweep")
        done, _ = ctx.progress()
        st["ranks"] = ctx.history(done)[0]
    finally:
        ctx.close()
    return st


def limited(P, blob=True, boundary=True):
    Q = dict(P)
    Q["flux_order"] = 2
    if blob:
        Q["psi0"] = smooth_initial_flux(Q)
    if not boundary:
        Q["bc_base"] = np.zeros_like(Q["bc_base"])
    return Q


@pytest.mark.parametrize("boundary", [True, False])
def test_fixed_limited_full(boundary):
    P = limited(cases.small_fixed(r=6), boundary=boundary)
    announce(f"fixed limited flux, boundary source {boundary}", P)
    g = run(P, _lib.INTEGRATOR_FIXED)
    o = dlra.solve_fixed(P)
    ed, ep = rel(g["dose"], o["dose"]), rel(g["psi"], o["psi"])
    print(f"[fixed limited] dose rel {ed:.3e}  psi rel {ep:.3e}", flush=True)
    assert ed <= 1e-10 and ep <= 1e-12


def test_bug_limited_heterogeneous():
    P = limited(cases.small_bug())
    announce("BUG limited flux, phantom", P)
    g = run(P, _lib.INTEGRATOR_BUG)
    o = dlra.solve_bug(P)
    ed = rel(g["dose"], o["dose"])
    print(f"[bug limited] dose rel {ed:.3e}", flush=True)
    assert np.array_equal(g["ranks"][:len(o["ranks"])], np.asarray(o["ranks"])[:len(g["ranks"])])
    assert ed <= 1e-10


def test_mixed_limited(monkeypatch):
    P = limited(cases.small_fixed(r=6))
    announce("mixed limited flux", P)
    monkeypatch.setattr(dlra, "RANK_TOL", MIXED_RANK_TOL)
    g = run(P, _lib.INTEGRATOR_FIXED, precision=_lib.PRECISION_MIXED, rank_tol=MIXED_RANK_TOL)
    o = dlra.solve_fixed(P)
    ed = rel(g["dose"], o["dose"])
    print(f"[mixed limited] dose rel vs FP64 oracle {ed:.3e}", flush=True)
    assert 0.0 < ed <= 1e-4


def test_tuned_limited_sweep():
    """A grid large enough to tune the limited sweep's launch plan."""
    from titus_amd.problem import INTEGRATOR_FIXED, build_problem
    from titus_amd.settings import Settings
    s = Settings(161, 161, 8, "waterBeamElectrons", "Electrons", a=0.0, b=8.0, c=0.0, d=8.0,
                 x0=4.0, y0=0.0, nPN=5, eMin=1.0, eMax=10.5, sigmaX=1.0, sigmaY=1.0,
                 sigmaO1Inv=2.0, sigmaO3Inv=2.0)
    P = limited(build_problem(s, "subset", INTEGRATOR_FIXED))
    announce("tuned limited sweep", P)
    assert P["C"] * P["Q"] >= 2 ** 18
    g = run(P, INTEGRATOR_FIXED, last_step=6)
    print(f"[tuned limited] plan {g['plan']}", flush=True)
    assert g["plan"]["tuned"] and not g["plan"]["nontemporal"]
    o = dlra.solve_fixed(P, steps=6)
    assert rel(g["dose"], o["dose"]) <= 1e-10
    assert rel(g["psi"], o["psi"]) <= 1e-12


def test_host_mirror_and_limits():
    from titus_amd.settings import Settings
    from titus_amd.solver import CSolveFirstCollisionSourceDLR, SolverCSD
    s = Settings(25, 25, 6, "lung", "Electrons", nPN=5, nEnergies=40)
    obj = SolverCSD(s)
    out = CSolveFirstCollisionSourceDLR(obj)
    P = dict(obj.problem(_lib.INTEGRATOR_FIXED))
    P["flux_order"] = 2
    o = dlra.solve_fixed(P)
    print(f"[CSolve mirror] dose rel {rel(out[4], o['dose']):.3e}", flush=True)
    assert rel(out[4], o["dose"]) <= 1e-10
    with pytest.raises(ValueError, match="lung"):
        CSolveFirstCollisionSourceDLR(SolverCSD(cases_settings_water()))
    # the adjoint needs the upwind flux
    ctx = _lib.Context(limited(cases.small_fixed(r=6)), _lib.INTEGRATOR_FIXED)
    try:
        with pytest.raises(_lib.TitusError, match="upwind"):
            ctx.set_checkpoints(4)
    finally:
        ctx.close()


def cases_settings_water():
    from titus_amd.settings import Settings
    return Settings(25, 25, 6, "waterBeamElectrons", "Electrons", nPN=5, nEnergies=10)
