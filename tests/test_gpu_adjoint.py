"""Device adjoint (titus_adjoint) against the oracle adjoint and finite differences.

The differentiated map is the fixed-rank solve (``SolveFirstCollisionSourceDLR``,
``code/SolverCSD.jl:891-1120``) as a function of the density.  No reference
adjoint exists; the pins are the oracle's reverse pass and its forward-mode
(dual-number) derivative (both pinned by central differences in
``tests/test_adjoint_oracle.py``) and central differences of the device
forward map on 32 random directions.
"""
import math

import numpy as np
import pytest

from oracle import adjoint as oadj
from oracle import tangent as otan
from titus_amd import _lib
from titus_amd.adjoint import piecewise_target, plan_objective_grad
from titus_amd.beamlets import beamlet_plan
from titus_amd.problem import INTEGRATOR_BUG, INTEGRATOR_FIXED, build_problem
from titus_amd.settings import Settings, ellipse_phantom

pytestmark = pytest.mark.gpu


def _beam(n, N, r, angle_deg=5.0, seed=0, integrator=INTEGRATOR_FIXED):
    rho = ellipse_phantom(n, n, 0.0, 10.0, 0.0, 10.0, seed)
    ang = math.radians(angle_deg)
    s = Settings(n + 1, n + 1, r, "waterBeamElectrons", "Electrons", a=0.0, b=10.0, c=0.0, d=10.0,
                 x0=4.0, y0=0.0, Omega1=math.sin(ang), Omega3=math.cos(ang), nPN=N, eMin=0.1,
                 eMax=12.5, sigmaX=0.25, sigmaY=0.25, sigmaO1Inv=1000.0, sigmaO3Inv=1000.0)
    s.set_density(rho)
    return build_problem(s, "subset", integrator)


def _device_grad(P, steps, Db, every):
    c = _lib.Context(P, INTEGRATOR_FIXED)
    try:
        c.set_checkpoints(every)
        c.solve(steps)
        d0 = c.state()["dose"]
        g = c.adjoint(Db)
        d1 = c.state()["dose"]
    finally:
        c.close()
    return d0, d1, g


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_adjoint_matches_oracle_40x40():
    """40 x 40 cells, P_7, rank 20, whole energy grid, checkpoints every 16."""
    P = _beam(39, 7, 20)
    rng = np.random.default_rng(0)
    Db = rng.standard_normal(P["C"])
    dose_o, tape = oadj.forward(P)
    g_o = oadj.adjoint(P, tape, Db)
    d0, d1, g = _device_grad(P, 0, Db, 16)
    print("dose rel", _rel(d0, dose_o), "grad rel", _rel(g, g_o))
    assert _rel(d0, dose_o) < 1e-10
    assert np.array_equal(d0, d1)                 # forward state restored
    assert _rel(g, g_o) < 1e-6


def test_adjoint_matches_forward_mode_40x40_p7():
    """The device gradient against the oracle's forward-mode (dual-number)
    derivative: <grad, v> = <dose_bar, d dose[v]> on three directions."""
    P = _beam(39, 7, 20)
    rng = np.random.default_rng(3)
    Db = rng.standard_normal(P["C"])
    _, tape = oadj.forward(P)
    _, _, g = _device_grad(P, 0, Db, 16)
    rho0 = np.asarray(P["density"])
    Lh, Fh = oadj.stencil_matrices(P), oadj.flux_matrices(P)
    for _ in range(3):
        v = rng.standard_normal(P["C"]) * rho0
        dd = otan.tangent(P, tape, v, Lh, Fh)
        assert abs(Db @ dd - g @ v) <= 1e-7 * abs(Db @ dd)


def test_checkpoint_interval_does_not_change_the_gradient():
    P = _beam(23, 5, 6)
    Db = np.random.default_rng(1).standard_normal(P["C"])
    grads = [_device_grad(P, 40, Db, every)[2] for every in (1, 7, 16, 1000)]
    for g in grads[1:]:
        assert np.array_equal(g, grads[0])


def test_adjoint_central_differences_32_directions():
    """<g, v> against (J(rho + h v) - J(rho - h v)) / 2h of the device solve."""
    P = _beam(23, 5, 6)
    steps = 30
    rng = np.random.default_rng(2)
    Db = rng.standard_normal(P["C"])
    _, _, g = _device_grad(P, steps, Db, 16)
    rho0 = np.asarray(P["density"])

    def J(rho):
        c = _lib.Context(dict(P, density=rho), INTEGRATOR_FIXED)
        try:
            c.solve(steps)
            return float(Db @ c.state()["dose"])
        finally:
            c.close()

    h = 1e-7 * np.linalg.norm(rho0)
    errs = []
    for _ in range(32):
        v = rng.standard_normal(P["C"]) * rho0
        v /= np.linalg.norm(v)
        fd = (J(rho0 + h * v) - J(rho0 - h * v)) / (2 * h)
        errs.append(abs(fd - g @ v) / max(abs(fd), 1e-3 * np.linalg.norm(g)))
    print("max rel error over 32 directions", max(errs))
    assert max(errs) < 1e-5


def test_plan_objective_gradients():
    """Four fixed-rank beamlets on a 40 x 40 phantom: dJ/dw and dJ/ds against
    central differences of the plan objective."""
    plan = beamlet_plan(n_beamlets=4, n=39, N=7, seed=3, rank_range=(4, 10),
                        integrator=INTEGRATOR_FIXED)
    problems = [P for _, P in plan]
    C = problems[0]["C"]
    rng = np.random.default_rng(4)
    w = rng.uniform(0.5, 1.5, len(problems))
    s = np.ones(C)
    # prescription at the plan dose's scale
    J0, wbar, sbar = plan_objective_grad(problems, w, s, np.zeros(C))
    target = piecewise_target(C, problems[0]["ny"], seed=5, scale=2.0 * math.sqrt(J0 / C))
    J, wbar, sbar = plan_objective_grad(problems, w, s, target)

    def Jf(ww, ss):
        return plan_objective_grad(problems, ww, ss, target)[0]

    hw = 1e-6
    for b in range(len(problems)):
        e = np.zeros(len(problems))
        e[b] = 1.0
        fd = (Jf(w + hw * e, s) - Jf(w - hw * e, s)) / (2 * hw)
        assert abs(fd - wbar[b]) < 1e-6 * max(abs(fd), 1e-3 * np.abs(wbar).max()), (b, fd, wbar[b])
    hs = 1e-7 * math.sqrt(C)
    for _ in range(2):
        v = rng.standard_normal(C)
        v /= np.linalg.norm(v)
        fd = (Jf(w, s + hs * v) - Jf(w, s - hs * v)) / (2 * hs)
        assert abs(fd - sbar @ v) < 1e-5 * max(abs(fd), 1e-3 * np.linalg.norm(sbar)), (fd, sbar @ v)


def test_adjoint_errors():
    P = _beam(15, 3, 4)
    c = _lib.Context(P, INTEGRATOR_FIXED)
    try:
        with pytest.raises(_lib.TitusError):
            c.adjoint(np.zeros(P["C"]))           # no checkpoints
        c.step(1)
        with pytest.raises(_lib.TitusError):
            c.set_checkpoints(4)                  # after the first step
    finally:
        c.close()
    c = _lib.Context(_beam(15, 3, 4, integrator=INTEGRATOR_BUG), INTEGRATOR_BUG)
    try:
        with pytest.raises(_lib.TitusError):
            c.set_checkpoints(4)                  # rank-adaptive: no adjoint
    finally:
        c.close()
