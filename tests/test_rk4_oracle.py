"""Pins of the oracle's restatement of the CUDA.jl RK4 variant
(``oracle.dlra.solve_rk4``, ``CudaSolveFirstCollisionSourceDLR4thOrder``,
``SolverCSD.jl:1422-1724``).

* Full-rank exactness: on a 6 x 6 grid with all 16 moments of P_3, rank 16
  (X spanning the interior cells, W square) and no scattering, one streaming
  step (step 2) equals the classical RK4 step of the dense central streaming
  equation u' = -L2x u Ax - L2y u Az, which for this linear equation is the
  Taylor polynomial u + dE F(u) + dE^2/2 F^2(u) + dE^3/6 F^3(u) + dE^4/24
  F^4(u), written with the stencil matrices assembled by the reference's
  loops.  A swapped operator pairing or the Roe terms left in fail it.
* The analytic flux: at tau = 0 it is the water beam's SetupIC profile on
  the interior, it moves with tau Omega, and entries whose beam factors fall
  below 1e-6 keep their old value (the reference's loop skips them).
"""
import numpy as np
import pytest

from oracle import dlra
from test_oracle import dense_stencils_reference_loops
from test_oracle_streaming import problem, rel, start
from titus_amd.problem import INTEGRATOR_FIXED


def taylor4(P, u):
    rho2d = np.asarray(P["density"]).reshape(P["nx"], P["ny"])
    _, _, L2x, L2y = dense_stencils_reference_loops(rho2d, P["dx"], P["dy"])

    def F(v):
        out = -L2x @ v @ P["Ax"] - L2y @ v @ P["Az"]
        out[P["bcells"], :] = 0.0
        return out
    dE = P["dE"]
    out, term = u.copy(), u.copy()
    for k in range(1, 5):
        term = dE / k * F(term)
        out = out + term
    return out


def rk4_pin():
    P = start(problem(INTEGRATOR_FIXED), 16)
    u0 = P["X0"] @ P["S0"] @ P["W0"].T
    out = dlra.solve_rk4(P, steps=2)
    return rel(out["X"] @ out["S"] @ out["W"].T, taylor4(P, u0))


def test_rk4_streaming_step_is_the_dense_rk4_step():
    assert rk4_pin() < 1e-12


def test_rk4_pin_fails_on_a_swapped_pairing(monkeypatch):
    orig = dlra._Ops.__init__

    def init(self, P):
        orig(self, P)
        self.Ax, self.Az = self.Az, self.Ax
    monkeypatch.setattr(dlra._Ops, "__init__", init)
    assert rk4_pin() > 1e-3


def test_rk4_step_differs_from_euler_and_from_the_roe_form():
    """The pin is sharp: an explicit-Euler step, or the step with the
    dissipative terms, is a different answer."""
    P = start(problem(INTEGRATOR_FIXED), 16)
    u0 = P["X0"] @ P["S0"] @ P["W0"].T
    out = dlra.solve_rk4(P, steps=2)
    u = out["X"] @ out["S"] @ out["W"].T
    euler = taylor4(dict(P, dE=P["dE"]), u0)
    rho2d = np.asarray(P["density"]).reshape(P["nx"], P["ny"])
    _, _, L2x, L2y = dense_stencils_reference_loops(rho2d, P["dx"], P["dy"])
    e1 = u0 - P["dE"] * (L2x @ u0 @ P["Ax"] + L2y @ u0 @ P["Az"])
    e1[P["bcells"], :] = 0.0
    assert rel(u, e1) > 1e-6 and rel(u, euler) < 1e-12
    full = dlra.solve_fixed(P, steps=2)
    assert rel(u, full["X"] @ full["S"] @ full["W"].T) > 1e-3


def test_analytic_flux_profile_motion_and_kept_entries():
    import math
    from test_psibeam import analytic, water_problem
    s, P = water_problem(60, Omega1=0.3, Omega3=math.sqrt(1 - 0.09))
    nx, ny, Q = P["nx"], P["ny"], P["Q"]
    X, Y = np.meshgrid(s.xMid, s.yMid, indexing="ij")
    # tau = 0 (eTrafo[0] = 0): the SetupIC profile on the interior
    psi = dlra.analytic_psi(P, np.zeros((P["C"], Q)), 0, 0.0)
    inner = np.zeros((nx, ny), bool)
    inner[2:nx - 2, 2:ny - 2] = True
    for q in range(Q):
        got = psi[:, q].reshape(nx, ny)
        ref = analytic(s, P["omega"][q], X, Y, 0.0)
        np.testing.assert_allclose(got[inner & (got > 0)], ref[inner & (got > 0)], rtol=1e-12)
    # a later step: the profile at tau, and a kept old value where the beam has left
    k = 1                                       # tau = eTrafo[1], the beam still inside
    tau = P["eTrafo"][k]
    marker = np.full((P["C"], Q), -7.0)
    psi = dlra.analytic_psi(P, marker.copy(), k, 0.3)
    assert (psi != -7.0).any() and (psi == -7.0).any()
    for q in range(Q):
        w = P["omega"][q]
        got = psi[:, q].reshape(nx, ny)
        ref = analytic(s, w, X, Y, tau, int_sigma=0.3)
        written = got != -7.0
        np.testing.assert_allclose(got[written], ref[written], rtol=1e-12)
        # written exactly where both beam factors reach 1e-6
        bx = dlra.normpdf(s.xMid - tau * w[0], s.x0, 0.1)
        by = dlra.normpdf(s.yMid - tau * w[2], s.y0, 0.1)
        assert np.array_equal(written, (bx[:, None] >= 1e-6) & (by[None, :] >= 1e-6))
