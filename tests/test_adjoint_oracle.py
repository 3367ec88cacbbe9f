"""The oracle's discrete adjoint against central finite differences and
against the forward-mode (dual-number) derivative of the same map.

``oracle.adjoint`` differentiates ``oracle.dlra.solve_fixed`` (the fixed-rank
loop of ``SolveFirstCollisionSourceDLR``, ``code/SolverCSD.jl:891-1120``) with
respect to the density.  There is no reference adjoint; the pin is the
forward map itself: <g, v> against (J(rho + h v) - J(rho - h v)) / 2h for
J = <dose_bar, dose>, on random directions v scaled by rho.  The step h
(1e-7 |rho|) sits far below the distance to the forward map's rank
decisions and far above rounding; the central difference then agrees with
the exact derivative to about 1e-7, so the bar of 1e-5 catches a dropped
term, a wrong sign or a transposed operand anywhere in the reverse pass
(each one moves the result by more than 1e-3).
"""
import math

import numpy as np
import pytest

from oracle import adjoint, dlra, tangent
from titus_amd.problem import INTEGRATOR_FIXED, build_problem
from titus_amd.settings import Settings, ellipse_phantom


def _beam(n, N, r, angle_deg=5.0, seed=0):
    rho = ellipse_phantom(n, n, 0.0, 10.0, 0.0, 10.0, seed)
    ang = math.radians(angle_deg)
    s = Settings(n + 1, n + 1, r, "waterBeamElectrons", "Electrons", a=0.0, b=10.0, c=0.0, d=10.0,
                 x0=4.0, y0=0.0, Omega1=math.sin(ang), Omega3=math.cos(ang), nPN=N, eMin=0.1,
                 eMax=12.5, sigmaX=0.25, sigmaY=0.25, sigmaO1Inv=1000.0, sigmaO3Inv=1000.0)
    s.set_density(rho)
    return build_problem(s, "subset", INTEGRATOR_FIXED)


def _fd_check(P, steps, ndir, seed, tol):
    rng = np.random.default_rng(seed)
    rho0 = np.array(P["density"])
    dose, tape = adjoint.forward(P, steps)
    ref = dlra.solve_fixed(P, steps)["dose"]
    assert np.array_equal(dose, ref)
    Db = rng.standard_normal(P["C"])
    g = adjoint.adjoint(P, tape, Db)

    def J(rho):
        return float(Db @ dlra.solve_fixed(dict(P, density=rho), steps)["dose"])

    h = 1e-7 * np.linalg.norm(rho0)
    errs = []
    for _ in range(ndir):
        v = rng.standard_normal(P["C"]) * rho0
        v /= np.linalg.norm(v)
        fd = (J(rho0 + h * v) - J(rho0 - h * v)) / (2 * h)
        errs.append(abs(fd - g @ v) / max(abs(fd), 1e-3 * np.linalg.norm(g)))
    assert max(errs) < tol, errs
    return g


def test_adjoint_fd_32_directions():
    """32 random directions on a 24 x 24, P_5, rank-6 beam over 30 steps."""
    P = _beam(23, 5, 6)
    _fd_check(P, 30, 32, 1, 1e-5)


def test_adjoint_fd_40x40_p7():
    """40 x 40 cells, P_7, rank 20, the whole energy grid: the kept and dropped
    singular values sit within a factor 3 of each other at the rank cut."""
    P = _beam(39, 7, 20)
    _fd_check(P, None, 3, 2, 1e-5)


def test_adjoint_matches_forward_mode_40x40_p7():
    """Dot-product test against the forward-mode (dual-number) derivative of
    the same map, 40 x 40 cells, P_7, rank 20, the whole energy grid:
    <dose_bar, d dose[v]> = <grad, v> with no difference step involved."""
    P = _beam(39, 7, 20)
    _, tape = adjoint.forward(P)
    Lh, Fh = adjoint.stencil_matrices(P), adjoint.flux_matrices(P)
    rng = np.random.default_rng(6)
    Db = rng.standard_normal(P["C"])
    g = adjoint.adjoint(P, tape, Db, Lh, Fh)
    rho0 = np.asarray(P["density"])
    for _ in range(3):
        v = rng.standard_normal(P["C"]) * rho0
        dd = tangent.tangent(P, tape, v, Lh, Fh)
        assert abs(Db @ dd - g @ v) <= 1e-9 * abs(g @ v)


def test_forward_mode_matches_differences():
    """The forward-mode derivative of the whole dose vector against central
    differences (24 x 24, P_5, rank 6, 30 steps)."""
    P = _beam(23, 5, 6)
    _, tape = adjoint.forward(P, 30)
    rng = np.random.default_rng(7)
    rho0 = np.asarray(P["density"])
    for _ in range(3):
        v = rng.standard_normal(P["C"]) * rho0
        dd = tangent.tangent(P, tape, v)
        h = 1e-7 * np.linalg.norm(rho0) / np.linalg.norm(v)
        fd = (dlra.solve_fixed(dict(P, density=rho0 + h * v), 30)["dose"]
              - dlra.solve_fixed(dict(P, density=rho0 - h * v), 30)["dose"]) / (2 * h)
        assert np.linalg.norm(dd - fd) <= 1e-5 * np.linalg.norm(fd)


def test_adjoint_is_linear_in_the_seed():
    P = _beam(15, 3, 4)
    _, tape = adjoint.forward(P, 12)
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(P["C"]), rng.standard_normal(P["C"])
    ga, gb = adjoint.adjoint(P, tape, a), adjoint.adjoint(P, tape, b)
    gab = adjoint.adjoint(P, tape, 2.0 * a - b)
    assert np.allclose(gab, 2.0 * ga - gb, rtol=1e-10, atol=1e-12 * np.abs(gab).max())


def test_orth_bar_matches_finite_differences():
    """The basis adjoint alone, on a rank-deficient K whose cut and
    completion are active: <Kbar, dK> against the central difference of a
    function of the projector X1 X1^T (invariant under X1 -> X1 Q)."""
    rng = np.random.default_rng(4)
    m, k, r = 30, 8, 8
    sv = np.array([3.0, 1.0, 0.5, 2e-2, 1e-3, 6e-7, 1.5e-7, 0.0])   # cut at 3e-7
    U = np.linalg.qr(rng.standard_normal((m, k)))[0]
    V = np.linalg.qr(rng.standard_normal((k, k)))[0]
    K = (U * sv) @ V.T
    cands = list(range(m))
    X1, rec = adjoint.orth(K, r, cands, False)
    B = rng.standard_normal((m, r))
    def f(Kp):
        Xp = dlra.range_basis(Kp, r, cands)
        return float(np.sum((Xp @ Xp.T) * (B @ B.T)))
    X1bar = 2.0 * (B @ B.T) @ X1
    Kb = adjoint.orth_bar(rec, X1bar)
    dK = rng.standard_normal((m, k))
    # the kept 6e-7 and dropped 1.5e-7 pair makes the second derivative
    # large (1 / (sigma_i^2 - sigma_j^2)), so the difference step is small
    h = 1e-10
    fd = (f(K + h * dK) - f(K - h * dK)) / (2 * h)
    assert abs(fd - np.sum(Kb * dK)) < 1e-5 * abs(fd)


@pytest.mark.parametrize("q", [0, 3])
def test_flux_matrices_match_the_sweep(q):
    P = _beam(11, 3, 2)
    F = adjoint.flux_matrices(P)
    rng = np.random.default_rng(q)
    psi = rng.standard_normal((P["C"], P["Q"]))
    ref = dlra.upwind_flux(psi, np.ones(P["C"]), np.asarray(P["omega"]), P["nx"], P["ny"],
                           P["dx"], P["dy"])
    qq = q % P["Q"]
    assert np.allclose(F[qq] @ psi[:, qq], ref[:, qq], rtol=1e-14, atol=1e-12)
