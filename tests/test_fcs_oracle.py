"""Pins of the full-rank P_N oracle (oracle/fullpn.py) and its host plumbing.

* ``solve_fcs`` -- against a cell-by-cell restatement of
  ``SolveFirstCollisionSource`` (``code/SolverCSD.jl:465-586``): dense stencil
  matrices assembled as ``stencils.jl`` does, the quadrant loops of
  ``solveFluxUpwind!``, the reference's loop over interior cells with its
  ``(i-1)*nx + j`` numbering and a dense solve of ``(I + dE D)`` per cell, and
  the dose line with its endpoint weights;
* the scattering step conserves u + M psi moment by moment when every
  sigma_l equals sigma_0 (D = 0) and streaming is off: the implicit update
  puts back into u exactly what the uncollided flux loses;
* ``setup_ic`` (SetupIC for "waterBeam") -- against the reference's loops.

The library's argument checks run here too (no device needed): non-square
grids and ghost cells other than the four edges are rejected before any GPU
call.
"""
import ctypes

import numpy as np
import pytest

import cases
from oracle import fullpn
from test_oracle import dense_stencils_reference_loops, upwind_loops
from titus_amd import physics as ph
from titus_amd.settings import Settings


def fcs_cell_loops(P, steps):
    """SolveFirstCollisionSource written the way the reference loops."""
    nx, ny, Q, Nm = P["nx"], P["ny"], P["Q"], P["Nm"]
    rho2d = P["density"].reshape(nx, ny)
    L1x, L1y, L2x, L2y = dense_stencils_reference_loops(rho2d, P["dx"], P["dy"])
    Ax, Az, aAx, aAz = P["Ax"], P["Az"], P["absAx"], P["absAz"]
    M, deg = P["M"], P["degree"]
    psi = np.zeros((nx, ny, Q)) if P.get("psi0") is None else P["psi0"].reshape(nx, ny, Q).copy()
    base = np.zeros((nx, ny, Q))
    for b, c in enumerate(P["bcells"]):
        base[c // ny, c % ny, :] = P["bc_base"][b, :]
    boundary = [(i - 1) * ny + j - 1 for i in range(1, nx + 1) for j in range(1, ny + 1)
                if i in (1, nx) or j in (1, ny)]
    u = np.zeros((nx * ny, Nm))
    uNew = u.copy()
    uOUnc = np.zeros(nx * ny)
    dose = np.zeros(nx * ny)
    nE = P["nE"]
    for n in range(2, steps + 2):
        sig = P["sigma"][n - 1]
        dE = P["eTrafo"][n - 1] - P["eTrafo"][n - 2]
        scale = P["fcs_bc_scale"][n - 2]
        for k in range(Q):
            for j in range(nx):
                psi[j, 0, k] = base[j, 0, k] * scale
                psi[j, ny - 1, k] = base[j, ny - 1, k] * scale
            for j in range(ny):
                psi[0, j, k] = base[0, j, k] * scale
                psi[nx - 1, j, k] = base[nx - 1, j, k] * scale
        Dvec = np.array([sig[deg[m]] for m in range(Nm)])
        D = np.diag(sig[0] - Dvec)
        flux = upwind_loops(psi, np.ones((nx, ny)), P["omega"], P["dx"], P["dy"])
        psi = psi - dE * flux
        psiNew = psi / (1 + dE * sig[0])
        uTilde = u - dE * L2x @ u @ Ax - dE * L2y @ u @ Az - dE * L1x @ u @ aAx - dE * L1y @ u @ aAz
        uTilde[boundary, :] = 0.0
        for i in range(2, nx):
            for j in range(2, ny):
                idx = (i - 1) * nx + j - 1
                uNew[idx, :] = np.linalg.solve(np.eye(Nm) + dE * D,
                                               uTilde[idx, :] + dE * np.diag(Dvec) @ M @ psiNew[i - 1, j - 1, :])
                uOUnc[idx] = psiNew[i - 1, j - 1, :] @ M[0, :]
        uNew[boundary, :] = 0.0
        dose += dE * (uNew[:, 0] + uOUnc) * P["SMid"][n - 2] / P["density"] / (1 + (n == 2 or n == nE))
        u = uNew.copy()
        u[boundary, :] = 0.0
        psi = psiNew.copy()
    return u, dose, psi.reshape(nx * ny, Q)


@pytest.mark.parametrize("steps_total", [None, 6])
def test_solve_fcs_matches_cell_loops(steps_total):
    rng = np.random.default_rng(4)
    n = 8
    P = cases.small_fcs(n=n, N=3, steps=steps_total if steps_total else 7,
                        density=rng.uniform(0.5, 1.8, size=(n, n)))
    steps = P["nE"] - 1 if steps_total is None else 4     # full grid: the last step weighs 1/2
    o = fullpn.solve_fcs(P, steps=steps)
    u, dose, psi = fcs_cell_loops(P, steps)
    assert np.abs(o["dose"]).max() > 0
    np.testing.assert_allclose(o["u"], u, rtol=1e-12, atol=1e-12 * np.abs(u).max())
    np.testing.assert_allclose(o["dose"], dose, rtol=1e-12, atol=1e-12 * np.abs(dose).max())
    np.testing.assert_allclose(o["psi"], psi, rtol=1e-12, atol=1e-12 * np.abs(psi).max())


def test_scattering_conserves_u_plus_M_psi():
    P = dict(cases.small_fcs(n=10, N=5, steps=12))
    rng = np.random.default_rng(5)
    for k in ("Ax", "Az", "absAx", "absAz"):
        P[k] = np.zeros_like(P[k])
    P["omega"] = np.zeros_like(P["omega"])            # no uncollided streaming
    P["bc_base"] = np.zeros_like(P["bc_base"])
    P["psi0"] = np.asfortranarray(rng.uniform(0.0, 1.0, size=(P["C"], P["Q"])))
    sig = np.array(P["sigma"])
    sig[:, 1:] = sig[:, :1]                           # sigma_l = sigma_0: D = 0
    P["sigma"] = sig
    o = fullpn.solve_fcs(P, trajectory=True)
    inner = P["interior"]
    ref = P["psi0"] @ P["M"].T
    for u, psi, _ in o["trajectory"]:
        np.testing.assert_allclose(u[inner] + psi[inner] @ P["M"].T, ref[inner], rtol=1e-12,
                                   atol=1e-12 * np.abs(ref).max())
    # with absorption (sigma_l < sigma_0 for l >= 1) the l >= 1 moments decay instead
    P["sigma"] = np.array(cases.small_fcs(n=10, N=5, steps=12)["sigma"])
    o2 = fullpn.solve_fcs(P)
    tot = o2["u"][inner] + o2["psi"][inner] @ P["M"].T
    np.testing.assert_allclose(tot[:, 0], ref[inner, 0], rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_setup_ic_matches_reference_loops():
    s = Settings(12, 12, 2, "waterBeam", "Protons", nPN=3, nEnergies=5)
    omega = np.array([[0.0, 0.0, 1.0], [0.004, 0.0, np.sqrt(1 - 0.004 ** 2)],
                      [-0.006, 0.001, np.sqrt(1 - 0.006 ** 2 - 0.001 ** 2)], [0.3, 0.2, 0.93]])
    got = ph.setup_ic(s, omega)
    nx, ny = s.NCellsX, s.NCellsY
    ref = np.zeros((nx, ny, len(omega)))
    for i in range(3, nx - 1):                        # i = 3:nx-2 (1-based)
        for j in range(3, ny - 1):
            space = ph.normpdf(s.xMid[i - 1], s.x0, .1) * ph.normpdf(s.yMid[j - 1], s.y0, .1)
            for k in range(len(omega)):
                ref[i - 1, j - 1, k] = (10 ** 5 * np.exp(-10000.0 * (s.Omega1 - omega[k, 0]) ** 2)
                                        * np.exp(-10000.0 * (s.Omega3 - omega[k, 2]) ** 2) * space)
    assert ref.max() > 0
    np.testing.assert_allclose(got, ref.reshape(nx * ny, -1), rtol=1e-14, atol=0)
    assert ph.setup_ic(Settings(12, 12, 2, "waterBeamElectrons", "Electrons"), omega) is None


def test_builder_and_oracle_reject_non_square():
    s = Settings(9, 11, 2, "waterBeamElectrons", "Electrons", nPN=3, nEnergies=5)
    from titus_amd.problem import build_fcs_problem
    with pytest.raises(ValueError, match="square"):
        build_fcs_problem(s)


def test_library_rejects_bad_full_solver_problems():
    from titus_amd import _lib
    lib = _lib.load()
    P = cases.small_fcs(n=8, N=3, steps=4)
    st, keep = _lib.make_fcs_problem(P)
    h = ctypes.c_void_p()
    st.ny = st.nx + 1
    code = lib.titus_fcs_create(ctypes.byref(h), ctypes.byref(st), 0)
    assert code == -1 and b"square" in lib.titus_last_error()
    st.ny = st.nx
    bad = np.array(keep["bcells"])
    bad[0] = P["interior"][0]
    keep["bad"] = np.ascontiguousarray(bad, dtype=np.int32)
    st.bcells = keep["bad"].ctypes.data_as(_lib._PI)
    code = lib.titus_fcs_create(ctypes.byref(h), ctypes.byref(st), 0)
    assert code == -1 and b"four edges" in lib.titus_last_error()
