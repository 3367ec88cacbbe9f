"""FP64 CPU oracle of the full-rank P_N solve -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it; the product path never does.

``solve_fcs`` restates ``SolveFirstCollisionSource(obj)``
(``code/SolverCSD.jl:465-586``) in numpy, T = Float64, for the operands of
``titus_amd.problem.build_fcs_problem``:

* the beam boundary source at ``energy[n]`` with ``PsiBeam``'s last argument
  n-1 (``:528-537``), on the four edges and every reduced ordinate;
* the uncollided update ``psi - dE flux(psi)`` with no 1/rho (``:549-551``),
  then ``psiNew = psi / (1 + dE sigma_0)`` everywhere (``:553``);
* the collided streaming ``uTilde = u - dE (L2x u Ax + L2y u Az + L1x u |Ax|
  + L1y u |Az|)`` with the density-scaled stencil matrices of
  ``stencils.jl`` (``:556``), ghost rows zeroed (``:557``);
* the per-cell implicit scattering ``(I + dE D) \\ (uTilde + dE Dvec M psiNew)``,
  D = diag(sigma_0 - Dvec), on interior cells (``:560-566``), cells numbered
  ``(i-1) nx + j`` as the reference does (square grids only);
* the dose ``dE (u[:,1] + uOUnc) SMid[n-1] / rho / (1 + [n == 2 or n == nE])``
  (``:575``), uOUnc = psiNew M[1,:] on interior cells, 0 on ghost cells.

Pinned by: the stencils and the upwind flux it shares with ``oracle.dlra``
(pinned there), a cell-by-cell restatement with explicit sparse stencil
matrices and a dense solve per cell on small grids, and the exact
conservation of u_0 + uOUnc by the scattering step when every sigma_l equals
sigma_0 and streaming is off (``tests/test_fcs_oracle.py``).
"""
import numpy as np

from .dlra import initial_psi, stencils, upwind_flux


def solve_fcs(P, steps=None, trajectory=False):
    """Returns dict(u, dose, psi, u_out = 0.5 sqrt(gamma[1]) u); with
    ``trajectory`` also the per-step list of (u, psi, dose)."""
    nx, ny = P["nx"], P["ny"]
    if nx != ny:
        raise ValueError("SolveFirstCollisionSource numbers cells (i-1)*nx + j: square grids only")
    C, Q, Nm = P["C"], P["Q"], P["Nm"]
    rho = np.asarray(P["density"], dtype=np.float64)
    ones = np.ones(C)
    bc = np.asarray(P["bcells"])
    interior = np.asarray(P["interior"])
    deg = np.asarray(P["degree"])
    omega, M = np.asarray(P["omega"]), np.asarray(P["M"])
    Ax, Az = np.asarray(P["Ax"]), np.asarray(P["Az"])
    aAx, aAz = np.asarray(P["absAx"]), np.asarray(P["absAz"])
    dx, dy = P["dx"], P["dy"]
    nE = P["nE"]
    last = nE - 1 if steps is None else min(nE - 1, steps)
    u = np.zeros((C, Nm))
    psi = initial_psi(P)
    dose = np.zeros(C)
    uOUnc = np.zeros(C)
    traj = []
    for s in range(1, last + 1):           # n = s + 1
        sig = P["sigma"][s]
        dE = P["dE_steps"][s - 1]
        psi[bc, :] = P["bc_base"] * P["fcs_bc_scale"][s - 1]
        Dvec = sig[deg]
        flux = upwind_flux(psi, ones, omega, nx, ny, dx, dy)
        psi = psi - dE * flux
        psiNew = psi / (1.0 + dE * sig[0])
        L1x, L1y, L2x, L2y = stencils(u, rho, nx, ny, dx, dy)
        uTilde = u - dE * (L2x @ Ax) - dE * (L2y @ Az) - dE * (L1x @ aAx) - dE * (L1y @ aAz)
        uTilde[bc, :] = 0.0
        uNew = np.zeros((C, Nm))
        rhs = uTilde[interior] + dE * (psiNew[interior] @ M.T) * Dvec[None, :]
        uNew[interior] = rhs / (1.0 + dE * (sig[0] - Dvec))[None, :]
        uOUnc[interior] = psiNew[interior] @ M[0, :]
        uNew[bc, :] = 0.0
        w = 2.0 if (s == 1 or s == nE - 1) else 1.0
        dose += dE * (uNew[:, 0] + uOUnc) * P["SMid"][s - 1] / rho / w
        u = uNew
        psi = psiNew
        if trajectory:
            traj.append((u.copy(), psi.copy(), dose.copy()))
    out = dict(u=u, dose=dose, psi=psi, u_out=0.5 * np.sqrt(2.0) * u)
    if trajectory:
        out["trajectory"] = traj
    return out
