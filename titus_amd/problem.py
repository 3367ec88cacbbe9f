"""Assemble the operands of the DLRA energy loop (the ``SolverCSD`` constructor).

``build_problem(settings)`` restates ``SolverCSD(settings)`` (``code/SolverCSD.jl:63-146``)
and the set-up part of the ``Solve*`` functions (ordinate reduction, initial
factors, ``SolverCSD.jl:904-961``) with T = Float64.  The result is a plain
dict of numpy arrays: exactly what crosses the C-ABI (``titus_problem_v1``).

Layouts (all FP64, 0-based):

* cells c = i*ny + j; ``density`` has length C = nx*ny;
* matrices are column-major as in Julia: ``X`` is C x r with each basis vector
  contiguous, ``W`` is Nm x r, ``S`` r x r; numpy arrays here are kept in
  Fortran order so that their buffers are already in that layout;
* the uncollided flux ``psi`` is stored C x Q (cell-major within each ordinate
  column), i.e. ``Mat2Vec(psi)`` of ``utils.jl:67-78``.
"""
import functools

import numpy as np

from . import physics as ph

INTEGRATOR_FIXED = 0      # SolveFirstCollisionSourceDLR
INTEGRATOR_BUG = 1        # SolveFirstCollisionSourceDLRBUGRejection (rejection off)
INTEGRATOR_PARALLEL = 2   # SolveFirstCollisionSourceDLRParallel


@functools.lru_cache(maxsize=8)
def angular_operands(N, moments, cs_phase):
    """P_N, Roe and transform matrices of degree N (shared by every beam)."""
    lk, full_idx, degree = ph.moment_set(N, moments)
    Axf, _, Azf = ph.pn_matrices_full(N)
    Ax = Axf[np.ix_(full_idx, full_idx)]
    Az = Azf[np.ix_(full_idx, full_idx)]
    absAx = ph.abs_matrix(Ax)
    absAz = ph.abs_matrix(Az)
    O, M, pts, w = ph.trafo_matrices(N, lk, cs_phase=cs_phase)
    for a in (Ax, Az, absAx, absAz, O, M, pts, w):
        a.setflags(write=False)
    return lk, degree, Ax, Az, absAx, absAz, O, M, pts, w


# problems whose beam enters through the boundary source (PsiBeam); the
# line-source family is an initial-value problem with PsiBeam = 0 and its own
# ordinate selection (SolverCSD.jl:907-911), not provided here
BEAM_PROBLEMS = ("waterBeam", "waterBeamElectrons", "lung", "liver", "slab", "phantom")


def build_problem(s, moments="subset", integrator=INTEGRATOR_FIXED):
    if s.problem not in BEAM_PROBLEMS:
        raise ValueError(f"problem {s.problem!r}: only the beam problems {BEAM_PROBLEMS} are provided")
    nx, ny = s.NCellsX, s.NCellsY
    C = nx * ny
    N = s.nPN
    csd = ph.CSDGrid(s)
    lk, degree, Ax, Az, absAx, absAz, O, M, pts, w = angular_operands(N, moments, bool(s.csPhase))

    # ordinate reduction: keep directions the beam reaches (SolverCSD.jl:913-925)
    psi_beam = ph.beam_angular(s, pts)
    idx = np.nonzero(psi_beam > s.floorPsi * psi_beam.max())[0]
    omega = pts[idx]
    Mred = M[:, idx]

    # boundary source: separable into a per-(cell, ordinate) base and a per-step
    # scale (SolverCSD.jl:977-988, PsiBeam :240-276)
    bidx = ph.boundary_index(nx, ny)
    bcells = np.unique(bidx)
    bi, bj = bcells // ny, bcells % ny
    spatial = ph.beam_spatial(s, s.xMid[bi], s.yMid[bj])
    bc_base = np.asfortranarray(spatial[:, None] * ph.beam_angular(s, omega)[None, :])
    bc_scale = np.array([ph.beam_energy_factor(s, csd.eGrid[n], csd.S[n])
                         for n in range(csd.nEnergies)])

    interior = np.ones(C, dtype=bool)
    interior[bcells] = False

    if integrator in (INTEGRATOR_BUG, INTEGRATOR_PARALLEL):
        r0 = s.r // 2                         # SolverCSD.jl:2013, 1728
    else:
        r0 = s.r
    if r0 < 1:
        raise ValueError("rank must be >= 1")
    if r0 > C or r0 > len(lk):
        raise ValueError("rank exceeds problem dimensions")

    # svd!(zeros) -> leading identity columns; ghost rows of X zeroed (Q6)
    X0 = np.zeros((C, r0), order="F")
    X0[np.arange(r0), np.arange(r0)] = 1.0
    X0[bcells, :] = 0.0
    W0 = np.zeros((len(lk), r0), order="F")
    W0[np.arange(r0), np.arange(r0)] = 1.0
    S0 = np.zeros((r0, r0), order="F")
    psi0 = ph.setup_ic(s, omega)

    return dict(
        nx=nx, ny=ny, C=C, dx=float(s.dx), dy=float(s.dy),
        density=np.ascontiguousarray(s.density.reshape(C)),
        N=N, Nm=len(lk), degree=degree.astype(np.int32), lk=lk,
        Ax=np.asfortranarray(Ax), Az=np.asfortranarray(Az),
        absAx=np.asfortranarray(absAx), absAz=np.asfortranarray(absAz),
        Q=len(idx), Qfull=len(w), omega=np.asfortranarray(omega),
        M=np.asfortranarray(Mred), O=np.asfortranarray(O), ordinate_index=idx,
        nE=csd.nEnergies, dE=csd.dE, eGrid=csd.eGrid, Sst=csd.S,
        eTrafo=csd.eTrafo, SMid=csd.SMid,
        sigma=np.ascontiguousarray(csd.sigma_table()),
        bcells=bcells.astype(np.int32), bc_base=bc_base, bc_scale=bc_scale,
        restore_cells=ph.boundary_restore_cells(nx, ny).astype(np.int32),
        interior=np.nonzero(interior)[0].astype(np.int32),
        integrator=integrator, r=r0, rank_cap=int(s.rank_cap),
        theta=float(s.theta), thetaIndex=float(s.thetaIndex), cEta=float(s.cEta),
        X0=X0, W0=W0, S0=S0, psi0=psi0,
        # the beam of the analytic uncollided flux (CUDA.jl variant, SolverCSD.jl:1437-1440)
        xMid=np.asarray(s.xMid, dtype=np.float64), yMid=np.asarray(s.yMid, dtype=np.float64),
        beam=dict(x0=float(s.x0), y0=float(s.y0), Omega1=float(s.Omega1), Omega3=float(s.Omega3)),
    )


def build_fcs_problem(s, moments="subset"):
    """Operands of ``SolveFirstCollisionSource(obj)`` (``SolverCSD.jl:465-586``):
    those of ``build_problem`` plus the per-step energy width
    ``eTrafo[n] - eTrafo[n-1]`` (``:527``), ``SMid[n-1]`` (``:575``) and the
    beam's energy factor at ``energy[n]`` with ``S[n-1]`` (``PsiBeam(...,
    energy[n], ..., n-1)``, ``:530-537``), each indexed by step - 1."""
    if s.NCellsX != s.NCellsY:
        raise ValueError("SolveFirstCollisionSource numbers cells (i-1)*nx + j (SolverCSD.jl:563): "
                         "square grids only")
    P = build_problem(s, moments, INTEGRATOR_FIXED)
    nE = P["nE"]
    P["dE_steps"] = np.diff(P["eTrafo"])
    P["fcs_bc_scale"] = np.array([ph.beam_energy_factor(s, P["eGrid"][n], P["Sst"][n - 1])
                                  for n in range(1, nE)])
    return P
