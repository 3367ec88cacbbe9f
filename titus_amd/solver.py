"""Host mirror of the reference's solver interface, backed by the HIP library.

Reference interface (Julia, ``code/SolverCSD.jl``):

* ``SolverCSD(settings)``                              -- ``:63-146``
* ``SolveFirstCollisionSourceDLR(obj)``                 -- ``:891-1120``
  returns ``(X*U, 0.5*sqrt(gamma[1])*Sigma, O*W*V, W*V, dose, psi)``
* ``SolveFirstCollisionSourceDLRBUGRejection(obj, recordGif=false)`` -- ``:2011-2374``
  returns the above plus ``[E; rank]``, ``[E eta]``, ``[E bound]`` before ``psi``
* ``SolveFirstCollisionSourceDLRParallel(obj, recordGif=false)`` -- ``:1726-2008``
  returns the fixed-rank tuple with ``rVec`` (2 x nEnergies) before ``psi``
* ``CSolveFirstCollisionSourceDLR(obj)``               -- ``:1122-1360``
  the fixed-rank solve with the limited flux ``solveFlux!``; same tuple
* ``SolveFirstCollisionSource(obj)``                   -- ``:465-586``
  returns ``(0.5*sqrt(gamma[1])*u, dose, psi)`` (the full-rank P_N solve)

Like the reference, ``Solve*`` mutates ``obj``: ``obj.dose`` accumulates and
``obj.settings.dE`` is overwritten with the grid step.  The energy loop runs
entirely in ``libtitus_hip.so`` on the GPU; this module only assembles the
operands (the work the reference's constructor does) and unpacks the results.
"""
import math

import numpy as np

from . import _lib
from .problem import (INTEGRATOR_BUG, INTEGRATOR_FIXED, INTEGRATOR_PARALLEL, build_fcs_problem,
                      build_problem)


# Rank cut of the range rule.  FP64: 1e-7 of the scale.  Mixed form: FP32
# arithmetic perturbs singular values by ~1e-7 relative, and the dose depends
# on which near-null directions enter the basis (both integrators keep at
# least two directions, and the fixed-rank one keeps all r), so the cut sits
# three decades above that noise.  Compare a mixed solve with an FP64 solve
# at the same cut.
RANK_TOL = 1e-7
MIXED_RANK_TOL = 1e-4

_PRECISIONS = {"fp64": _lib.PRECISION_FP64, "mixed": _lib.PRECISION_MIXED,
               "mixed_full_fp32": _lib.PRECISION_MIXED_FULL_FP32}


class SolverCSD:
    """Operands of the CSD transport solve (``SolverCSD.jl:20-147``).

    ``precision``: "fp64" (plain form, FP64 throughout), "mixed" (psi stored
    and swept in FP32, the rest FP64) or "mixed_full_fp32" (streaming stencil,
    psi contractions and Galerkin products in FP32 as well).  ``moments``: "subset" (k >= 0 real harmonics, the 2D-exact
    reduction) or "full" ((N+1)^2 as the reference).  ``rank_tol``: None picks
    the defaults above.
    """

    def __init__(self, settings, device=0, precision="fp64", moments="subset", rank_tol=None):
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        self.settings = settings
        self.device = device
        self.moments = moments
        self.precision = precision
        self.rank_tol = rank_tol
        nx, ny = settings.NCellsX, settings.NCellsY
        self.dose = np.zeros(nx * ny)
        self.gamma = np.array([2.0 / (2 * n + 1) for n in range(settings.nPN + 1)])
        self._problem = {}

    def problem(self, integrator):
        if integrator not in self._problem:
            self._problem[integrator] = build_problem(self.settings, self.moments, integrator)
        return self._problem[integrator]

    @property
    def qReduced(self):
        return self.problem(INTEGRATOR_FIXED)["omega"]

    @property
    def MReduced(self):
        return self.problem(INTEGRATOR_FIXED)["M"]

    def tolerance(self, integrator):
        if self.rank_tol is not None:
            return self.rank_tol
        return RANK_TOL if self.precision == "fp64" else MIXED_RANK_TOL

    def context(self, integrator):
        """A fresh device context for ``integrator`` (steps run through it)."""
        return _lib.Context(self.problem(integrator), integrator, device=self.device,
                            precision=_PRECISIONS[self.precision],
                            rank_tol=self.tolerance(integrator))


def _finish(obj, P, st):
    S = st["S"]
    U, sig, Vt = np.linalg.svd(S)
    V = Vt.T
    X, W = st["X"], st["W"]
    obj.dose += st["dose"]
    obj.settings.dE = P["dE"]
    nx, ny = obj.settings.NCellsX, obj.settings.NCellsY
    psi = st["psi"].reshape(nx, ny, -1)              # Vec2Mat per ordinate
    return X @ U, 0.5 * math.sqrt(obj.gamma[0]) * sig, P["O"] @ (W @ V), W @ V, obj.dose, psi


def SolveFirstCollisionSourceDLR(obj):
    """Fixed-rank first-collision-source DLRA solve on the GPU."""
    P = obj.problem(INTEGRATOR_FIXED)
    ctx = obj.context(INTEGRATOR_FIXED)
    try:
        ctx.solve(0)
        st = ctx.state()
    finally:
        ctx.close()
    return _finish(obj, P, st)


def CSolveFirstCollisionSourceDLR(obj):
    """The reference's GPU variant of the fixed-rank solve
    (``SolverCSD.jl:1122-1360``): ``SolveFirstCollisionSourceDLR`` with the
    second-order limited uncollided flux ``solveFlux!`` (``:1240``).  It
    defines ``psi`` only for the lung, lungOrig, liver and validation
    problems (``:1143-1151``); other problems raise, as there."""
    if obj.settings.problem not in ("lung", "lungOrig", "liver", "validation"):
        raise ValueError("CSolveFirstCollisionSourceDLR defines psi only for the lung, lungOrig, "
                         "liver and validation problems (SolverCSD.jl:1143-1151)")
    P = dict(obj.problem(INTEGRATOR_FIXED))
    P["flux_order"] = 2
    ctx = _lib.Context(P, INTEGRATOR_FIXED, device=obj.device, precision=_PRECISIONS[obj.precision],
                       rank_tol=obj.tolerance(INTEGRATOR_FIXED))
    try:
        ctx.solve(0)
        st = ctx.state()
    finally:
        ctx.close()
    return _finish(obj, P, st)


def CudaSolveFirstCollisionSourceDLR4thOrder(obj, recordGif=False):
    """The reference's CUDA.jl RK4 variant (``SolverCSD.jl:1422-1724``) on the
    GPU: fixed rank, the central streaming terms only (L2x with Ax, L2y with
    Az) integrated by RK4 in the K-, L- and S-steps, and the analytic
    uncollided flux of its pencil beam instead of a sweep.  Returns
    ``(X U, Sigma/sqrt2, O W V, W V, dose, psi)`` with psi C x Q, as there."""
    P = obj.problem(INTEGRATOR_FIXED)
    ctx = _lib.Context(P, _lib.INTEGRATOR_RK4, device=obj.device, precision=_PRECISIONS[obj.precision],
                       rank_tol=obj.tolerance(INTEGRATOR_FIXED))
    try:
        ctx.set_analytic_beam(P)
        ctx.solve(0)
        st = ctx.state()
    finally:
        ctx.close()
    out = _finish(obj, P, st)
    return (*out[:5], st["psi"])


def SolveFirstCollisionSourceDLRBUGRejection(obj, recordGif=False, rejection=False):
    """Rank-adaptive augmented-BUG solve on the GPU.  ``rejection=True`` turns
    on the step rejection of ``SolverCSD.jl:2204-2233, 2282-2301`` with
    ``settings.cEta`` (a rejected step restarts from its own start at the
    enlarged rank; see ``titus_set_rejection``)."""
    P = obj.problem(INTEGRATOR_BUG)
    ctx = obj.context(INTEGRATOR_BUG)
    try:
        if rejection:
            ctx.set_rejection(obj.settings.cEta)
        ctx.solve(0)
        st = ctx.state()
        done, _ = ctx.progress()
        ranks, etas = ctx.history(done)
    finally:
        ctx.close()
    out = _finish(obj, P, st)
    E = P["eGrid"][1:done + 1]
    rank_hist = np.vstack([E, ranks])
    eta_E = P["eGrid"][2:done + 1]                    # eta exists from step 2 on
    eta_hist = np.column_stack([eta_E, etas[:len(eta_E)]]) if done > 1 else np.zeros((0, 2))
    return (*out[:5], rank_hist, eta_hist, np.zeros((0, 2)), out[5])


def SolveFirstCollisionSourceDLRParallel(obj, recordGif=False):
    """Parallel rank-adaptive BUG solve on the GPU (``SolverCSD.jl:1726-2008``).

    Returns ``(X U, Sigma/sqrt2, O W V, W V, dose, rVec, psi)``; ``rVec`` is
    2 x nEnergies with the energy-step time t (the sum of the steps before
    it) and the rank after each step, and the initial rank in both rows of
    the first column, as the reference fills it (``:1805, 1945-1948``).
    """
    P = obj.problem(INTEGRATOR_PARALLEL)
    ctx = obj.context(INTEGRATOR_PARALLEL)
    try:
        ctx.solve(0)
        st = ctx.state()
        done, _ = ctx.progress()
        ranks, _ = ctx.history(done)
    finally:
        ctx.close()
    out = _finish(obj, P, st)
    rvec = np.full((2, P["nE"]), float(P["r"]))
    rvec[0, 1:done + 1] = np.arange(done) * P["dE"]
    rvec[1, 1:done + 1] = ranks[:done]
    return (*out[:5], rvec, out[5])


def SolveFirstCollisionSource(obj):
    """Full-rank P_N first-collision-source solve on the GPU
    (``SolverCSD.jl:465-586``): the dense C x Nm moments u with explicit-Euler
    streaming and the implicit per-cell scattering solve, FP64.

    Returns ``(0.5*sqrt(gamma[1])*u, dose, psi)`` as the reference does, with
    u C x Nm and psi nx x ny x Q; ``obj.dose`` accumulates the dose.  Square
    grids only (the reference's scattering loop numbers cells (i-1)*nx + j).
    """
    if obj.precision != "fp64":
        raise ValueError("SolveFirstCollisionSource runs in FP64 only")
    if "fcs" not in obj._problem:
        obj._problem["fcs"] = build_fcs_problem(obj.settings, obj.moments)
    P = obj._problem["fcs"]
    ctx = _lib.FcsContext(P, device=obj.device)
    try:
        ctx.solve(0)
        st = ctx.state()
    finally:
        ctx.close()
    obj.dose += st["dose"]
    obj.settings.dE = P["dE"]
    nx, ny = obj.settings.NCellsX, obj.settings.NCellsY
    return 0.5 * math.sqrt(obj.gamma[0]) * st["u"], obj.dose, st["psi"].reshape(nx, ny, -1)
