"""Pencil-beamlet plans: many independent beams on one phantom.

The reference solves one beam per call (``code/SolverCSD.jl:891, 2011``).  A
treatment plan is a weighted sum of beamlet doses, D(w) = sum_b w_b D_b; each
beamlet is an independent rank-adaptive DLRA solve with its own entry point,
direction, energy and rank cap.  ``beamlet_plan`` draws such a plan from a
seed; ``solve_plan`` runs every beamlet on the GPU concurrently through
``titus_solve_batch_lockstep`` (the kernels of all beamlets that reach the
same point of the energy step run as one launch) or ``titus_solve_batch``
(one stream per beamlet) and returns the per-beamlet doses.
"""
import math

import numpy as np

from . import _lib
from .problem import INTEGRATOR_BUG, build_problem
from .settings import Settings, ellipse_phantom
from .solver import MIXED_RANK_TOL, RANK_TOL


def beamlet_plan(n_beamlets=64, n=128, N=15, seed=0, rank_range=(5, 40), e_range=(6.0, 20.0),
                 max_angle_deg=15.0, length=10.0, sigma_x=0.25, angular_inv_width=1000.0,
                 integrator=INTEGRATOR_BUG):
    """Seeded beamlet plan on an n x n ellipse phantom (as the 200x200 case).

    Entry points uniform along the y = 0 and y = length edges (alternating),
    in the inner 80% of the edge; directions within +-max_angle_deg of the
    inward normal; kinetic energies uniform in e_range (MeV); rank caps
    uniform integers in rank_range (initial rank cap // 2 for the rank-adaptive
    integrator; the fixed-rank one keeps the cap throughout).  The angular width
    sigma_Omega^-1 = 1000 (about 2 degrees) keeps every tilted beam on several
    P_15 ordinates.  Returns a list of (Settings, problem dict).
    """
    rng = np.random.default_rng(seed)
    rho = ellipse_phantom(n, n, 0.0, length, 0.0, length, seed)
    plan = []
    for b in range(n_beamlets):
        bottom = (b % 2 == 0)
        x0 = rng.uniform(0.1 * length, 0.9 * length)
        ang = math.radians(rng.uniform(-max_angle_deg, max_angle_deg))
        ekin = rng.uniform(*e_range)
        cap = int(rng.integers(rank_range[0], rank_range[1] + 1))
        s = Settings(n + 1, n + 1, cap, "waterBeamElectrons", "Electrons",
                     a=0.0, b=length, c=0.0, d=length,
                     x0=x0, y0=0.0 if bottom else length,
                     Omega1=math.sin(ang), Omega3=math.cos(ang) if bottom else -math.cos(ang),
                     nPN=N, eMin=0.1, eMax=ekin + 0.5, sigmaX=sigma_x, sigmaY=sigma_x,
                     sigmaO1Inv=angular_inv_width, sigmaO3Inv=angular_inv_width,
                     theta=1e-2, thetaIndex=1.0, cfl=1.5)
        s.set_density(rho)
        plan.append((s, build_problem(s, "subset", integrator)))
    return plan


LOCKSTEP_GROUPS = 4   # groups of a lockstep batch (host work of one overlaps the others' kernels)


def solve_plan(plan, precision=_lib.PRECISION_MIXED, nthreads=0, device=0, rank_tol=None,
               lockstep=True, groups=LOCKSTEP_GROUPS):
    """Solve every beamlet concurrently; returns (doses C x B, steps per beamlet).

    ``lockstep``: shared launches (``titus_solve_batch_lockstep``, all beamlets
    at once in ``groups`` groups; ``nthreads`` > 0 limits how many run at
    once); otherwise one stream per beamlet over ``nthreads`` host workers
    (``titus_solve_batch``).  Both give the same doses bit for bit.
    ``rank_tol`` None: the default cut of the precision (``solver.RANK_TOL`` or
    ``solver.MIXED_RANK_TOL``).
    """
    if rank_tol is None:
        rank_tol = MIXED_RANK_TOL if precision == _lib.PRECISION_MIXED else RANK_TOL
    ctxs = [_lib.Context(P, INTEGRATOR_BUG, precision=precision, device=device, rank_tol=rank_tol)
            for _, P in plan]
    try:
        if lockstep:
            _lib.solve_batch_lockstep(ctxs, nthreads, groups)
        else:
            _lib.solve_batch(ctxs, nthreads)
        doses = np.stack([c.state()["dose"] for c in ctxs], axis=1)
        steps = np.array([c.progress()[0] for c in ctxs])
    finally:
        for c in ctxs:
            c.close()
    return doses, steps


def plan_dose(doses, weights):
    """Weighted plan dose D(w) = sum_b w_b D_b."""
    return doses @ np.asarray(weights, dtype=np.float64)
