"""Gradients of a plan-dose objective: the rrule of the beamlet plan.

The reference has no derivatives.  This module is the host side of the
device adjoint (``titus_set_checkpoints`` / ``titus_adjoint_batch`` in
``include/titus_hip.h``), in the shape of a ChainRulesCore ``rrule`` /
``torch.autograd.Function``: the forward solves every beamlet with
checkpoints, the pullback seeds each beamlet's adjoint with its share of
dJ/dD.

Objective (the plan-optimisation form):

    J(w, s) = || sum_b w_b D_b(rho0 o s) - D_target ||^2

over the beamlet weights w and per-voxel density scalings s (rho = rho0 o s).
D_b is the fixed-rank DLRA dose of beamlet b (``SolveFirstCollisionSourceDLR``,
``code/SolverCSD.jl:891-1120``) on the shared phantom; the energy grids and
beam sources of the beamlets are held fixed.  Gradients:

    dJ/dw_b = 2 <D - D_target, D_b>,
    dJ/ds   = rho0 o sum_b d<2 w_b (D - D_target), D_b>/d rho   (device adjoint).
"""
import numpy as np

from . import _lib


def piecewise_target(n_cells, ny, blocks=4, seed=0, scale=1.0):
    """Seeded piecewise-constant prescription: a blocks x blocks partition of
    the nx x ny grid (nx = n_cells / ny), one level per block drawn from
    U[0, scale)."""
    nx = n_cells // ny
    rng = np.random.default_rng(seed)
    levels = rng.uniform(0.0, scale, (blocks, blocks))
    bi = np.minimum(np.arange(nx) * blocks // nx, blocks - 1)
    bj = np.minimum(np.arange(ny) * blocks // ny, blocks - 1)
    return levels[bi[:, None], bj[None, :]].reshape(-1)


def solve_with_checkpoints(problems, every=16, nthreads=0, device=0, lockstep=True, groups=8):
    """FP64 fixed-rank solves of every problem dict, checkpointed for the
    adjoint: with shared launches (``lockstep``; ``nthreads`` > 0 limits how
    many run at once) or one stream per context over ``nthreads`` host
    workers.  Returns (contexts, doses C x B); the caller closes the
    contexts."""
    ctxs = [_lib.Context(P, P["integrator"], precision=_lib.PRECISION_FP64, device=device)
            for P in problems]
    try:
        for c in ctxs:
            c.set_checkpoints(every)
        if lockstep:
            _lib.solve_batch_lockstep(ctxs, nthreads, groups)
        else:
            _lib.solve_batch(ctxs, nthreads)
        doses = np.stack([c.state()["dose"] for c in ctxs], axis=1)
    except Exception:
        for c in ctxs:
            c.close()
        raise
    return ctxs, doses


def plan_objective_rrule(problems, weights, scale, target, every=16, nthreads=0, device=0,
                         lockstep=True, groups=8):
    """J and its pullback for the plan objective above.

    ``problems``: fixed-rank problem dicts (``beamlet_plan(...,
    integrator=INTEGRATOR_FIXED)``) sharing one phantom; their "density" is
    taken as rho0.  Returns (J, pullback) with pullback(Jbar) -> (wbar, sbar).
    The contexts stay alive until the pullback has run.  ``lockstep``: the
    forward solves and the adjoints with shared launches (the same values
    bit for bit as one stream per beamlet).
    """
    w = np.asarray(weights, dtype=np.float64)
    s = np.asarray(scale, dtype=np.float64)
    rho0 = np.asarray(problems[0]["density"], dtype=np.float64)
    scaled = [dict(P, density=rho0 * s) for P in problems]
    ctxs, doses = solve_with_checkpoints(scaled, every, nthreads, device, lockstep, groups)
    resid = doses @ w - np.asarray(target, dtype=np.float64)
    J = float(resid @ resid)

    def pullback(Jbar=1.0, per_beamlet=False):
        """(wbar, sbar); with per_beamlet also each beamlet's share of sbar
        (C x B) and the forward's decision hashes."""
        try:
            wbar = Jbar * 2.0 * (doses.T @ resid)
            bars = Jbar * 2.0 * resid[:, None] * w[None, :]
            grads = _lib.adjoint_batch(ctxs, bars, nthreads, lockstep, groups)   # d/d rho, C x B
            sbar = rho0 * grads.sum(axis=1)
            if per_beamlet:
                return wbar, sbar, rho0[:, None] * grads, [c.decisions()[0] for c in ctxs]
            return wbar, sbar
        finally:
            for c in ctxs:
                c.close()

    pullback.doses = doses
    pullback.residual = resid
    return J, pullback


def plan_objective_grad(problems, weights, scale, target, every=16, nthreads=0, device=0,
                        lockstep=True, groups=8):
    """(J, dJ/dw, dJ/ds) in one call."""
    J, pb = plan_objective_rrule(problems, weights, scale, target, every, nthreads, device,
                                 lockstep, groups)
    wbar, sbar = pb(1.0)
    return J, wbar, sbar
