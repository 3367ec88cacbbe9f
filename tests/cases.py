"""Problem instances shared by the parity tests (seeded, small enough for the oracle)."""
from titus_amd.problem import INTEGRATOR_BUG, INTEGRATOR_FIXED, INTEGRATOR_PARALLEL, build_problem
from titus_amd.settings import Settings, ellipse_phantom, heterogeneous_phantom, water_slab_pencil_beam


def small_fixed(n=24, N=5, r=6, steps=None, angular=2.0, density=None):
    """Beam on a coarse grid, fixed rank.  A broad angular profile (``angular`` =
    sigma_Omega^-1) keeps ordinates of every (Omega_x, Omega_z) sign quadrant."""
    s = Settings(n + 1, n + 1, r, "waterBeamElectrons", "Electrons", a=0.0, b=3.0, c=0.0, d=3.0,
                 x0=1.5, y0=0.0, nPN=N, eMin=0.5, sigmaX=0.4, sigmaY=0.4, eMax=12.5,
                 sigmaO1Inv=angular, sigmaO3Inv=angular,
                 nEnergies=None if steps is None else steps + 1)
    if density is not None:
        s.set_density(density)
    return build_problem(s, "subset", INTEGRATOR_FIXED)


def small_bug(n=30, N=7, cap=12, seed=3, steps=None, angular=2.0):
    """Heterogeneous ellipse phantom, rank-adaptive integrator, broad beam."""
    s = heterogeneous_phantom(n=n, N=N, seed=seed, r=cap, eKin=8.0)
    s.sigmaO1Inv = s.sigmaO3Inv = angular
    if steps is not None:
        s.nEnergies = steps + 1
    return build_problem(s, "subset", INTEGRATOR_BUG)


def small_parallel(n=30, N=7, cap=12, seed=3, steps=None, angular=2.0):
    """The small_bug phantom for the parallel integrator."""
    s = heterogeneous_phantom(n=n, N=N, seed=seed, r=cap, eKin=8.0)
    s.sigmaO1Inv = s.sigmaO3Inv = angular
    if steps is not None:
        s.nEnergies = steps + 1
    return build_problem(s, "subset", INTEGRATOR_PARALLEL)


def slab_beam():
    """The 80x80, P_9, rank-20 water-slab pencil beam (fixed rank)."""
    return build_problem(water_slab_pencil_beam(), "subset", INTEGRATOR_FIXED)


def phantom_bug(n=200, N=21, cap=40, seed=0):
    """The 200x200, P_21 heterogeneous phantom (rank-adaptive)."""
    return build_problem(heterogeneous_phantom(n=n, N=N, seed=seed, r=cap), "subset", INTEGRATOR_BUG)


def small_fcs(n=24, N=5, steps=None, angular=2.0, density=None, moments="subset", problem="waterBeamElectrons"):
    """The small_fixed beam for the full-rank P_N solver (square grid)."""
    from titus_amd.problem import build_fcs_problem
    s = Settings(n + 1, n + 1, 2, problem, "Electrons", a=0.0, b=3.0, c=0.0, d=3.0,
                 x0=1.5, y0=0.0, nPN=N, eMin=0.5, sigmaX=0.4, sigmaY=0.4, eMax=12.5,
                 sigmaO1Inv=angular, sigmaO3Inv=angular,
                 nEnergies=None if steps is None else steps + 1)
    if density is not None:
        s.set_density(density)
    return build_fcs_problem(s, moments)


def phantom_fcs(n=48, N=9, seed=3, steps=None, angular=2.0):
    """Heterogeneous ellipse phantom for the full-rank P_N solver."""
    from titus_amd.problem import build_fcs_problem
    s = heterogeneous_phantom(n=n, N=N, seed=seed, r=2, eKin=8.0)
    s.sigmaO1Inv = s.sigmaO3Inv = angular
    if steps is not None:
        s.nEnergies = steps + 1
    return build_fcs_problem(s, "subset")
