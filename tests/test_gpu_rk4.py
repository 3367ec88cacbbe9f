"""The CUDA.jl RK4 variant (TITUS_INTEGRATOR_RK4,
CudaSolveFirstCollisionSourceDLR4thOrder, SolverCSD.jl:1422-1724) on the GPU
against the oracle's restatement (oracle.dlra.solve_rk4, pinned by
tests/test_rk4_oracle.py).  Bar: dose relative L2 <= 1e-10 (FP64); psi (the
analytic flux) to rounding."""
import numpy as np
import pytest

import cases
from oracle import dlra
from titus_amd import _lib
from titus_amd.problem import INTEGRATOR_FIXED, build_problem
from titus_amd.settings import Settings, water_slab_pencil_beam

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def run(P, steps=0, precision=_lib.PRECISION_FP64, rank_tol=1e-7):
    ctx = _lib.Context(P, _lib.INTEGRATOR_RK4, precision=precision, rank_tol=rank_tol)
    try:
        ctx.set_analytic_beam(P)
        ctx.solve(steps)
        return ctx.state()
    finally:
        ctx.close()


def aimed(s):
    """The settings with the beam direction set to the quadrature ordinate
    closest to it: the analytic flux's angular factor exp(-1e4 |Omega -
    w_q|^2) (width fixed in the variant) is then 1 on that ordinate instead of
    vanishing between ordinates."""
    from titus_amd import physics as ph
    pts = ph.quadrature(s.nPN + 1)[0]
    q = int(np.argmin((pts[:, 0] - s.Omega1) ** 2 + (pts[:, 2] - s.Omega3) ** 2))
    s.Omega1, s.Omega3 = float(pts[q, 0]), float(pts[q, 2])
    return build_problem(s, "subset", INTEGRATOR_FIXED)


def scaled(P, f=1e-4):
    """Cross sections times f.  With the electron tables dE sigma_0 is about
    3000 per step, so the variant's exp(-int sigma) underflows to 0 after the
    first step (as in the reference); scaled down, the analytic flux lives
    through the whole grid and every part of the step is exercised."""
    return dict(P, sigma=np.asarray(P["sigma"]) * f)


def water(n, N, r, steps, problem="waterBeam"):
    s = Settings(n + 1, n + 1, r, problem, "Electrons", a=0.0, b=3.0, c=0.0, d=3.0, x0=1.5,
                 y0=0.6, nPN=N, eMin=0.5, eMax=12.5, sigmaO1Inv=2.0, sigmaO3Inv=2.0, cfl=0.8,
                 nEnergies=steps + 1)
    return aimed(s)


@pytest.mark.parametrize("case", ["waterBeam 40^2 P_5 r6", "electrons 32^2 P_7 r8",
                                  "slab 80^2 P_9 r20"])
def test_rk4_matches_the_oracle(case):
    if case.startswith("waterBeam"):
        P = scaled(water(40, 5, 6, 60))
    elif case.startswith("electrons"):
        P = scaled(water(32, 7, 8, 40, problem="waterBeamElectrons"))
    else:
        P = scaled(aimed(water_slab_pencil_beam()))
    print(f"\n[case] rk4 {case}: C={P['C']} Nm={P['Nm']} Q={P['Q']} nE={P['nE']}", flush=True)
    g = run(P)
    o = dlra.solve_rk4(P)
    ed, ep = rel(g["dose"], o["dose"]), rel(g["psi"], o["psi"])
    eu = rel(g["X"] @ g["S"] @ g["W"].T, o["X"] @ o["S"] @ o["W"].T)
    print(f"[rk4 {case}] dose rel {ed:.3e} u rel {eu:.3e} psi rel {ep:.3e} "
          f"max dose {np.abs(o['dose']).max():.3e}", flush=True)
    assert np.abs(o["dose"]).max() > 1e-3
    assert ep <= 1e-13
    assert ed <= 1e-10


def test_rk4_needs_its_beam_and_is_2d():
    P = water(16, 3, 2, 4)
    ctx = _lib.Context(P, _lib.INTEGRATOR_RK4)
    try:
        with pytest.raises(_lib.TitusError) as e:
            ctx.step(1)                                  # no analytic beam set
        assert e.value.code == -5
    finally:
        ctx.close()
    ctx = _lib.Context(P, _lib.INTEGRATOR_FIXED)
    try:
        with pytest.raises(_lib.TitusError):
            ctx.set_analytic_beam(P)                     # not the RK4 integrator
    finally:
        ctx.close()
