"""Rejection of the rank-adaptive step on the GPU (titus_set_rejection)
against the oracle with the same c_eta (``SolverCSD.jl:2204-2233,
2282-2301``; semantics in ``oracle.dlra.solve_bug``).

Bar: the same number of rejections, equal rank histories, dose relative L2
<= 1e-10 and psi equal to rounding (FP64).
"""
import numpy as np
import pytest

import cases
from oracle import dlra
from titus_amd import _lib

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def run(P, c_eta, steps, precision=_lib.PRECISION_FP64, rank_tol=1e-7):
    ctx = _lib.Context(P, _lib.INTEGRATOR_BUG, precision=precision, rank_tol=rank_tol)
    try:
        ctx.set_rejection(c_eta)
        ctx.solve(steps)
        st = ctx.state()
        done, _ = ctx.progress()
        st["ranks"] = ctx.history(done)[0]
        st["rejections"] = ctx.rejections()
    finally:
        ctx.close()
    return st


@pytest.mark.parametrize("cap,c_eta", [(12, 5.0), (12, 1e-2), (16, 1e-2), (16, 1e-6)])
def test_rejection_matches_the_oracle(cap, c_eta):
    P = cases.small_bug(steps=30, cap=cap)
    print(f"\n[case] rejection cap {cap} c_eta {c_eta:g}: C={P['C']} Nm={P['Nm']}", flush=True)
    g = run(P, c_eta, 30)
    o = dlra.solve_bug(dict(P, rejection=c_eta), steps=30)
    ed = rel(g["dose"], o["dose"])
    print(f"[rejection cap {cap} c_eta {c_eta:g}] rejections {g['rejections']} / {o['rejections']} "
          f"dose rel {ed:.3e}", flush=True)
    assert g["rejections"] == o["rejections"] >= 1
    assert np.array_equal(g["ranks"], np.asarray(o["ranks"]))
    assert rel(g["psi"], o["psi"]) <= 1e-13
    assert ed <= 1e-10


def test_rejection_on_the_phantom_and_mixed_form():
    """The 48 x 48 phantom at P_9, 80 steps: FP64 with rejection against the
    oracle; the mixed form with rejection runs and stays finite."""
    from titus_amd.problem import INTEGRATOR_BUG, build_problem
    from titus_amd.settings import heterogeneous_phantom
    s = heterogeneous_phantom(n=48, N=9, seed=3, r=16, eKin=8.0)
    s.sigmaO1Inv = s.sigmaO3Inv = 2.0
    P = build_problem(s, "subset", INTEGRATOR_BUG)
    print(f"\n[case] rejection phantom 48x48 P_9: C={P['C']}", flush=True)
    g = run(P, 5.0, 80)
    o = dlra.solve_bug(dict(P, rejection=5.0), steps=80)
    ed = rel(g["dose"], o["dose"])
    m = run(P, 5.0, 80, _lib.PRECISION_MIXED, 1e-4)
    print(f"[rejection phantom] rejections {g['rejections']} dose rel {ed:.3e}; mixed rejections "
          f"{m['rejections']}", flush=True)
    assert g["rejections"] == o["rejections"]
    assert np.array_equal(g["ranks"], np.asarray(o["ranks"]))
    assert ed <= 1e-10
    assert np.all(np.isfinite(m["dose"]))


def test_rejection_refusals():
    Pf = cases.small_fixed(steps=3)
    c = _lib.Context(Pf, _lib.INTEGRATOR_FIXED)
    try:
        with pytest.raises(_lib.TitusError):
            c.set_rejection(5.0)                       # fixed rank: no rejection
    finally:
        c.close()
    Pb = cases.small_bug(steps=3, cap=8)
    c = _lib.Context(Pb, _lib.INTEGRATOR_BUG)
    try:
        with pytest.raises(_lib.TitusError):
            c.set_rejection(-1.0)
        c.step(1)
        with pytest.raises(_lib.TitusError):
            c.set_rejection(5.0)                       # after the first step
    finally:
        c.close()
