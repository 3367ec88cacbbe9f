"""3D grids on the GPU against the 3D oracle (the 3D proton configuration).

Cells c = (i*ny + j)*nz + k; all (N+1)^2 moments; six stencil operators; the
3D sweep, streaming stencil and MFMA Galerkin products.  FP64 parity at the
2D bar (dose relative L2 <= 1e-10), the mixed form within 1e-4 of the FP64
oracle at the mixed rank cut.
"""
import sys

import numpy as np
import pytest

from oracle import dlra
from titus_amd import _lib
from titus_amd.problem import INTEGRATOR_BUG, INTEGRATOR_FIXED
from titus_amd.problem3d import Settings3D, build_problem3d, proton_cube
from titus_amd.solver import MIXED_RANK_TOL

pytestmark = pytest.mark.gpu


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def announce(name, P):
    print(f"\n[case] {name}: {P['nx']}x{P['ny']}x{P['nz']} Nm={P['Nm']} Q={P['Q']} "
          f"nE={P['nE']} r={P['r']}", flush=True)
    sys.stdout.flush()


def run(P, integ, last=0, precision=_lib.PRECISION_FP64, rank_tol=1e-7):
    ctx = _lib.Context(P, integ, precision=precision, rank_tol=rank_tol)
    try:
        ctx.solve(last)
        st = ctx.state()
        done, rank = ctx.progress()
        st.update(done=done, rank=rank, ranks=ctx.history(done)[0], plan=ctx.tuning("sweep"))
    finally:
        ctx.close()
    return st


def broad_cube(n=10, N=3, r=6, integ=INTEGRATOR_FIXED, seed=1):
    s = Settings3D(n=(n, n, n), h=0.1, r=r, nPN=N, nEnergies=301, sigmaO1Inv=2.0, sigmaO3Inv=2.0)
    rng = np.random.default_rng(seed)
    s.density = 0.6 + 0.8 * rng.random((n, n, n))
    return build_problem3d(s, integ)


@pytest.mark.parametrize("steps", [1, 2, 5])
def test_3d_fixed_first_steps(steps):
    P = broad_cube()
    announce(f"3d fixed {steps} steps", P)
    g = run(P, INTEGRATOR_FIXED, steps)
    o = dlra.solve_fixed(P, steps=steps)
    print(f"[3d fixed {steps}] dose {rel(g['dose'], o['dose']):.2e} psi {rel(g['psi'], o['psi']):.2e}",
          flush=True)
    assert rel(g["psi"], o["psi"]) <= 1e-13
    assert rel(g["dose"], o["dose"]) <= 1e-10
    ug = g["X"] @ g["S"] @ g["W"].T
    uo = o["X"] @ o["S"] @ o["W"].T
    assert rel(ug, uo) <= 1e-10


def test_3d_proton_cube_full_solve():
    P = build_problem3d(proton_cube(n=12, N=5, r=8))
    announce("3d proton cube 12^3", P)
    g = run(P, INTEGRATOR_FIXED)
    o = dlra.solve_fixed(P)
    e = rel(g["dose"], o["dose"])
    print(f"[3d proton 12^3] dose {e:.2e}", flush=True)
    assert g["done"] == P["nE"] - 1
    assert e <= 1e-10


def test_3d_bug_small():
    P = broad_cube(n=9, N=3, r=8, integ=INTEGRATOR_BUG, seed=2)
    announce("3d bug", P)
    g = run(P, INTEGRATOR_BUG, 60)
    o = dlra.solve_bug(P, steps=60)
    e = rel(g["dose"], o["dose"])
    print(f"[3d bug] dose {e:.2e} ranks {list(g['ranks'][:12])}", flush=True)
    assert list(g["ranks"]) == list(o["ranks"])
    assert e <= 1e-10


def test_3d_mixed_form(monkeypatch):
    P = build_problem3d(proton_cube(n=12, N=5, r=8))
    announce("3d mixed", P)
    monkeypatch.setattr(dlra, "RANK_TOL", MIXED_RANK_TOL)
    g = run(P, INTEGRATOR_FIXED, precision=_lib.PRECISION_MIXED, rank_tol=MIXED_RANK_TOL)
    o = dlra.solve_fixed(P)
    e = rel(g["dose"], o["dose"])
    print(f"[3d mixed] dose vs FP64 oracle {e:.2e}", flush=True)
    assert 0.0 < e <= 1e-4


def test_3d_tuned_sweep():
    P = broad_cube(n=40, N=3, r=4)
    announce("3d tuned sweep", P)
    assert P["C"] * P["Q"] >= 2 ** 18
    g = run(P, INTEGRATOR_FIXED, 3)
    print(f"[3d tuned sweep] plan {g['plan']}", flush=True)
    assert g["plan"]["tuned"]
    o = dlra.solve_fixed(P, steps=3)
    assert rel(g["psi"], o["psi"]) <= 1e-13
    assert rel(g["dose"], o["dose"]) <= 1e-10


@pytest.mark.parametrize("integ,r", [(INTEGRATOR_FIXED, 1), (INTEGRATOR_BUG, 2)])
def test_3d_smallest_cube_lowest_order(integ, r):
    """A 5^3 grid (27 interior cells) with P_1 (four moments): rank 1 fixed,
    and the BUG with cap 2 (augmented bases of four columns = all moments)."""
    P = broad_cube(n=5, N=1, r=r, integ=integ, seed=4)
    announce(f"3d 5^3 P_1, integrator {integ}, r {r}", P)
    g = run(P, integ, 20)
    o = dlra.solve_bug(P, steps=20) if integ == INTEGRATOR_BUG else dlra.solve_fixed(P, steps=20)
    ed = rel(g["dose"], o["dose"])
    print(f"[3d smallest] ranks {list(g['ranks'])} dose {ed:.2e}", flush=True)
    assert np.abs(o["dose"]).max() > 0 and ed <= 1e-10
    if integ == INTEGRATOR_BUG:
        assert np.array_equal(g["ranks"], np.asarray(o["ranks"])[:len(g["ranks"])])
