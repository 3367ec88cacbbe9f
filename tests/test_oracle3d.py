"""Pins of the 3D extension of the oracle (the 3D proton configuration).

The reference snapshot is 2D, so the 3D operators have no reference to match.
They are pinned here by
* a loop restatement of the 3D stencils and of the 3D upwind flux;
* the nz -> 2D limit: on fields constant along k, the x/y operators equal the
  2D ones slice by slice and the k operators vanish;
* mirror symmetry of the uncollided flux and of the full-rank (dense) solve
  for a beam centred in the cross-section (x -> -x with Omega_x -> -Omega_x,
  z -> -z with Omega_y -> -Omega_y);
* the DLRA fixed-rank solve equals the dense splitting at full rank after one
  step (the scattering substeps are exact there).
"""
import numpy as np
import pytest

from oracle import dlra
from titus_amd.problem3d import Settings3D, build_problem3d, proton_cube


def _loop_stencils3(U, rho, nx, ny, nz, dx, dy, dz):
    C, m = U.shape
    out = [np.zeros((C, m)) for _ in range(6)]
    c = lambda i, j, k: (i * ny + j) * nz + k
    for i in range(1, nx - 1):
        for j in range(1, ny - 1):
            for k in range(1, nz - 1):
                v = lambda ii, jj, kk: U[c(ii, jj, kk)] / rho[c(ii, jj, kk)]
                o = c(i, j, k)
                out[0][o] = (v(i, j, k) - 0.5 * v(i - 1, j, k) - 0.5 * v(i + 1, j, k)) / dx
                out[1][o] = (v(i, j, k) - 0.5 * v(i, j - 1, k) - 0.5 * v(i, j + 1, k)) / dy
                out[2][o] = (v(i, j, k) - 0.5 * v(i, j, k - 1) - 0.5 * v(i, j, k + 1)) / dz
                out[3][o] = (v(i + 1, j, k) - v(i - 1, j, k)) / (2 * dx)
                out[4][o] = (v(i, j + 1, k) - v(i, j - 1, k)) / (2 * dy)
                out[5][o] = (v(i, j, k + 1) - v(i, j, k - 1)) / (2 * dz)
    return out


def test_stencils3_loop_restatement():
    rng = np.random.default_rng(0)
    nx, ny, nz = 4, 5, 6
    U = rng.standard_normal((nx * ny * nz, 3))
    rho = 0.5 + rng.random(nx * ny * nz)
    a = dlra.stencils3(U, rho, nx, ny, nz, 0.3, 0.4, 0.5)
    b = _loop_stencils3(U, rho, nx, ny, nz, 0.3, 0.4, 0.5)
    for x, y in zip(a, b):
        assert np.allclose(x, y, rtol=1e-13, atol=1e-13)


def test_stencils3_2d_limit():
    rng = np.random.default_rng(1)
    nx, ny, nz = 6, 7, 5
    U2 = rng.standard_normal((nx * ny, 2))
    rho2 = 0.5 + rng.random(nx * ny)
    U3 = np.repeat(U2, nz, axis=0)                   # constant along k
    rho3 = np.repeat(rho2, nz)
    L3 = dlra.stencils3(U3, rho3, nx, ny, nz, 0.2, 0.3, 0.4)
    L2 = dlra.stencils(U2, rho2, nx, ny, 0.2, 0.3)
    inner = slice(1, nz - 1)
    for o3, o2 in ((0, 0), (1, 1), (3, 2), (4, 3)):
        got = L3[o3].reshape(nx, ny, nz, 2)[:, :, inner]
        want = np.repeat(L2[o2].reshape(nx, ny, 1, 2), nz - 2, axis=2)
        assert np.allclose(got, want, rtol=1e-13, atol=1e-13)
    for o3 in (2, 5):
        assert np.allclose(L3[o3], 0.0, atol=1e-12)


def test_upwind_flux3_loop_and_2d_limit():
    rng = np.random.default_rng(2)
    nx, ny, nz, Q = 5, 4, 6, 8
    psi = rng.random((nx * ny * nz, Q))
    rho = 0.5 + rng.random(nx * ny * nz)
    om = rng.standard_normal((Q, 3))
    om /= np.linalg.norm(om, axis=1)[:, None]
    F = dlra.upwind_flux3(psi, rho, om, nx, ny, nz, 0.2, 0.3, 0.4)
    c = lambda i, j, k: (i * ny + j) * nz + k
    v = psi / rho[:, None]
    for q in range(Q):
        for i in range(1, nx - 1):
            for j in range(1, ny - 1):
                for k in range(1, nz - 1):
                    o = c(i, j, k)
                    fx = (v[o, q] - v[c(i - 1, j, k), q]) if om[q, 0] >= 0 else (v[c(i + 1, j, k), q] - v[o, q])
                    fy = (v[o, q] - v[c(i, j - 1, k), q]) if om[q, 2] >= 0 else (v[c(i, j + 1, k), q] - v[o, q])
                    fz = (v[o, q] - v[c(i, j, k - 1), q]) if om[q, 1] >= 0 else (v[c(i, j, k + 1), q] - v[o, q])
                    want = fx * om[q, 0] / 0.2 + fy * om[q, 2] / 0.3 + fz * om[q, 1] / 0.4
                    assert abs(F[o, q] - want) <= 1e-12 * (1 + abs(want))
    # k-uniform field: equals the 2D flux slice by slice
    p2 = rng.random((nx * ny, Q))
    r2 = 0.5 + rng.random(nx * ny)
    F3 = dlra.upwind_flux3(np.repeat(p2, nz, axis=0), np.repeat(r2, nz), om, nx, ny, nz, 0.2, 0.3, 0.4)
    F2 = dlra.upwind_flux(p2, r2, om, nx, ny, 0.2, 0.3)
    got = F3.reshape(nx, ny, nz, Q)[:, :, 1:-1]
    assert np.allclose(got, np.repeat(F2.reshape(nx, ny, 1, Q), nz - 2, axis=2), rtol=1e-13, atol=1e-13)


def _mirror_problem(n=9, N=5, r=6):
    s = Settings3D(n=(n, n, n), h=0.1, r=r, nPN=N, nEnergies=301,
                   sigmaO1Inv=2.0, sigmaO3Inv=2.0)
    return build_problem3d(s)


def _mirror_map(P):
    """Cell permutation i -> nx-1-i, k -> nz-1-k, and the matching ordinate
    permutation Omega_x -> -Omega_x, Omega_y -> -Omega_y."""
    nx, ny, nz = P["nx"], P["ny"], P["nz"]
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = (((nx - 1 - i) * ny + j) * nz + (nz - 1 - k)).reshape(-1)
    om = P["omega"]
    target = om * np.array([-1.0, -1.0, 1.0])
    perm = np.array([int(np.argmin(np.linalg.norm(om - t, axis=1))) for t in target])
    assert np.allclose(om[perm], target, atol=1e-12)
    return cells, perm


def test_uncollided_flux_mirror_symmetry():
    P = _mirror_problem()
    cells, perm = _mirror_map(P)
    o = dlra.solve_dense(P, steps=40, stream=False)
    psi = o["psi"]
    assert np.allclose(psi[cells][:, perm], psi, rtol=1e-10, atol=1e-12 * np.abs(psi).max())


def test_dense_3d_dose_mirror_symmetry():
    P = _mirror_problem()
    cells, _ = _mirror_map(P)
    o = dlra.solve_dense(P, steps=30)
    d = o["dose"]
    assert np.abs(d).max() > 0
    assert np.allclose(d[cells], d, rtol=1e-9, atol=1e-12 * np.abs(d).max())


def test_dlra_3d_full_rank_one_step_equals_dense():
    s = Settings3D(n=(6, 6, 6), h=0.1, r=16, nPN=3, nEnergies=301, sigmaO1Inv=2.0, sigmaO3Inv=2.0)
    P = build_problem3d(s)
    assert P["r"] == P["Nm"]                         # full rank in the angle
    g = dlra.solve_fixed(P, steps=1)
    d = dlra.solve_dense(P, steps=1)
    assert np.linalg.norm(g["dose"] - d["dose"]) <= 1e-12 * np.linalg.norm(d["dose"])


def test_proton_cube_grid():
    """300 steps over 90 MeV protons: dE = 0.0213 g/cm^2 (CFL 0.71 at 0.3 g/cc)."""
    P = build_problem3d(proton_cube(n=8, N=3, r=4))
    assert P["nE"] == 301
    assert abs(P["dE"] - 0.0213) < 5e-4
    assert P["Nm"] == 16
    assert np.all(np.isfinite(P["sigma"]))
