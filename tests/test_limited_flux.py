"""Pins of the second-order limited uncollided flux (``oracle.dlra.limited_flux``,
the reference's ``solveFlux!`` / ``slopefit`` / ``minmod``,
``code/SolverCSD.jl:278-386``).

* a per-ordinate loop restatement with the reference's quadrant blocks and
  its s1..s4 windows, written with scalar ``min``/``max``;
* invariants: constant psi/rho streams nothing; on linear data the limiter is
  inactive and the flux is the exact derivative Omega . grad(psi/rho); at a
  local extremum the limited slope is zero;
* the solve with ``flux_order = 2`` differs from the upwind solve (the switch
  reaches the sweep) and stays finite.
"""
import numpy as np

import cases
from oracle import dlra


def minmod(x, y):
    if x < 0:
        return min(max(y, x), 0.0)
    return min(max(y, 0.0), x)


def slopefit(l, c, r):
    return minmod(2.0 * (r - c), minmod(0.5 * (r - l), 2.0 * (c - l)))


def limited_loops(psi3, rho2d, omega, dx, dy):
    nx, ny, nq = psi3.shape
    phi = psi3 / rho2d[:, :, None]
    flux = np.zeros_like(psi3)
    for q in range(nq):
        ox, oz = omega[q, 0], omega[q, 2]
        for j in range(2, ny - 2):
            for i in range(2, nx - 2):
                if oz >= 0:
                    s1, s2, s3, s4 = phi[i, j - 2, q], phi[i, j - 1, q], phi[i, j, q], phi[i, j + 1, q]
                    north = s3 + 0.5 * slopefit(s2, s3, s4)
                    south = s2 + 0.5 * slopefit(s1, s2, s3)
                else:
                    s1, s2, s3, s4 = phi[i, j - 1, q], phi[i, j, q], phi[i, j + 1, q], phi[i, j + 2, q]
                    north = s3 - 0.5 * slopefit(s2, s3, s4)
                    south = s2 - 0.5 * slopefit(s1, s2, s3)
                if ox >= 0:
                    s1, s2, s3, s4 = phi[i - 2, j, q], phi[i - 1, j, q], phi[i, j, q], phi[i + 1, j, q]
                    east = s3 + 0.5 * slopefit(s2, s3, s4)
                    west = s2 + 0.5 * slopefit(s1, s2, s3)
                else:
                    s1, s2, s3, s4 = phi[i - 1, j, q], phi[i, j, q], phi[i + 1, j, q], phi[i + 2, j, q]
                    east = s3 - 0.5 * slopefit(s2, s3, s4)
                    west = s2 - 0.5 * slopefit(s1, s2, s3)
                flux[i, j, q] = ox / dx * (east - west) + oz / dy * (north - south)
    return flux


def _directions(nq, rng):
    v = rng.standard_normal((nq, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_limited_flux_matches_loop_restatement():
    rng = np.random.default_rng(11)
    nx, ny, nq, dx, dy = 9, 8, 12, 0.3, 0.45
    rho = rng.uniform(0.5, 2.0, size=(nx, ny))
    psi3 = rng.standard_normal((nx, ny, nq))
    omega = _directions(nq, rng)
    assert len(set(zip(omega[:, 0] >= 0, omega[:, 2] >= 0))) == 4
    got = dlra.limited_flux(psi3.reshape(nx * ny, nq), rho.reshape(-1), omega, nx, ny, dx, dy)
    ref = limited_loops(psi3, rho, omega, dx, dy).reshape(nx * ny, nq)
    np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-13)


def test_limited_flux_invariants():
    rng = np.random.default_rng(12)
    nx, ny, nq, dx, dy = 10, 9, 8, 0.2, 0.3
    rho = rng.uniform(0.5, 2.0, size=nx * ny)
    omega = _directions(nq, rng)
    inner = np.zeros((nx, ny), bool)
    inner[2:-2, 2:-2] = True
    inner = inner.reshape(-1)
    # constant psi/rho: no flux
    const = dlra.limited_flux(np.tile(rho[:, None], (1, nq)) * 2.5, rho, omega, nx, ny, dx, dy)
    assert np.abs(const).max() < 1e-12
    # linear psi/rho = a x + b y: the exact derivative, limiter inactive
    a, b = 1.7, -0.6
    x = np.repeat(np.arange(nx) * dx, ny)
    y = np.tile(np.arange(ny) * dy, nx)
    lin = (rho * (a * x + b * y))[:, None] * np.ones((1, nq))
    f = dlra.limited_flux(lin, rho, omega, nx, ny, dx, dy)
    want = (a * omega[:, 0] + b * omega[:, 2])[None, :] * np.ones((nx * ny, 1))
    np.testing.assert_allclose(f[inner], want[inner], rtol=1e-11, atol=1e-11)
    assert np.all(f[~inner] == 0.0)
    # a local maximum: both limited slopes vanish, the faces are the cell values
    assert dlra.slopefit(np.array(1.0), np.array(3.0), np.array(2.0)) == 0.0
    assert dlra.slopefit(np.array(1.0), np.array(2.0), np.array(4.0)) == 1.5


def smooth_initial_flux(P, seed=0):
    """A smooth blob of uncollided flux inside the grid (cells 3..n-4)."""
    nx, ny, Q = P["nx"], P["ny"], P["Q"]
    rng = np.random.default_rng(seed)
    x, y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    blob = np.exp(-((x - nx / 2) ** 2 + (y - ny / 3) ** 2) / (0.05 * nx * ny))
    blob[:3, :] = blob[-3:, :] = 0.0
    blob[:, :3] = blob[:, -3:] = 0.0
    return np.asfortranarray(blob.reshape(-1, 1) * rng.uniform(10.0, 50.0, size=(1, Q)))


def test_solve_with_limited_flux_differs_and_is_finite():
    """The boundary beam never crosses the first interior ring under the
    limited flux (ring cells have no flux, and the limited slope of a zero
    cell next to the source is zero), so the comparison starts from an
    initial flux inside the grid and no boundary source."""
    P = dict(cases.small_fixed(r=6))
    P["psi0"] = smooth_initial_flux(P)
    P["bc_base"] = np.zeros_like(P["bc_base"])
    o1 = dlra.solve_fixed(P, steps=20)
    P["flux_order"] = 2
    o2 = dlra.solve_fixed(P, steps=20)
    assert np.all(np.isfinite(o2["dose"]))
    d = np.linalg.norm(o2["dose"] - o1["dose"]) / np.linalg.norm(o1["dose"])
    assert 1e-4 < d < 0.5
