"""Pins of the host-side set-up (P_N matrices, Roe matrices, quadrature,
transforms, energy grid, stencils, boundary bookkeeping).

Each test checks against a fixture of the reference's own tree, a closed form
or an independent construction -- not against a retyping of the same formula.
"""
import math
import os

import numpy as np
import pytest

from titus_amd import physics as ph
from titus_amd.settings import Settings, ellipse_phantom

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --- P_N matrices against the reference's validationData CSVs (N = 7) -------
@pytest.mark.parametrize("axis", ["x", "y"])
def test_flux_splitting_matches_reference_csv(axis):
    Ax, Ay, _ = ph.pn_matrices_full(7)
    A = Ax if axis == "x" else Ay
    absA = ph.abs_matrix(A)
    plus = np.loadtxt(os.path.join(GOLDEN, f"A{axis}Plus.csv"), delimiter=",")
    minus = np.loadtxt(os.path.join(GOLDEN, f"A{axis}Minus.csv"), delimiter=",")
    assert plus.shape == (64, 64)
    # the CSVs carry 6 significant digits
    np.testing.assert_allclose((A + absA) / 2, plus, atol=6e-7)
    np.testing.assert_allclose((A - absA) / 2, minus, atol=6e-7)


def test_flux_splitting_csv_catches_transposed_or_sign_errors():
    Ax, _, _ = ph.pn_matrices_full(7)
    plus = np.loadtxt(os.path.join(GOLDEN, "AxPlus.csv"), delimiter=",")
    assert np.abs((-Ax + ph.abs_matrix(Ax)) / 2 - plus).max() > 1e-2


@pytest.mark.parametrize("N", [1, 3, 7, 9])
def test_streaming_matrices_are_ordinate_moments(N):
    """Closed form: Ax = M diag(Omega_x) O and Az = M diag(Omega_z) O on the
    tensor quadrature of order N+1 (exact for degree <= 2N+1)."""
    lk, _, _ = ph.moment_set(N, "full")
    Ax, Ay, Az = ph.pn_matrices_full(N)
    O, M, pts, w = ph.trafo_matrices(N, lk, cs_phase=False)
    np.testing.assert_allclose(M @ O, np.eye(len(lk)), atol=1e-12)
    np.testing.assert_allclose(M @ (pts[:, 0:1] * O), Ax, atol=1e-12)
    np.testing.assert_allclose(M @ (pts[:, 1:2] * O), Ay, atol=1e-12)
    np.testing.assert_allclose(M @ (pts[:, 2:3] * O), Az, atol=1e-12)
    # the Condon-Shortley phase flips the x-streaming sign only
    O2, M2, _, _ = ph.trafo_matrices(N, lk, cs_phase=True)
    np.testing.assert_allclose(M2 @ (pts[:, 0:1] * O2), -Ax, atol=1e-12)
    np.testing.assert_allclose(M2 @ (pts[:, 2:3] * O2), Az, atol=1e-12)


def test_quadrature_closed_forms():
    pts, w, mu, phi = ph.quadrature(2)
    np.testing.assert_allclose(sorted(mu), [-1 / math.sqrt(3), 1 / math.sqrt(3)], atol=1e-15)
    for n in (2, 6, 10, 22):
        pts, w, mu, phi = ph.quadrature(n)
        assert pts.shape == (2 * n * n, 3)
        assert abs(w.sum() - 4 * math.pi) < 1e-12
        np.testing.assert_allclose(np.linalg.norm(pts, axis=1), 1.0, atol=1e-14)
        # integrates mu^2 exactly: 4 pi / 3
        assert abs((w * pts[:, 2] ** 2).sum() - 4 * math.pi / 3) < 1e-12


def test_abs_matrix_invariants():
    _, _, Az = ph.pn_matrices_full(9)
    Ax, _, _ = ph.pn_matrices_full(9)
    for A in (Ax, Az):
        B = ph.abs_matrix(A)
        np.testing.assert_allclose(B, B.T, atol=1e-15)
        np.testing.assert_allclose(B @ B, A @ A, atol=1e-9)          # |A|^2 = A^2
        assert np.linalg.eigvalsh(B).min() > -1e-9                    # PSD
        np.testing.assert_allclose(A @ B, B @ A, atol=1e-9)           # commute


@pytest.mark.parametrize("N", [3, 9, 21])
def test_k_nonnegative_subset_decouples(N):
    """Ax and Az couple k >= 0 moments only among themselves (2D reduction)."""
    Ax, _, Az = ph.pn_matrices_full(N)
    _, sub, _ = ph.moment_set(N, "subset")
    rest = np.setdiff1d(np.arange(Ax.shape[0]), sub)
    for A in (Ax, Az):
        assert np.abs(A[np.ix_(sub, rest)]).max() == 0.0
        assert np.abs(A[np.ix_(rest, sub)]).max() == 0.0
    assert len(sub) == (N + 1) * (N + 2) // 2


def test_real_harmonics_orthonormal_and_low_orders():
    lk, _, _ = ph.moment_set(3, "full")
    O, M, pts, w = ph.trafo_matrices(3, lk)
    Y00 = 1 / math.sqrt(4 * math.pi)
    np.testing.assert_allclose(O[:, 0], Y00, atol=1e-15)
    # Y_{1,1} = sqrt(3/4pi) Omega_x, Y_{1,0} = sqrt(3/4pi) Omega_z, Y_{1,-1} ~ Omega_y
    c = math.sqrt(3 / (4 * math.pi))
    idx = {t: i for i, t in enumerate(lk)}
    np.testing.assert_allclose(O[:, idx[(1, 1)]], c * pts[:, 0], atol=1e-14)
    np.testing.assert_allclose(O[:, idx[(1, 0)]], c * pts[:, 2], atol=1e-14)
    np.testing.assert_allclose(O[:, idx[(1, -1)]], c * pts[:, 1], atol=1e-14)


# --- energy grid -----------------------------------------------------------
def test_transformed_energy_closed_form():
    """Constant stopping power S: tau(E) = (E - E0)/S exactly (trapezoid)."""
    s = Settings(11, 11, 2, "waterBeamElectrons", "Electrons", eMin=0.1, eMax=10.5, nEnergies=11)
    g = ph.CSDGrid(s)
    E, St = g.E_tab, np.full_like(g.S_tab, 2.0)
    tr = np.concatenate([[0.0], np.cumsum(np.diff(E) / 2 * (1 / St[1:] + 1 / St[:-1]))])
    np.testing.assert_allclose(tr, (E - E[0]) / 2.0, rtol=1e-12, atol=1e-15)
    # the grid's own transform is the same trapezoid on the real table
    assert g.E_transformed[0] == 0.0 and np.all(np.diff(g.E_transformed) > 0)


@pytest.mark.parametrize("eMax,tau", [(21.5, 9.902), (21.0, 9.704), (15.0, 7.206)])
def test_tau_range_of_electron_grid(eMax, tau):
    s = Settings(11, 11, 2, "waterBeamElectrons", "Electrons", eMin=0.1, eMax=eMax, nEnergies=11)
    g = ph.CSDGrid(s)
    assert abs(g.eTrafo[-1] - tau) < 6e-3
    # eGrid runs from eMax down to eMin, S follows the table
    assert abs(g.eGrid[0] - eMax) < 1e-9 and abs(g.eGrid[-1] - 0.1) < 1e-9
    np.testing.assert_allclose(g.S, np.interp(g.eGrid, g.E_tab, g.S_tab))


def test_sigma_interpolation_is_piecewise_linear_and_throws():
    s = Settings(11, 11, 2, "waterBeamElectrons", "Electrons", nPN=5, eMin=0.1, nEnergies=11)
    g = ph.CSDGrid(s)
    E = g.E_sigmaTab
    # at table nodes the table values come back; midpoints average
    for i in (10, 50, 80):
        np.testing.assert_allclose(g.sigma_at_energy(E[i]), g.sigma_tab[i, :6], rtol=1e-14)
        mid = 0.5 * (E[i] + E[i + 1])
        np.testing.assert_allclose(g.sigma_at_energy(mid), 0.5 * (g.sigma_tab[i, :6] + g.sigma_tab[i + 1, :6]),
                                   rtol=1e-12)
    with pytest.raises(ValueError):
        g.sigma_at_energy(150.0)


def test_reference_grid_rule_gives_stable_step():
    from titus_amd.settings import water_slab_pencil_beam
    s = water_slab_pencil_beam()
    g = ph.CSDGrid(s)
    assert g.nEnergies == 243
    assert g.dE / s.dx < 0.7


# --- spatial bookkeeping ----------------------------------------------------
def test_boundary_index_and_restore_quirk():
    nx, ny = 5, 7
    b = ph.boundary_index(nx, ny)
    assert len(b) == 2 * nx + 2 * ny
    ij = {(c // ny, c % ny) for c in b}
    assert ij == {(i, j) for i in range(nx) for j in range(ny) if i in (0, nx - 1) or j in (0, ny - 1)}
    # linear index c into a column-major (nx, ny, nq) array: ordinate 0 at
    # (c mod nx, c div nx); for a square grid that is the same ghost-cell set
    r = ph.boundary_restore_cells(6, 6)
    assert set(r) == set(ph.boundary_index(6, 6))
    r57 = ph.boundary_restore_cells(nx, ny)
    expect = set()
    arr = np.zeros((nx, ny, 2), order="F")
    flat = arr.reshape(-1, order="F")
    for c in b:
        flat[c] = 1.0
    for i in range(nx):
        for j in range(ny):
            if arr[i, j, 0]:
                expect.add(i * ny + j)
    assert set(r57) == expect and not arr[:, :, 1].any()


def test_settings_grid_and_phantom():
    s = Settings(81, 81, 20, "waterBeamElectrons", "Electrons", a=0.0, b=5.0, c=0.0, d=5.0)
    assert s.NCellsX == 80 and abs(s.dx - 5 / 79) < 1e-15
    np.testing.assert_allclose(s.xMid[[0, -1]], [0.0, 5.0], atol=1e-12)
    p1 = ellipse_phantom(50, 50, 0, 10, 0, 10, seed=4)
    p2 = ellipse_phantom(50, 50, 0, 10, 0, 10, seed=4)
    assert np.array_equal(p1, p2)
    assert set(np.unique(p1)) <= {1.0, 0.25, 1.85}
    with pytest.raises(AttributeError):
        Settings(11, 11, 2, "waterBeam", "Electrons", notAField=1)
