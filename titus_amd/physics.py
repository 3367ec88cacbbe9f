"""Discretisation operands of the CSD transport step (host-side setup).

Each function restates one builder of the reference and cites it:

* energy grid and stopping power      -- ``code/CSD.jl:27-95``
* scattering moments sigma_l(E)        -- ``code/CSD.jl:97-127``
* P_N streaming matrices Ax, Ay, Az    -- ``code/PNSystem.jl:58-194``
* Roe matrices abs(A)                  -- ``code/SolverCSD.jl:83-97``
* tensor quadrature                    -- ``code/quadratures/standardQuadrature.jl:7-37``
* moment <-> ordinate transforms O, M  -- ``code/quadratures/Quadrature.jl:173-208``,
                                          ``code/utils.jl:12-41``
* boundary cell list                   -- ``code/SolverCSD.jl:108-131``
* beam boundary source                 -- ``code/SolverCSD.jl:240-276``

Moments.  The reference carries all (N+1)^2 real-SH moments (``PNSystem.jl:28``)
ordered by ``GlobalIndex(l,k) = l^2 + k + l``.  In 2D the moments with k < 0
(odd in Omega_y) decouple exactly from the k >= 0 ones under Ax and Az, and a
beam symmetric in Omega_y never excites them, so this build carries the k >= 0
subset, ordered l-major with k = 0..l (index l(l+1)/2 + k): (N+1)(N+2)/2
moments.  ``moment_set(N, "full")`` gives the reference's full ordering.

Cells.  Cell (i, j) (0-based, i along x) has linear index c = i*ny + j, the
reference's ``vectorIndex`` (``utils.jl:98-100``) shifted to 0-based.
"""
import json
import functools
import math
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def load_tables(particle):
    name = "electrons.json" if particle == "Electrons" else "protons.json"
    with open(os.path.join(_DATA, name)) as f:
        d = json.load(f)
    return {k: (np.asarray(v, dtype=np.float64) if k != "provenance" else v)
            for k, v in d.items()}


def interp_throw(xt, yt, x):
    """Piecewise-linear interpolation that raises outside the table.

    Interpolations.jl ``LinearInterpolation(...; extrapolation_bc=Throw())``.
    """
    x = np.asarray(x, dtype=np.float64)
    lo, hi = xt[0], xt[-1]
    if np.any(x < lo - 1e-12 * abs(lo)) or np.any(x > hi + 1e-12 * abs(hi)):
        raise ValueError(f"energy outside table range [{lo}, {hi}]")
    return np.interp(x, xt, yt)


class CSDGrid:
    """Transformed-energy grid (``CSD.jl:27-95``).

    tau(E) = int_{E_tab[0]}^{E} dE'/S(E') by the trapezoid rule on the table;
    ``eTrafo`` is uniform in [0, tau(eMax) - tau(eMin)] with ``nEnergies``
    points; ``eGrid = E(tau(eMax) - eTrafo)``; ``S = S(eGrid)``.
    """

    def __init__(self, settings):
        s = settings
        t = load_tables(s.particle)
        if s.particle == "Protons":
            # MaterialParametersProtons.jl:2802-2808: energies shifted to total energy
            S_tab = t["S_tab_PSTAR"]
            E_tab = t["E_tab_PSTAR"] + s.eRest
            self.E_sigmaTab = t["E_sigmaTab"] + s.eRest
            self.sigma_tab = t["sigmaEl_OInt_ICRU"]
            minE = max(s.eMin, 0.001) + s.eRest if s.eMin < s.eRest else s.eMin
        else:
            S_tab, E_tab = t["S_tab"], t["E_tab"]
            self.E_sigmaTab = t["E_sigmaTab"]
            self.sigma_tab = t["sigma_tab"]
            minE = s.eMin
        maxE = s.eMax
        self.E_tab, self.S_tab = E_tab, S_tab
        E_tr = np.zeros(len(E_tab))
        for i in range(1, len(E_tab)):
            E_tr[i] = E_tr[i - 1] + (E_tab[i] - E_tab[i - 1]) / 2 * (1.0 / S_tab[i] + 1.0 / S_tab[i - 1])
        self.E_transformed = E_tr
        eMaxTrafo = float(interp_throw(E_tab, E_tr, maxE))
        eMinTrafo = float(interp_throw(E_tab, E_tr, minE))
        nE = s.nEnergies if s.nEnergies is not None else int(math.ceil(maxE / s.dE))
        if nE < 3:
            raise ValueError("need at least 3 energy points")
        self.eTrafo = np.linspace(0.0, eMaxTrafo - eMinTrafo, nE)
        self.eGrid = interp_throw(E_tr, E_tab, eMaxTrafo - self.eTrafo)
        self.S = interp_throw(E_tab, S_tab, self.eGrid)
        dE = np.diff(self.eTrafo)
        eGridMid = interp_throw(E_tr, E_tab, eMaxTrafo - (self.eTrafo[:-1] + 0.5 * dE))
        self.SMid = interp_throw(E_tab, S_tab, eGridMid)
        self.dE = float(self.eTrafo[1] - self.eTrafo[0])   # SolverCSD.jl:960
        self.nEnergies = nE
        self.settings = s

    def sigma_at_energy(self, energy):
        """sigma_l(E), l = 0..nPN (``CSD.jl:97-127``)."""
        s = self.settings
        if s.particle == "Protons":
            return self._sigma_protons(energy)
        if energy <= 5e-5:
            energy = 5e-5 + 1e-9
        if s.nPN + 1 > self.sigma_tab.shape[1]:
            raise ValueError("table holds only 40 Legendre moments")
        return np.array([float(interp_throw(self.E_sigmaTab, self.sigma_tab[:, i], energy))
                         for i in range(s.nPN + 1)])

    def _sigma_protons(self, energy):
        """Proton branch (``CSD.jl:98-115``): O-16 elastic (water's oxygen mass
        fraction 0.888106, dropped below 7 MeV kinetic), proton-proton nuclear
        (hydrogen fraction 0.111894) and single Coulomb scattering, each
        interpolated linearly in total energy."""
        s = self.settings
        if energy <= 0.001 + s.eRest:
            energy = 0.001 + s.eRest
        if s.nPN + 1 > self.sigma_tab.shape[1]:
            raise ValueError("tables hold only 40 Legendre moments")
        pp, ce = proton_moment_tables()
        E_tab = self.E_tab
        y = np.zeros(s.nPN + 1)
        for i in range(s.nPN + 1):
            v_pp = float(interp_throw(E_tab, pp[:, i], energy))
            v_ce = float(interp_throw(E_tab, ce[:, i], energy))
            if energy < 7.0 + s.eRest:
                y[i] = 0.88810600 * 0 + 0.11189400 * v_pp + v_ce
            else:
                v_o = float(interp_throw(self.E_sigmaTab, self.sigma_tab[:, i], energy))
                y[i] = 0.88810600 * v_o + 0.11189400 * v_pp + v_ce
        return y

    def sigma_table(self):
        """sigma_l(eGrid[n]) for every grid point: nEnergies x (nPN+1)."""
        return np.stack([self.sigma_at_energy(e) for e in self.eGrid])


# ---------------------------------------------------------------------------
# proton cross-section moments (MaterialParametersProtons.jl)
# ---------------------------------------------------------------------------
# scattering angles (degrees) of the tabulated elastic cross sections
# (MaterialParametersProtons.jl:26)
OMEGA_SIGMA_TAB = np.arange(5.0, 181.0, 5.0)


def legendre_table(lmax, x):
    """P_l(x) for l = 0..lmax (rows) at the points x (columns), by recurrence."""
    x = np.asarray(x, dtype=np.float64)
    P = np.zeros((lmax + 1, x.size))
    P[0] = 1.0
    if lmax >= 1:
        P[1] = x
    for l in range(1, lmax):
        P[l + 1] = ((2 * l + 1) * x * P[l] - l * P[l - 1]) / (l + 1)
    return P


def integrate_xs_poly(N, mu, sigma):
    """Legendre moments of tabulated cross sections (``integrateXS_Poly``,
    ``MaterialParametersProtons.jl:2812-2831``): xi[n, i] = -sum_k (mu_{k+1} -
    mu_k)/2 (sigma[n, k] P_i(mu_k) + sigma[n, k+1] P_i(mu_{k+1})), i < N."""
    P = legendre_table(N - 1, mu)                      # N x nMu
    dmu = np.diff(mu)                                   # nMu - 1
    t = sigma[:, :-1, None] * P.T[None, :-1, :] + sigma[:, 1:, None] * P.T[None, 1:, :]
    return -0.5 * np.einsum("k,nki->ni", dmu, t)


@functools.lru_cache(maxsize=1)
def proton_moment_tables():
    """(sigma_pp, sigma_ce): 40 Legendre moments on the PSTAR energies of the
    proton-proton nuclear elastic cross section (``Proton_Proton_nuclear``,
    ``MaterialParametersProtons.jl:2834-2849``: 0.315 E^-1.126 + 3.78e-6 E for
    E >= 10 MeV kinetic, isotropic) and of single Coulomb scattering on water
    (``Rutherford``, ``:2888-2902``, with the reference's unit conventions)."""
    t = load_tables("Protons")
    E = np.asarray(t["E_tab_PSTAR"], dtype=np.float64)      # kinetic MeV
    mu = np.cos(np.radians(OMEGA_SIGMA_TAB))
    s_pp = np.where(E >= 10, 0.315 * E ** (-1.126) + 3.78e-6 * E, 0.0)
    sig_pp = np.repeat(s_pp[:, None], OMEGA_SIGMA_TAB.size, axis=1)
    z_eff, z_p = 7.42, 1.0
    e, eps0 = 1.602176634e-19, 8.8541878128e-12
    sig_ce = ((4 * math.pi * eps0) ** -2 * (z_eff * z_p * e ** 2 / (4 * E[:, None])) ** 2
              / np.sin(np.radians(OMEGA_SIGMA_TAB[None, :] / 2)) ** 4)
    return integrate_xs_poly(40, mu, sig_pp), integrate_xs_poly(40, mu, sig_ce)


# ---------------------------------------------------------------------------
# P_N system (PNSystem.jl)
# ---------------------------------------------------------------------------
def global_index(l, k):
    """``GlobalIndex`` (PNSystem.jl:124-128), 0-based."""
    return l * l + k + l


def _A(l, k): return math.sqrt(((l - k + 1) * (l + k + 1)) / ((2 * l + 3) * (2 * l + 1)))
def _B(l, k): return math.sqrt(((l - k) * (l + k)) / ((2 * l + 1) * (2 * l - 1)))
def _C(l, k): return math.sqrt(((l + k + 1) * (l + k + 2)) / ((2 * l + 3) * (2 * l + 1)))
def _D(l, k): return math.sqrt(((l - k) * (l - k - 1)) / ((2 * l + 1) * (2 * l - 1)))
def _E(l, k): return math.sqrt(((l - k + 1) * (l - k + 2)) / ((2 * l + 3) * (2 * l + 1)))
def _F(l, k): return math.sqrt(((l + k) * (l + k - 1)) / ((2 * l + 1) * (2 * l - 1)))


def _Ct(l, k):
    if k < 0:
        return 0.0
    return math.sqrt(2) * _C(l, k) if k == 0 else _C(l, k)


def _Dt(l, k):
    if k < 0:
        return 0.0
    return math.sqrt(2) * _D(l, k) if k == 0 else _D(l, k)


def _Et(l, k): return math.sqrt(2) * _E(l, k) if k == 1 else _E(l, k)
def _Ft(l, k): return math.sqrt(2) * _F(l, k) if k == 1 else _F(l, k)
def _sgn(k): return 1 if k >= 0 else -1
def _kp(k): return k + _sgn(k)
def _km(k): return k - _sgn(k)


def _safe(fn, l, k):
    # the reference evaluates the coefficient functions only where the target
    # index is in range; out-of-range l gives a negative sqrt argument there
    try:
        return fn(l, k)
    except ValueError:
        return float("nan")


def pn_matrices_full(N):
    """Dense Ax, Ay, Az of size (N+1)^2 (``PNSystem.jl:142-194``)."""
    n = global_index(N, N) + 1
    Ax = np.zeros((n, n)); Ay = np.zeros((n, n)); Az = np.zeros((n, n))
    for l in range(N + 1):
        for k in range(-l, l + 1):
            i = global_index(l, k)

            def put(M, j, v):
                if 0 <= j < n:
                    M[i, j] = v
            if k != -1:
                put(Ax, global_index(l - 1, _km(k)), 0.5 * _safe(_Ct, l - 1, abs(k) - 1))
                put(Ax, global_index(l + 1, _km(k)), -0.5 * _safe(_Dt, l + 1, abs(k) - 1))
            put(Ax, global_index(l - 1, _kp(k)), -0.5 * _safe(_Et, l - 1, abs(k) + 1))
            put(Ax, global_index(l + 1, _kp(k)), 0.5 * _safe(_Ft, l + 1, abs(k) + 1))
            if k != 1:
                put(Ay, global_index(l - 1, -_km(k)), -0.5 * _sgn(k) * _safe(_Ct, l - 1, abs(k) - 1))
                put(Ay, global_index(l + 1, -_km(k)), 0.5 * _sgn(k) * _safe(_Dt, l + 1, abs(k) - 1))
            put(Ay, global_index(l - 1, -_kp(k)), -0.5 * _sgn(k) * _safe(_Et, l - 1, abs(k) + 1))
            put(Ay, global_index(l + 1, -_kp(k)), 0.5 * _sgn(k) * _safe(_Ft, l + 1, abs(k) + 1))
            put(Az, global_index(l - 1, k), _safe(_A, l - 1, k))
            put(Az, global_index(l + 1, k), _safe(_B, l + 1, k))
    for M in (Ax, Ay, Az):
        if np.isnan(M).any():
            raise AssertionError("P_N coefficient evaluated out of range")
    return Ax, Ay, Az


def moment_set(N, kind="subset"):
    """(l, k) pairs and their positions in the reference's full ordering."""
    if kind == "full":
        lk = [(l, k) for l in range(N + 1) for k in range(-l, l + 1)]
    elif kind == "subset":
        lk = [(l, k) for l in range(N + 1) for k in range(0, l + 1)]
    else:
        raise ValueError(kind)
    full_idx = np.array([global_index(l, k) for l, k in lk])
    degree = np.array([l for l, _ in lk])
    return lk, full_idx, degree


def abs_matrix(A, drop=1e-10):
    """Roe matrix V |Lambda| V^-1 with entries |v| <= drop removed (``SolverCSD.jl:84-97``).

    A is symmetric, so the eigen-decomposition is orthogonal (``eigh``).
    """
    lam, V = np.linalg.eigh(A)
    out = (V * np.abs(lam)) @ V.T
    out = 0.5 * (out + out.T)
    out[np.abs(out) <= drop] = 0.0
    return out


# ---------------------------------------------------------------------------
# quadrature and transforms
# ---------------------------------------------------------------------------
def quadrature(n):
    """Tensor Gauss-Legendre x equispaced-phi quadrature (``standardQuadrature.jl:7-37``).

    2 n^2 points; point q = j*n + k for mu-node k and phi-node j, phi_j =
    (j + 1/2) pi / n.  Returns (Q x 3 directions, weights summing to 4 pi).
    """
    mu, gw = np.polynomial.legendre.leggauss(n)
    phi = (np.arange(2 * n) + 0.5) * math.pi / n
    sq = np.sqrt(1.0 - mu ** 2)
    pts = np.empty((2 * n * n, 3))
    w = np.empty(2 * n * n)
    for j in range(2 * n):
        sl = slice(j * n, (j + 1) * n)
        pts[sl, 0] = sq * math.cos(phi[j])
        pts[sl, 1] = sq * math.sin(phi[j])
        pts[sl, 2] = mu
        w[sl] = 2.0 * math.pi / n * gw * 0.5
    return pts, w, mu, phi


def _assoc_legendre(lmax, x):
    """P_l^m(x) for 0 <= m <= l <= lmax, without the Condon-Shortley phase."""
    P = np.zeros((lmax + 1, lmax + 1))
    s = math.sqrt(max(0.0, 1.0 - x * x))
    pmm = 1.0
    for m in range(lmax + 1):
        if m > 0:
            pmm *= (2 * m - 1) * s
        P[m, m] = pmm
        if m + 1 <= lmax:
            P[m + 1, m] = x * (2 * m + 1) * pmm
        for l in range(m + 2, lmax + 1):
            P[l, m] = ((2 * l - 1) * x * P[l - 1, m] - (l + m - 1) * P[l - 2, m]) / (l - m)
    return P


def real_sph(mu, phi, lk, cs_phase=False):
    """Real spherical harmonics Y_{l,k}(mu, phi) (``utils.jl:12-41``).

    Orthonormal on the sphere: Y_l0 = N_l0 P_l; Y_lk = sqrt2 N_l|k| P_l^|k| cos(k phi)
    for k > 0 and sqrt2 N_l|k| P_l^|k| sin(|k| phi) for k < 0, with
    N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!).  ``cs_phase`` multiplies by (-1)^|k|.
    """
    lmax = max(l for l, _ in lk)
    P = _assoc_legendre(lmax, mu)
    out = np.empty(len(lk))
    for t, (l, k) in enumerate(lk):
        m = abs(k)
        nrm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.exp(math.lgamma(l - m + 1) - math.lgamma(l + m + 1)))
        v = nrm * P[l, m]
        if k > 0:
            v *= math.sqrt(2) * math.cos(k * phi)
        elif k < 0:
            v *= math.sqrt(2) * math.sin(m * phi)
        if cs_phase and (m % 2 == 1):
            v = -v
        out[t] = v
    return out


def trafo_matrices(N, lk, cs_phase=False):
    """O (Q x Nm) = Y(Omega_q) and M = O^T diag(w) (``Quadrature.jl:173-208``)."""
    n = N + 1
    pts, w, mu, phi = quadrature(n)
    O = np.empty((2 * n * n, len(lk)))
    for k in range(n):
        for j in range(2 * n):
            O[j * n + k] = real_sph(mu[k], phi[j], lk, cs_phase)
    M = O.T * w[None, :]
    return O, M, pts, w


# ---------------------------------------------------------------------------
# grid helpers
# ---------------------------------------------------------------------------
def boundary_index(nx, ny):
    """Ghost-cell list in the reference's order, corners twice (``SolverCSD.jl:108-131``)."""
    out = []
    for i in range(nx):
        out.append(i * ny + 0)
        out.append(i * ny + ny - 1)
    for j in range(ny):
        out.append(0 * ny + j)
        out.append((nx - 1) * ny + j)
    return np.array(out, dtype=np.int64)


def boundary_restore_cells(nx, ny):
    """Cells whose ordinate-1 flux the reference restores after the sweep.

    ``psi[obj.boundaryIdx] .= psiBC`` (``SolverCSD.jl:1002,1005``) indexes the
    column-major (nx, ny, nq) array ``psi`` linearly with the row-major cell
    index c = (i-1) ny + j.  Linear index c (<= nx ny) addresses ordinate 1 at
    (i', j') = ((c-1) mod nx + 1, (c-1) div nx + 1).  Returns the 0-based cell
    indices i'*ny + j' of those entries (unique, sorted).
    """
    lin = boundary_index(nx, ny)               # 0-based c
    ip = lin % nx
    jp = lin // nx
    return np.unique(ip * ny + jp)


def normpdf(x, mu, sigma):
    """``utils.jl:43-45``."""
    return 1.0 / (sigma * math.sqrt(2 * math.pi)) * np.exp(-(x - mu) ** 2 / 2 / sigma ** 2)


def beam_angular(s, omega):
    """Angular factor of ``PsiBeam`` (``SolverCSD.jl:240-276``)."""
    return (1e5 * np.exp(-s.sigmaO1Inv * (s.Omega1 - omega[:, 0]) ** 2)
            * np.exp(-s.sigmaO3Inv * (s.Omega3 - omega[:, 2]) ** 2))


def beam_spatial(s, x, y):
    if s.beamModel == "normpdf":
        return normpdf(x, s.x0, s.sigmaX) * normpdf(y, s.y0, s.sigmaY)
    return np.exp(-s.sigmaXInv * (x - s.x0) ** 2) * np.exp(-s.sigmaYInv * (y - s.y0) ** 2)


def beam_energy_factor(s, E, S_n):
    """Energy-dependent factor of ``PsiBeam``: S[n] * densityMin (* exp term).

    The water-beam branch multiplies by a local ``densityMin = 1.0``
    (``SolverCSD.jl:267,271``); the lung branch by ``settings.densityMin``.
    """
    if s.beamModel == "normpdf":
        return S_n * 1.0
    return np.exp(-s.sigmaEInv * (s.eMax - E) ** 2) * S_n * s.densityMin


def setup_ic(s, omega):
    """Initial uncollided flux of ``SetupIC`` (``SolverCSD.jl:155-185``) on the
    reduced ordinates ``omega`` (Q x 3): C x Q (Fortran order), or None when it
    is zero.

    "waterBeam" starts with the beam already inside the slab: cells i, j in
    3..n-2 (1-based) hold 1e5 exp(-1e4 (Omega1 - w_x)^2) exp(-1e4 (Omega3 -
    w_z)^2) normpdf(x, x0, 0.1) normpdf(y, y0, 0.1) (the angular widths and
    the 0.1 cm spot are fixed there, not taken from the settings).  The other
    beam problems start from zero (``IC``, ``settings.jl:341-365``, is zero for
    them).
    """
    if s.problem != "waterBeam":
        return None
    nx, ny = s.NCellsX, s.NCellsY
    ang = (1e5 * np.exp(-10000.0 * (s.Omega1 - omega[:, 0]) ** 2)
           * np.exp(-10000.0 * (s.Omega3 - omega[:, 2]) ** 2))
    psi = np.zeros((nx, ny, len(omega)))
    i = np.arange(2, nx - 2)
    j = np.arange(2, ny - 2)
    space = normpdf(s.xMid[i], s.x0, 0.1)[:, None] * normpdf(s.yMid[j], s.y0, 0.1)[None, :]
    psi[2:nx - 2, 2:ny - 2, :] = space[:, :, None] * ang[None, None, :]
    return np.asfortranarray(psi.reshape(nx * ny, len(omega)))
