"""Operands of the 3D energy loop (the 3D proton configuration).

The reference snapshot is two-dimensional (``SolverCSD.jl:413-414,1027,1040``
pair the x and y cells with Omega_x and Omega_z; ``Ay`` is built by
``PNSystem.jl:142-194`` but never used).  This module extends the same
discretisation to three dimensions, as the extension the 3D configuration
asks for:

* cells c = (i*ny + j)*nz + k (k fastest); with nz = 1 this is the 2D
  numbering c = i*ny + j;
* spatial axes pair with angular components as in 2D plus one: i <-> Omega_x
  (``Ax``), j <-> Omega_z (``Az``), k <-> Omega_y (``Ay``); a 2D problem is the
  nz = 1 case;
* all (N+1)^2 real-harmonic moments (in 3D the k >= 0 subset no longer
  decouples), Roe matrices |A| = V |Lambda| V^-1 of all three;
* ghost cells: every cell on the six faces (rows zeroed in X, psi set by the
  beam source), as the 2D ghost frame;
* the beam enters the j = 0 face travelling along +Omega_z, with the 2D water
  beam's separable profile extended by the third axis:
  psi = PsiBeam_angular(Omega) G(x) G(y) G(z) S[n], G = normpdf(., centre,
  sigma);
* no ordinate-1 restore (the 2D reference's transposed-index quirk has no 3D
  counterpart).

Layouts match the 2D problem dict of ``titus_amd.problem`` plus ``nz``,
``dz``, ``Ay`` and ``absAy``.
"""
import functools
import math

import numpy as np

from . import physics as ph
from .problem import INTEGRATOR_BUG, INTEGRATOR_FIXED


class Settings3D:
    """Grid, energy, P_N, rank and beam of a 3D solve (fields as ``Settings``)."""

    def __init__(self, n=(64, 64, 64), h=0.1, r=30, particle="Protons", **kw):
        self.NCellsX, self.NCellsY, self.NCellsZ = (int(v) for v in n)
        if min(n) < 3:
            raise ValueError("need at least 3 cells per direction")
        self.particle = particle
        self.eRest = 938.26 if particle == "Protons" else 0.5
        self.dx = self.dy = self.dz = float(h)
        # cell centres on [0, (n-1) h] (settings.jl:316-320: centres on the edges)
        self.xMid = np.arange(self.NCellsX) * self.dx
        self.yMid = np.arange(self.NCellsY) * self.dy
        self.zMid = np.arange(self.NCellsZ) * self.dz
        self.r = int(r)
        self.nPN = 13
        self.eKin = 90.0
        self.eMin = 0.001
        self.nEnergies = 301
        self.cfl = 1.5
        self.theta, self.thetaIndex, self.cEta = 1e-2, 1.0, 5.0
        self.x0 = 0.5 * (self.NCellsX - 1) * self.dx
        self.y0 = 0.0
        self.z0 = 0.5 * (self.NCellsZ - 1) * self.dz
        self.sigma = 0.4
        self.Omega1, self.Omega3 = 0.0, 1.0
        self.sigmaO1Inv = self.sigmaO3Inv = 1000.0
        self.floorPsi = 1e-17
        self.csPhase = False
        self.beamModel = "normpdf"
        self.density = np.ones((self.NCellsX, self.NCellsY, self.NCellsZ))
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(f"unknown 3D setting {k!r}")
            setattr(self, k, v)
        self.eMax = self.eKin + self.eRest
        self.dE = None                      # the grid comes from nEnergies

    @property
    def rank_cap(self):
        return self.r


def smooth_random_density(n, seed, sigma_vox=2.0, lo=0.3, hi=1.6):
    """Seeded N(0,1) noise, Gaussian-filtered (sigma_vox voxels, periodic
    boundary), min-max mapped to [lo, hi] g/cc."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal(n), sigma=sigma_vox, mode="wrap")
    f = (f - f.min()) / (f.max() - f.min())
    return lo + (hi - lo) * f


@functools.lru_cache(maxsize=4)
def angular_operands3d(N, cs_phase):
    """Full-moment P_N, Roe and transform matrices of degree N."""
    lk, full_idx, degree = ph.moment_set(N, "full")
    Axf, Ayf, Azf = ph.pn_matrices_full(N)
    Ax, Ay, Az = (A[np.ix_(full_idx, full_idx)] for A in (Axf, Ayf, Azf))
    absAx, absAy, absAz = ph.abs_matrix(Ax), ph.abs_matrix(Ay), ph.abs_matrix(Az)
    O, M, pts, w = ph.trafo_matrices(N, lk, cs_phase=cs_phase)
    out = (lk, degree, Ax, Ay, Az, absAx, absAy, absAz, O, M, pts, w)
    for a in out[2:]:
        a.setflags(write=False)
    return out


def face_cells(nx, ny, nz):
    """All cells on the six faces (0-based c = (i*ny + j)*nz + k), sorted."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    on = (i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1) | (k == 0) | (k == nz - 1)
    return np.nonzero(on.reshape(-1))[0]


def build_problem3d(s, integrator=INTEGRATOR_FIXED):
    nx, ny, nz = s.NCellsX, s.NCellsY, s.NCellsZ
    C = nx * ny * nz
    N = s.nPN
    csd = ph.CSDGrid(s)
    lk, degree, Ax, Ay, Az, absAx, absAy, absAz, O, M, pts, w = angular_operands3d(N, bool(s.csPhase))

    psi_beam = ph.beam_angular(s, pts)
    idx = np.nonzero(psi_beam > s.floorPsi * psi_beam.max())[0]
    omega = pts[idx]
    Mred = M[:, idx]

    bcells = face_cells(nx, ny, nz)
    bi = bcells // (ny * nz)
    bj = (bcells // nz) % ny
    bk = bcells % nz
    spatial = (ph.normpdf(s.xMid[bi], s.x0, s.sigma) * ph.normpdf(s.yMid[bj], s.y0, s.sigma)
               * ph.normpdf(s.zMid[bk], s.z0, s.sigma))
    bc_base = np.asfortranarray(spatial[:, None] * ph.beam_angular(s, omega)[None, :])
    bc_scale = np.array([ph.beam_energy_factor(s, csd.eGrid[n], csd.S[n])
                         for n in range(csd.nEnergies)])

    interior = np.ones(C, dtype=bool)
    interior[bcells] = False

    r0 = s.r // 2 if integrator == INTEGRATOR_BUG else s.r
    if r0 < 1 or r0 > C or r0 > len(lk):
        raise ValueError("rank out of range")
    X0 = np.zeros((C, r0), order="F")
    X0[np.arange(r0), np.arange(r0)] = 1.0
    X0[bcells, :] = 0.0
    W0 = np.zeros((len(lk), r0), order="F")
    W0[np.arange(r0), np.arange(r0)] = 1.0
    S0 = np.zeros((r0, r0), order="F")

    F = np.asfortranarray
    return dict(
        nx=nx, ny=ny, nz=nz, C=C, dx=s.dx, dy=s.dy, dz=s.dz,
        density=np.ascontiguousarray(np.asarray(s.density, dtype=np.float64).reshape(C)),
        N=N, Nm=len(lk), degree=degree.astype(np.int32), lk=lk,
        Ax=F(Ax), Ay=F(Ay), Az=F(Az), absAx=F(absAx), absAy=F(absAy), absAz=F(absAz),
        Q=len(idx), Qfull=len(w), omega=F(omega), M=F(Mred), O=F(O), ordinate_index=idx,
        nE=csd.nEnergies, dE=csd.dE, eGrid=csd.eGrid, Sst=csd.S,
        sigma=np.ascontiguousarray(csd.sigma_table()),
        bcells=bcells.astype(np.int32), bc_base=bc_base, bc_scale=bc_scale,
        restore_cells=np.zeros(0, dtype=np.int32),
        interior=np.nonzero(interior)[0].astype(np.int32),
        integrator=integrator, r=r0, rank_cap=int(s.rank_cap),
        theta=float(s.theta), thetaIndex=float(s.thetaIndex), cEta=float(s.cEta),
        X0=X0, W0=W0, S0=S0,
    )


def proton_cube(n=64, N=13, r=30, steps=300, seed=0, sigma_vox=2.0, **kw):
    """The 3D proton configuration: 90 MeV protons, 0.4 cm spot, into an n^3
    grid of 0.1 cm voxels with a seeded smooth density in [0.3, 1.6] g/cc,
    P_N with all (N+1)^2 moments, fixed rank r, `steps` energy steps."""
    s = Settings3D(n=(n, n, n), h=0.1, r=r, nPN=N, nEnergies=steps + 1, **kw)
    s.density = smooth_random_density((n, n, n), seed, sigma_vox)
    return s
