"""Problem configuration: the reference's ``Settings`` struct plus explicit fields.

Mirrors ``code/settings.jl:5-338``.  Field names follow the reference so that a
caller of ``Settings(Nx, Ny, r, problem, particle)`` finds the same attributes
(``NCellsX``, ``dx``, ``eMax``, ``nPN``, ``xMid``, ``density``, ``ϑ`` spelled
``theta`` ...).  What the reference hard-codes is an explicit field here:

* ``nPN``        -- P_N degree (reference: 21 at ``settings.jl:332``);
* ``eMin``       -- lower energy bound in MeV (reference: 0.001 at ``CSD.jl:58``);
* ``nEnergies``  -- number of energy grid points (reference derives it from
                    ``ceil(eMax/dE)`` at ``CSD.jl:68``); ``None`` keeps that rule;
* beam shape     -- ``beamModel`` ("normpdf" as in the water-beam presets,
                    ``SolverCSD.jl:263-271``; "exp" as in the lung presets,
                    ``SolverCSD.jl:242-247,275``), ``sigmaX``/``sigmaY`` (cm),
                    ``sigmaO1Inv``/``sigmaO3Inv``/``sigmaEInv``;
* ``rMax``       -- rank cap of the adaptive integrator (reference: ``r``,
                    ``SolverCSD.jl:3373``);
* ``phantomImage`` -- path of the image the "lung" / "liver" presets read
                    (reference: ``Lung.png`` / ``liver_cut.jpg`` in the working
                    directory, ``settings.jl:266-310``); ``None`` keeps a unit
                    density;
* ``csPhase``    -- sign convention of the real spherical harmonics.  With
                    ``False`` the ordinate transform satisfies
                    ``M diag(Omega_x) O = +Ax`` (physically consistent x-streaming);
                    ``True`` gives ``-Ax``.  The reference inherits this from
                    SphericalHarmonics.jl 0.1.19 (``code/utils.jl:4``).
"""
import math

import numpy as np


class Settings:
    """Grid, energy, P_N, rank and beam configuration (``settings.jl:67-338``)."""

    def __init__(self, Nx=102, Ny=102, r=15, problem="LineSource", particle="Electrons",
                 **overrides):
        # settings.jl:68-73
        if particle == "Protons":
            eRest = 938.26
        elif particle == "Electrons":
            eRest = 0.5
        else:
            raise ValueError(f"unknown particle {particle!r}")
        self.Nx, self.Ny = int(Nx), int(Ny)
        self.NCellsX, self.NCellsY = self.Nx - 1, self.Ny - 1
        if self.NCellsX < 3 or self.NCellsY < 3:
            raise ValueError("need at least 3 cells per direction")
        self.problem, self.particle, self.eRest = problem, particle, eRest
        self.r = int(r)
        # settings.jl:78-96 defaults
        self.a, self.b, self.c, self.d = 0.0, 14.5, 0.0, 14.5
        self.density = np.ones((self.NCellsX, self.NCellsY))
        self.sigmaS, sigmaA = 0.0, 0.0
        self.eMax = 1.0
        self.x0, self.y0 = 0.5 * self.b, 1.0 * self.d
        self.Omega1, self.Omega3 = -1.0, -1.0
        self.densityMin = 0.2
        self.thetaIndex = 1.0
        self.theta = 1e-2
        self.cfl = 1.5
        self.cEta = 5.0
        # explicit fields for the reference's hard-coded constants
        self.nPN = 21
        self.eMin = 0.001
        self.nEnergies = None
        self.beamModel = "exp"
        self.sigmaX, self.sigmaY = 0.1, 0.1
        self.sigmaO1Inv, self.sigmaO3Inv = 0.0, 75.0
        self.sigmaXInv, self.sigmaYInv, self.sigmaEInv = 20.0, 20.0, 100.0
        self.rMax = None
        self.csPhase = False
        self.floorPsi = 1e-17          # SolverCSD.jl:906
        self.phantomImage = overrides.pop("phantomImage", None)

        if problem in ("waterBeam", "waterBeamElectrons"):
            # settings.jl:175-242 (density inserts of those presets omitted: the
            # builder's configs set their own density maps)
            self.a, self.b, self.c = 0.0, 2.0, 0.0
            self.d = 3.0 if problem == "waterBeamElectrons" else 8.0
            self.sigmaS = 1.0
            self.cfl = 0.99 / math.sqrt(2) * (2.0 if problem == "waterBeamElectrons" else 120.5)
            eKin = 21.0 if problem == "waterBeamElectrons" else 85.0
            self.eMax = eKin + eRest
            self.thetaIndex, self.theta = 0.0, 0.3
            self.x0, self.y0 = 0.5 * self.b, 0.05 * self.d
            self.Omega1, self.Omega3 = -0.0, 1.0
            self.beamModel = "normpdf"
            self.sigmaO1Inv = self.sigmaO3Inv = 10000.0
            self.densityMin = 1.0
        elif problem == "LineSource":
            self.a, self.b, self.c, self.d = -1.5, 1.5, -1.5, 1.5
            self.sigmaS = 1.0
            self.cfl = 0.99 / math.sqrt(2)
            self.thetaIndex, self.theta = 0.0, 0.3
        elif problem in ("lung", "liver"):
            # settings.jl:266-310: density from an image, 1.85 g/cc at white,
            # floored at densityMin; then the preset's domain and beam
            self.densityMin = 0.05
            if self.phantomImage is not None:
                self.density = image_density(self.phantomImage, self.NCellsX, self.NCellsY,
                                             self.densityMin)
            self.cfl = 1.5
            self.Omega1, self.Omega3 = -1.0, -1.0
            if problem == "lung":
                self.b = self.d = 14.5
                self.eMax = math.sqrt(21.0 ** 2 + eRest ** 2)
                self.x0, self.y0 = 0.5 * self.b, 1.0 * self.d
            else:
                self.b = self.d = 35.0
                self.eMax = 60.0
                self.x0, self.y0 = 1.0 * self.b, 0.35 * self.d
        elif problem in ("slab", "phantom"):
            # "lung"-style beam (SolverCSD.jl:242-247): exp profiles, energy factor
            pass
        else:
            raise ValueError(f"problem {problem!r} is not provided by this build")

        for k, v in overrides.items():
            if not hasattr(self, k):
                raise AttributeError(f"Settings has no field {k!r}")
            setattr(self, k, v)
        self.sigmaT = sigmaA + self.sigmaS
        self._grid()

    def _grid(self):
        """settings.jl:315-329: cell centres on the domain edges, ghost shift."""
        x = np.linspace(self.a, self.b, self.NCellsX)
        self.dx = x[1] - x[0]
        x = np.concatenate([[x[0] - self.dx], x]) + self.dx / 2
        self.x = x
        self.xMid = x[:-1] + 0.5 * self.dx
        y = np.linspace(self.c, self.d, self.NCellsY)
        self.dy = y[1] - y[0]
        y = np.concatenate([[y[0] - self.dy], y]) + self.dy / 2
        self.y = y
        self.yMid = y[:-1] + 0.5 * self.dy
        self.density = np.asarray(self.density, dtype=np.float64)
        if self.density.shape != (self.NCellsX, self.NCellsY):
            raise ValueError("density must be NCellsX x NCellsY")
        self.dE = self.cfl * min(self.dx, self.dy) * float(self.density.min())

    def set_density(self, density):
        """Replace the density map (NCellsX x NCellsY, g/cc) and refresh dE."""
        self.density = np.asarray(density, dtype=np.float64)
        self._grid()

    @property
    def rank_cap(self):
        return self.r if self.rMax is None else int(self.rMax)


# ---------------------------------------------------------------------------
# image phantoms (settings.jl:266-310)
# ---------------------------------------------------------------------------
def image_gray(path):
    """``Float64.(Gray.(load(path)))``: an (rows x columns) array in [0, 1].

    8-bit images; gray values are k/255.  Colour images use the ITU-R BT.601
    luma weights in single precision rounded back to 8 bits, as Colors.jl's
    RGB-to-Gray conversion; where R = G = B (the shipped Lung.png) every
    weighting gives k/255 exactly, so only colour inputs depend on the rule.
    """
    from PIL import Image
    im = Image.open(path)
    if im.mode == "L":
        return np.asarray(im, dtype=np.float64) / 255.0
    a = np.asarray(im.convert("RGB"), dtype=np.float32) / np.float32(255.0)
    g = np.float32(0.299) * a[..., 0] + np.float32(0.587) * a[..., 1] + np.float32(0.114) * a[..., 2]
    return np.rint(g.astype(np.float64) * 255.0) / 255.0


def image_density(image, NCellsX, NCellsY, density_min, scale=1.85):
    """density[i, j] = max(scale * img[floor(i/NCellsX*nx), floor(j/NCellsY*ny)],
    density_min) with 1-based i, j and image indices (settings.jl:276-281);
    `image` is a path or an (nx x ny) gray array."""
    img = image_gray(image) if isinstance(image, (str, bytes)) or hasattr(image, "__fspath__") \
        else np.asarray(image, dtype=np.float64)
    nx, ny = img.shape[:2]
    ii = [int(math.floor(i / NCellsX * nx)) - 1 for i in range(1, NCellsX + 1)]
    jj = [int(math.floor(j / NCellsY * ny)) - 1 for j in range(1, NCellsY + 1)]
    if min(ii) < 0 or min(jj) < 0:
        raise ValueError("image smaller than the grid: the reference's index map starts at 0")
    return np.maximum(scale * img[np.ix_(ii, jj)], density_min)


# ---------------------------------------------------------------------------
# seeded phantoms (the builder's synthetic densities; no reference counterpart)
# ---------------------------------------------------------------------------
def ellipse_phantom(nx, ny, a, b, c, d, seed, n_lung=(3, 6), n_bone=(1, 3),
                    rho_bg=1.0, rho_lung=0.25, rho_bone=1.85):
    """Heterogeneous density: background, random lung ellipses, then bone ellipses.

    Centres uniform in the inner 80% of the domain, semi-axes U[0.5, 2.0] cm
    scaled to the domain, rotation U[0, pi).  Later ellipses override earlier.
    ``numpy.random.default_rng(seed)`` drives all draws in a fixed order.
    """
    rng = np.random.default_rng(seed)
    xs = np.linspace(a, b, nx)
    ys = np.linspace(c, d, ny)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    rho = np.full((nx, ny), rho_bg)
    scale = min(b - a, d - c) / 10.0
    for count, val in ((rng.integers(n_lung[0], n_lung[1] + 1), rho_lung),
                       (rng.integers(n_bone[0], n_bone[1] + 1), rho_bone)):
        for _ in range(int(count)):
            cx = rng.uniform(a + 0.1 * (b - a), b - 0.1 * (b - a))
            cy = rng.uniform(c + 0.1 * (d - c), d - 0.1 * (d - c))
            sa, sb = rng.uniform(0.5, 2.0, size=2) * scale
            th = rng.uniform(0.0, math.pi)
            u = (X - cx) * math.cos(th) + (Y - cy) * math.sin(th)
            v = -(X - cx) * math.sin(th) + (Y - cy) * math.cos(th)
            rho[(u / sa) ** 2 + (v / sb) ** 2 <= 1.0] = val
    return rho


# ---------------------------------------------------------------------------
# the configurations this build measures
# ---------------------------------------------------------------------------
def water_slab_pencil_beam(n=80, N=9, r=20, steps=None):
    """Electron pencil beam, 21 MeV kinetic, into a homogeneous 5x5 cm water slab.

    80x80 cells, P_9, fixed rank 20, energy grid down to 0.1 MeV.  ``steps=None``
    uses the reference's grid rule ``nEnergies = ceil(eMax/dE)`` with
    ``dE = cfl min(dx,dy) min(rho)`` and the water-beam cfl 0.99*sqrt(2)
    (``CSD.jl:68``, ``settings.jl:182,329``): 243 points, step/dx = 0.65.  A
    fixed 100 steps gives step/dx = 1.56, where the explicit upwind scheme
    diverges.  Beam centred
    at x = 2.5 cm on the y = 0 edge, direction +y (Omega_z = 1), lateral Gaussian
    sigma 0.5 cm, angular width as the reference's water-beam presets
    (sigma_Omega^-1 = 1e4).  eMax = eKin + eRest as in ``settings.jl:183-184``.
    """
    return Settings(n + 1, n + 1, r, "waterBeamElectrons", "Electrons",
                    a=0.0, b=5.0, c=0.0, d=5.0, x0=2.5, y0=0.0, nPN=N, eMin=0.1,
                    nEnergies=None if steps is None else steps + 1,
                    sigmaX=0.5, sigmaY=0.5, eMax=21.0 + 0.5)


def heterogeneous_phantom(n=200, N=21, steps=None, seed=0, r=40, eKin=15.0):
    """Rank-adaptive case: 10x10 cm ellipse phantom, 15 MeV electrons, P_21.

    200x200 cells, energy grid down to 0.1 MeV by the reference's rule with the
    lung preset's cfl 1.5 (``steps=None``; 823 points for the default phantom,
    step/(dx rho_min) = 0.70; a fixed 200 steps gives 2.9 in the 0.25 g/cc
    cells and diverges), truncation tolerance 1e-2
    (relative, ``thetaIndex = 1``), rank cap ``r`` (initial rank r/2 as the
    reference's BUG solver, ``SolverCSD.jl:2013``).  Pencil beam at x = 5 cm
    on the y = 0 edge, direction +y, lateral sigma 0.5 cm.
    """
    s = Settings(n + 1, n + 1, r, "waterBeamElectrons", "Electrons",
                 a=0.0, b=10.0, c=0.0, d=10.0, x0=5.0, y0=0.0, nPN=N, eMin=0.1,
                 nEnergies=None if steps is None else steps + 1,
                 sigmaX=0.5, sigmaY=0.5, eMax=eKin + 0.5,
                 theta=1e-2, thetaIndex=1.0, cfl=1.5)
    s.set_density(ellipse_phantom(n, n, s.a, s.b, s.c, s.d, seed))
    return s
