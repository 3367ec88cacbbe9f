"""ctypes binding of ``libtitus_hip.so`` (the C-ABI of ``include/titus_hip.h``).

The library is built in-tree (``python -m titus_amd.build``).  Loading fails
loudly when it is missing: there is no CPU fallback for the energy step.
"""
import ctypes
import os
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtitus_hip.so")

ABI_VERSION = 3
INTEGRATOR_FIXED, INTEGRATOR_BUG, INTEGRATOR_PARALLEL, INTEGRATOR_RK4 = 0, 1, 2, 3
PRECISION_FP64, PRECISION_MIXED, PRECISION_MIXED_FULL_FP32 = 0, 1, 2

_P = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int32)

ERRORS = {
    -1: "TITUS_E_ARG", -2: "TITUS_E_HIP", -3: "TITUS_E_NONFINITE", -4: "TITUS_E_RANK",
    -5: "TITUS_E_STATE", -6: "TITUS_E_NODEVICE",
}

# every symbol include/titus_hip.h declares
EXPORTED = ("titus_create", "titus_get_cfl", "titus_set_analytic_beam", "titus_set_rejection",
            "titus_get_rejections",
            "titus_step", "titus_solve", "titus_solve_batch", "titus_solve_batch_lockstep",
            "titus_progress", "titus_get_state",
            "titus_get_history", "titus_set_timing", "titus_kernel_stats", "titus_get_tuning",
            "titus_get_decisions", "titus_set_checkpoints", "titus_adjoint", "titus_adjoint_batch",
            "titus_adjoint_batch_lockstep",
            "titus_destroy", "titus_fcs_create", "titus_fcs_step", "titus_fcs_solve",
            "titus_fcs_get_state", "titus_fcs_set_timing", "titus_fcs_kernel_stats",
            "titus_fcs_get_plan", "titus_fcs_destroy", "titus_galerkin", "titus_galerkin_noalloc",
            "titus_last_error",
            "titus_abi_version")


class TitusError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class ProblemV1(ctypes.Structure):
    """``titus_problem_v1``."""
    _fields_ = [
        ("version", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("density", _P),
        ("nm", ctypes.c_int32), ("degree", _PI), ("Ax", _P), ("Az", _P), ("absAx", _P),
        ("absAz", _P),
        ("nq", ctypes.c_int32), ("omega", _P), ("M", _P),
        ("ne", ctypes.c_int32), ("dE", ctypes.c_double), ("stopping", _P),
        ("nsigma", ctypes.c_int32), ("sigma", _P),
        ("nb", ctypes.c_int32), ("bcells", _PI), ("bc_base", _P), ("bc_scale", _P),
        ("nrestore", ctypes.c_int32), ("restore_cells", _PI),
        ("ncand", ctypes.c_int32), ("cand_cells", _PI),
        ("integrator", ctypes.c_int32), ("precision", ctypes.c_int32), ("r", ctypes.c_int32),
        ("rank_cap", ctypes.c_int32), ("theta", ctypes.c_double),
        ("theta_index", ctypes.c_double), ("rank_tol", ctypes.c_double),
        ("X0", _P), ("S0", _P), ("W0", _P),
        ("nz", ctypes.c_int32), ("dz", ctypes.c_double), ("Ay", _P), ("absAy", _P),
        ("psi0", _P), ("flux_order", ctypes.c_int32),
    ]


class GridV1(ctypes.Structure):
    """``titus_grid_v1`` (inv_density is a device address)."""
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("inv_density", ctypes.c_void_p)]


class BeamV1(ctypes.Structure):
    """``titus_beam_v1`` (the RK4 integrator's analytic beam)."""
    _fields_ = [("x0", ctypes.c_double), ("y0", ctypes.c_double), ("Omega1", ctypes.c_double),
                ("Omega3", ctypes.c_double), ("xmid", _P), ("ymid", _P), ("tau", _P)]


class FcsProblemV1(ctypes.Structure):
    """``titus_fcs_problem_v1``."""
    _fields_ = [
        ("version", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("density", _P),
        ("nm", ctypes.c_int32), ("degree", _PI), ("Ax", _P), ("Az", _P), ("absAx", _P),
        ("absAz", _P),
        ("nq", ctypes.c_int32), ("omega", _P), ("M", _P),
        ("ne", ctypes.c_int32), ("dE", _P), ("stopping_mid", _P),
        ("nsigma", ctypes.c_int32), ("sigma", _P),
        ("nb", ctypes.c_int32), ("bcells", _PI), ("bc_base", _P), ("bc_scale", _P),
        ("psi0", _P),
    ]


_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m titus_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.titus_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ProblemV1), ctypes.c_int32]
    lib.titus_step.argtypes = [vp, ctypes.c_int32]
    lib.titus_get_cfl.argtypes = [vp, _P]
    lib.titus_set_rejection.argtypes = [vp, ctypes.c_double]
    lib.titus_set_analytic_beam.argtypes = [vp, ctypes.POINTER(BeamV1)]
    lib.titus_get_rejections.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    lib.titus_solve.argtypes = [vp, ctypes.c_int32]
    lib.titus_solve_batch.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.c_int32, _PI]
    lib.titus_solve_batch_lockstep.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_int32, _PI, ctypes.POINTER(ctypes.c_int64), _PI]
    lib.titus_progress.argtypes = [vp, _PI, _PI]
    lib.titus_get_state.argtypes = [vp, _P, _P, _P, _P, _P]
    lib.titus_get_history.argtypes = [vp, _PI, _P, ctypes.c_int32]
    lib.titus_set_timing.argtypes = [vp, ctypes.c_int32]
    lib.titus_kernel_stats.argtypes = [vp, ctypes.c_char_p, _P, ctypes.POINTER(ctypes.c_int64),
                                       _P, _P]
    lib.titus_get_tuning.argtypes = [vp, ctypes.c_char_p, _PI, ctypes.c_int32]
    lib.titus_get_decisions.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.POINTER(ctypes.c_int64)]
    lib.titus_set_checkpoints.argtypes = [vp, ctypes.c_int32]
    lib.titus_adjoint.argtypes = [vp, _P, _P]
    lib.titus_adjoint_batch_lockstep.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.POINTER(_P), ctypes.POINTER(_P),
                                                 _PI]
    lib.titus_adjoint_batch.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.c_int32,
                                        ctypes.POINTER(_P), ctypes.POINTER(_P), _PI]
    lib.titus_destroy.argtypes = [vp]
    lib.titus_destroy.restype = None
    lib.titus_fcs_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(FcsProblemV1), ctypes.c_int32]
    lib.titus_fcs_step.argtypes = [vp, ctypes.c_int32]
    lib.titus_fcs_solve.argtypes = [vp, ctypes.c_int32]
    lib.titus_fcs_get_state.argtypes = [vp, _P, _P, _P]
    lib.titus_fcs_set_timing.argtypes = [vp, ctypes.c_int32]
    lib.titus_fcs_kernel_stats.argtypes = [vp, ctypes.c_char_p, _P, ctypes.POINTER(ctypes.c_int64),
                                           _P, _P]
    lib.titus_fcs_get_plan.argtypes = [vp, _PI, ctypes.c_int32]
    lib.titus_fcs_destroy.argtypes = [vp]
    lib.titus_fcs_destroy.restype = None
    lib.titus_galerkin.argtypes = [ctypes.POINTER(GridV1), vp, ctypes.c_int32, ctypes.c_int32, vp, vp]
    lib.titus_galerkin.restype = ctypes.c_int
    lib.titus_galerkin_noalloc.argtypes = lib.titus_galerkin.argtypes
    lib.titus_galerkin_noalloc.restype = ctypes.c_int
    lib.titus_last_error.restype = ctypes.c_char_p
    lib.titus_abi_version.restype = ctypes.c_int32
    for name in ("titus_create", "titus_get_cfl", "titus_set_analytic_beam", "titus_set_rejection",
                 "titus_get_rejections",
                 "titus_step", "titus_solve", "titus_solve_batch", "titus_solve_batch_lockstep",
                 "titus_progress", "titus_get_state", "titus_get_history", "titus_set_timing",
                 "titus_kernel_stats", "titus_get_tuning", "titus_get_decisions",
                 "titus_set_checkpoints",
                 "titus_adjoint", "titus_adjoint_batch", "titus_adjoint_batch_lockstep",
                 "titus_fcs_create", "titus_fcs_step",
                 "titus_fcs_solve", "titus_fcs_get_state", "titus_fcs_set_timing",
                 "titus_fcs_kernel_stats", "titus_fcs_get_plan"):
        getattr(lib, name).restype = ctypes.c_int
    if lib.titus_abi_version() != ABI_VERSION:
        raise RuntimeError("libtitus_hip.so ABI version mismatch")
    _lib = lib
    return lib


def check(code):
    if code != 0:
        msg = load().titus_last_error()
        raise TitusError(code, msg.decode() if msg else "")


def _dp(a):
    return a.ctypes.data_as(_P)


def _ip(a):
    return a.ctypes.data_as(_PI)


def make_problem(P, integrator, precision=PRECISION_FP64, rank_tol=1e-7, initial=None):
    """Fill a ``ProblemV1`` from ``titus_amd.problem.build_problem`` output.

    Returns (struct, keepalive) -- the arrays must outlive ``titus_create``.
    """
    f64 = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float64))
    i32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    keep = {
        "density": f64(P["density"]), "degree": i32(P["degree"]),
        "Ax": f64(P["Ax"]), "Az": f64(P["Az"]), "absAx": f64(P["absAx"]), "absAz": f64(P["absAz"]),
        "omega": f64(P["omega"]), "M": f64(P["M"]), "stopping": f64(P["Sst"]),
        "sigma": np.ascontiguousarray(P["sigma"], dtype=np.float64),
        "bcells": i32(P["bcells"]), "bc_base": f64(P["bc_base"]), "bc_scale": f64(P["bc_scale"]),
        "restore": i32(P["restore_cells"]), "cand": i32(P["interior"]),
    }
    s = ProblemV1()
    s.version = ABI_VERSION
    s.nx, s.ny, s.dx, s.dy = P["nx"], P["ny"], P["dx"], P["dy"]
    s.density = _dp(keep["density"])
    s.nm, s.degree = P["Nm"], _ip(keep["degree"])
    s.Ax, s.Az = _dp(keep["Ax"]), _dp(keep["Az"])
    s.absAx, s.absAz = _dp(keep["absAx"]), _dp(keep["absAz"])
    s.nq, s.omega, s.M = P["Q"], _dp(keep["omega"]), _dp(keep["M"])
    s.ne, s.dE, s.stopping = P["nE"], P["dE"], _dp(keep["stopping"])
    s.nsigma, s.sigma = keep["sigma"].shape[1], _dp(keep["sigma"])
    s.nb, s.bcells = len(keep["bcells"]), _ip(keep["bcells"])
    s.bc_base, s.bc_scale = _dp(keep["bc_base"]), _dp(keep["bc_scale"])
    s.nrestore, s.restore_cells = len(keep["restore"]), _ip(keep["restore"])
    s.ncand, s.cand_cells = len(keep["cand"]), _ip(keep["cand"])
    s.integrator, s.precision = integrator, precision
    s.r, s.rank_cap = P["r"], P["rank_cap"]
    s.theta, s.theta_index, s.rank_tol = P["theta"], P["thetaIndex"], rank_tol
    s.nz, s.dz = int(P.get("nz", 1)), float(P.get("dz", 1.0))
    if s.nz > 1:
        keep["Ay"], keep["absAy"] = f64(P["Ay"]), f64(P["absAy"])
        s.Ay, s.absAy = _dp(keep["Ay"]), _dp(keep["absAy"])
    if initial is not None:
        for name in ("X0", "S0", "W0"):
            keep[name] = f64(initial[name])
            setattr(s, name, _dp(keep[name]))
    if P.get("psi0") is not None:
        keep["psi0"] = f64(P["psi0"])
        s.psi0 = _dp(keep["psi0"])
    s.flux_order = int(P.get("flux_order", 1))
    return s, keep


def make_fcs_problem(P):
    """Fill a ``FcsProblemV1`` from ``titus_amd.problem.build_fcs_problem``
    output.  Returns (struct, keepalive)."""
    f64 = lambda a: np.asfortranarray(np.asarray(a, dtype=np.float64))
    i32 = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    keep = {
        "density": f64(P["density"]), "degree": i32(P["degree"]),
        "Ax": f64(P["Ax"]), "Az": f64(P["Az"]), "absAx": f64(P["absAx"]), "absAz": f64(P["absAz"]),
        "omega": f64(P["omega"]), "M": f64(P["M"]), "dE": f64(P["dE_steps"]),
        "smid": f64(P["SMid"]), "sigma": np.ascontiguousarray(P["sigma"], dtype=np.float64),
        "bcells": i32(P["bcells"]), "bc_base": f64(P["bc_base"]),
        "bc_scale": f64(P["fcs_bc_scale"]),
    }
    s = FcsProblemV1()
    s.version = ABI_VERSION
    s.nx, s.ny, s.dx, s.dy = P["nx"], P["ny"], P["dx"], P["dy"]
    s.density = _dp(keep["density"])
    s.nm, s.degree = P["Nm"], _ip(keep["degree"])
    s.Ax, s.Az = _dp(keep["Ax"]), _dp(keep["Az"])
    s.absAx, s.absAz = _dp(keep["absAx"]), _dp(keep["absAz"])
    s.nq, s.omega, s.M = P["Q"], _dp(keep["omega"]), _dp(keep["M"])
    s.ne, s.dE, s.stopping_mid = P["nE"], _dp(keep["dE"]), _dp(keep["smid"])
    s.nsigma, s.sigma = keep["sigma"].shape[1], _dp(keep["sigma"])
    s.nb, s.bcells = len(keep["bcells"]), _ip(keep["bcells"])
    s.bc_base, s.bc_scale = _dp(keep["bc_base"]), _dp(keep["bc_scale"])
    if P.get("psi0") is not None:
        keep["psi0"] = f64(P["psi0"])
        s.psi0 = _dp(keep["psi0"])
    return s, keep


def solve_batch(contexts, nthreads=0):
    """Run every context to the end of its grid concurrently (titus_solve_batch).

    Returns the per-context status codes; raises on the first failure.
    """
    lib = load()
    n = len(contexts)
    arr = (ctypes.c_void_p * n)(*[c.h for c in contexts])
    status = np.zeros(n, dtype=np.int32)
    code = lib.titus_solve_batch(arr, n, nthreads, _ip(status))
    check(code)
    return status


def solve_batch_lockstep(contexts, width=0, groups=1):
    """Run every context to the end of its grid with shared launches
    (titus_solve_batch_lockstep): up to `width` solves (0: all) advance
    together in `groups` groups, and in a group each kernel runs once for all
    of its solves.

    Returns (per-context status codes, {"flushes", "launches", "ops", "flush_s",
    "stream_wait_s", "launches_lower_bound", "tiles"}), "tiles" holding each
    group's tuned tile choices; raises on the first failure.
    """
    lib = load()
    n = len(contexts)
    arr = (ctypes.c_void_p * n)(*[c.h for c in contexts])
    status = np.zeros(n, dtype=np.int32)
    stats = (ctypes.c_int64 * 6)()
    g = max(1, min(groups if groups > 0 else 1, width if width > 0 else n, max(n, 1)))
    tiles = np.zeros(4 * g, dtype=np.int32)
    code = lib.titus_solve_batch_lockstep(arr, n, width, groups, _ip(status), stats, _ip(tiles))
    check(code)
    return status, {"flushes": int(stats[0]), "launches": int(stats[1]), "ops": int(stats[2]),
                    "flush_s": stats[3] * 1e-9, "stream_wait_s": stats[4] * 1e-9,
                    "launches_lower_bound": int(stats[5]),
                    "tiles": [{"gemm_tiles_per_wave": int(t[0]), "kstep_tiles_per_wave": int(t[1]),
                               "sweep_rows": int(t[2]), "sweep_nontemporal": int(t[3])}
                              for t in tiles.reshape(g, 4)]}


def adjoint_batch(contexts, dose_bars, nthreads=0, lockstep=False, groups=4):
    """titus_adjoint over many contexts concurrently; dose_bars is C x n
    (column i seeds context i).  Returns the density gradients, C x n.
    ``lockstep``: shared launches (titus_adjoint_batch_lockstep, ``nthreads``
    > 0 limiting how many run at once), else one stream per context over
    ``nthreads`` host workers (titus_adjoint_batch); the same gradients bit
    for bit."""
    lib = load()
    n = len(contexts)
    bars = [np.ascontiguousarray(dose_bars[:, i], dtype=np.float64) for i in range(n)]
    grads = [np.zeros(c.C) for c in contexts]
    arr = (ctypes.c_void_p * n)(*[c.h for c in contexts])
    pb = (_P * n)(*[_dp(b) for b in bars])
    pg = (_P * n)(*[_dp(g) for g in grads])
    status = np.zeros(n, dtype=np.int32)
    if lockstep:
        check(lib.titus_adjoint_batch_lockstep(arr, n, nthreads, groups, pb, pg, _ip(status)))
    else:
        check(lib.titus_adjoint_batch(arr, n, nthreads, pb, pg, _ip(status)))
    return np.stack(grads, axis=1) if n else np.zeros((0, 0))


class Context:
    """Owner of one ``titus_ctx`` (a device-resident solver state)."""

    def __init__(self, P, integrator, precision=PRECISION_FP64, device=0, rank_tol=1e-7,
                 initial=None):
        lib = load()
        self._lib = lib
        self.C, self.nm, self.nq = P["C"], P["Nm"], P["Q"]
        self.ne = P["nE"]
        st, keep = make_problem(P, integrator, precision, rank_tol, initial)
        h = ctypes.c_void_p()
        check(lib.titus_create(ctypes.byref(h), ctypes.byref(st), device))
        del keep
        self.h = h
        cfl = self.cfl()
        if cfl > 1.0:
            warnings.warn(f"CFL = {cfl:.3g} > 1: the explicit streaming update may be unstable "
                          "(the reference prints this number, SolverCSD.jl:963)", RuntimeWarning,
                          stacklevel=2)

    def cfl(self):
        """dE / min(dx, dy[, dz]) * max(1 / density) of the problem."""
        v = ctypes.c_double()
        check(self._lib.titus_get_cfl(self.h, ctypes.byref(v)))
        return v.value

    def set_analytic_beam(self, P):
        """The analytic beam of the RK4 integrator from a problem dict
        (``beam``, ``xMid``, ``yMid``, ``eTrafo``)."""
        keep = [np.ascontiguousarray(P[k], dtype=np.float64) for k in ("xMid", "yMid", "eTrafo")]
        b = P["beam"]
        st = BeamV1(b["x0"], b["y0"], b["Omega1"], b["Omega3"], _dp(keep[0]), _dp(keep[1]),
                    _dp(keep[2]))
        check(self._lib.titus_set_analytic_beam(self.h, ctypes.byref(st)))

    def set_rejection(self, c_eta):
        """Turn the rank-adaptive step's rejection on (c_eta > 0) or off (0)."""
        check(self._lib.titus_set_rejection(self.h, float(c_eta)))

    def rejections(self):
        n = ctypes.c_int64()
        check(self._lib.titus_get_rejections(self.h, ctypes.byref(n)))
        return n.value

    def step(self, s):
        check(self._lib.titus_step(self.h, s))

    def solve(self, last_step=0):
        check(self._lib.titus_solve(self.h, last_step))

    def progress(self):
        a, b = ctypes.c_int32(), ctypes.c_int32()
        check(self._lib.titus_progress(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def state(self):
        _, r = self.progress()
        dose = np.zeros(self.C)
        X = np.zeros((self.C, r), order="F")
        S = np.zeros((r, r), order="F")
        W = np.zeros((self.nm, r), order="F")
        psi = np.zeros((self.C, self.nq), order="F")
        check(self._lib.titus_get_state(self.h, _dp(dose), _dp(X), _dp(S), _dp(W), _dp(psi)))
        return dict(dose=dose, X=X, S=S, W=W, psi=psi)

    def history(self, n):
        ranks = np.zeros(n, dtype=np.int32)
        etas = np.zeros(n)
        check(self._lib.titus_get_history(self.h, _ip(ranks), _dp(etas), n))
        return ranks, etas

    def set_timing(self, period):
        """Time every `period`-th step (True: every step; 0 / False: off)."""
        check(self._lib.titus_set_timing(self.h, int(period)))

    KERNELS = ("bc", "dose", "sweep", "stream_K", "galerkin", "gram", "gemm", "chol", "psi", "angular")
    HOST_PHASES = ("host:step", "host:wait", "host:range_finish", "host:truncate", "host:range_w",
                   "host:galerkin_w", "host:stream_LS")

    def kernel_stats(self, name):
        """(total ms, launches, algorithmic bytes, flops) since timing was enabled."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        by, fl = ctypes.c_double(), ctypes.c_double()
        check(self._lib.titus_kernel_stats(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n),
                                           ctypes.byref(by), ctypes.byref(fl)))
        return ms.value, n.value, by.value, fl.value

    def tuning(self, kernel="sweep"):
        """Launch plan of `kernel`: dict(rows, groups, tuned, nontemporal) for the sweep."""
        v = np.zeros(4, dtype=np.int32)
        check(self._lib.titus_get_tuning(self.h, kernel.encode(), _ip(v), 4))
        return {"rows": int(v[0]), "groups": int(v[1]), "tuned": bool(v[2]),
                "nontemporal": bool(v[3])}

    def decisions(self):
        """(hash, count) of the discrete decisions taken so far."""
        h, n = ctypes.c_uint64(), ctypes.c_int64()
        check(self._lib.titus_get_decisions(self.h, ctypes.byref(h), ctypes.byref(n)))
        return h.value, n.value

    def set_checkpoints(self, every):
        """Keep the state every `every` steps of the coming solve (for adjoint)."""
        check(self._lib.titus_set_checkpoints(self.h, int(every)))

    def adjoint(self, dose_bar):
        """d <dose_bar, dose> / d density (C values) for the steps solved."""
        bar = np.ascontiguousarray(dose_bar, dtype=np.float64)
        if bar.shape != (self.C,):
            raise ValueError("dose_bar must have one value per cell")
        grad = np.zeros(self.C)
        check(self._lib.titus_adjoint(self.h, _dp(bar), _dp(grad)))
        return grad

    def close(self):
        if getattr(self, "h", None):
            self._lib.titus_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FcsContext:
    """Owner of one ``titus_fcs`` (the full-rank P_N solver on the device)."""

    KERNELS = ("bc", "sweep", "collide", "dose")

    def __init__(self, P, device=0):
        lib = load()
        self._lib = lib
        self.C, self.nm, self.nq, self.ne = P["C"], P["Nm"], P["Q"], P["nE"]
        st, keep = make_fcs_problem(P)
        h = ctypes.c_void_p()
        check(lib.titus_fcs_create(ctypes.byref(h), ctypes.byref(st), device))
        del keep
        self.h = h

    def step(self, s):
        check(self._lib.titus_fcs_step(self.h, s))

    def solve(self, last_step=0):
        check(self._lib.titus_fcs_solve(self.h, last_step))

    def state(self):
        """dict(u C x nm (unscaled), dose C, psi C x nq)."""
        u = np.zeros((self.C, self.nm), order="F")
        dose = np.zeros(self.C)
        psi = np.zeros((self.C, self.nq), order="F")
        check(self._lib.titus_fcs_get_state(self.h, _dp(u), _dp(dose), _dp(psi)))
        return dict(u=u, dose=dose, psi=psi)

    def set_timing(self, on=True):
        check(self._lib.titus_fcs_set_timing(self.h, int(bool(on))))

    def kernel_stats(self, name):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        by, fl = ctypes.c_double(), ctypes.c_double()
        check(self._lib.titus_fcs_kernel_stats(self.h, name.encode(), ctypes.byref(ms),
                                               ctypes.byref(n), ctypes.byref(by), ctypes.byref(fl)))
        return ms.value, n.value, by.value, fl.value

    def plan(self):
        """Operator plan of the collided kernel."""
        v = np.zeros(7, dtype=np.int32)
        check(self._lib.titus_fcs_get_plan(self.h, _ip(v), 7))
        return {"column_blocks_per_wave": int(v[0]), "groups": int(v[1]),
                "nonzero_blocks": int(v[2]), "dense_blocks": int(v[3]), "tuned": bool(v[4]),
                "cell_tiles_per_wave": int(v[5]), "k_steps_per_block": int(v[6])}

    def close(self):
        if getattr(self, "h", None):
            self._lib.titus_fcs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
