"""titus_galerkin (device-pointer Galerkin products X^T L_o X) against the
oracle's stencils, 2D and 3D, FP64.

The oracle's stencils are pinned by dense assembly with the reference's loops
(tests/test_oracle.py, tests/test_oracle3d.py).  Bar: 1e-13 relative to
sum |terms| (the MFMA reduction sums in a different order than numpy).
"""
import ctypes
import sys

import numpy as np
import pytest

from oracle import dlra
from titus_amd import _lib

pytestmark = pytest.mark.gpu


class DeviceArray:
    """A device copy of a float64 array, allocated through the HIP runtime the
    library itself links (libamdhip64 of /opt/rocm; torch ships a runtime of
    its own, whose pointers this one does not own)."""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            _lib.load()                                     # libamdhip64 is loaded with it
            h = ctypes.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipSetDevice.argtypes = [ctypes.c_int]
            assert h.hipSetDevice(0) == 0
            cls._hip = h
        return cls._hip

    def __init__(self, a=None, n=None):
        h = self.hip()
        a = None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel(order="K")
        self.n = a.size if a is not None else n
        self.ptr = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.ptr), max(self.n, 1) * 8) == 0
        if a is not None:
            assert h.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0      # host to device

    def get(self):
        out = np.empty(self.n)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0   # device to host
        return out

    def __del__(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None


def device(a):
    return DeviceArray(a)


def galerkin(fn, nx, ny, nz, h, irho_d, X, r):
    """Call `fn` (titus_galerkin or a variant) on device copies; returns the
    2 nd blocks as an array (nops, r, r) of column-major r x r blocks."""
    C = X.shape[0]
    Xd = DeviceArray(np.asfortranarray(X).ravel(order="F"))    # column-major C x r
    nops = 4 if nz <= 1 else 6
    out = DeviceArray(n=nops * r * r)
    g = _lib.GridV1(nx, ny, nz, h[0], h[1], h[2], irho_d.ptr.value)
    _lib.check(fn(ctypes.byref(g), Xd.ptr, C, r, out.ptr, None))
    res = out.get()                                  # hipMemcpy waits for the default stream
    return res.reshape(nops, r, r).transpose(0, 2, 1)


def oracle(nx, ny, nz, h, rho, X):
    if nz <= 1:
        Ls = dlra.stencils(X, rho, nx, ny, h[0], h[1])
    else:
        Ls = dlra.stencils3(X, rho, nx, ny, nz, h[0], h[1], h[2])
    ref = np.stack([X.T @ L for L in Ls])
    scale = np.stack([np.abs(X).T @ np.abs(L) for L in Ls])
    return ref, scale


@pytest.mark.parametrize("shape,r", [((40, 37, 1), 9), ((200, 200, 1), 18), ((12, 10, 11), 30),
                                     ((64, 64, 64), 30), ((5, 4, 1), 1),
                                     # 2D grids of at least 2^17 cells: strips walked row by
                                     # row (vector loads; ragged strips with per-cell loads;
                                     # partial tiles; 7 interior rows, one band of four per
                                     # workgroup); more strips than workgroup slots: cell runs
                                     ((512, 512, 1), 16), ((400, 330, 1), 18), ((366, 365, 1), 9),
                                     ((9, 16000, 1), 5), ((3, 43700, 1), 5)])
def test_galerkin_products_match_the_oracle(shape, r):
    nx, ny, nz = shape
    print(f"\n[case] titus_galerkin {nx}x{ny}x{nz} r={r}", flush=True)
    C = nx * ny * nz
    rng = np.random.default_rng(C + r)
    rho = rng.uniform(0.3, 1.6, C)
    h = (0.1, 0.07, 0.05)
    X = rng.standard_normal((C, r))
    got = galerkin(_lib.load().titus_galerkin, nx, ny, nz, h, device(1.0 / rho), X, r)
    ref, scale = oracle(nx, ny, nz, h, rho, X)
    err = np.max(np.abs(got - ref) / scale)
    print(f"[galerkin {nx}x{ny}x{nz} r={r}] max error / sum|terms| {err:.2e}", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("shape,r", [((200, 200, 1), 18), ((64, 64, 64), 12), ((9, 8, 1), 5),
                                     ((6, 7, 8), 7), ((512, 512, 1), 20)])
def test_vector_and_per_cell_stencil_paths_agree_bitwise(shape, r):
    """The stencil products read B/rho with vector loads when the leading
    dimension and the unit-stride axis allow it, and cell by cell otherwise;
    both form the same values in the same order, so a leading dimension of
    C + 1 (per-cell loads) must give the same bits as C (vector loads)."""
    nx, ny, nz = shape
    print(f"\n[case] galerkin vector vs per-cell loads {nx}x{ny}x{nz} r={r}", flush=True)
    C = nx * ny * nz
    rng = np.random.default_rng(7 * C + r)
    rho = rng.uniform(0.3, 1.6, C)
    h = (0.1, 0.07, 0.05)
    X = rng.standard_normal((C, r))
    irho = device(1.0 / rho)
    nops = 4 if nz <= 1 else 6
    res = []
    for ld in (C, C + 1 if (C + 1) % 4 else C + 2):
        Xp = np.zeros((ld, r))
        Xp[:C] = X
        Xd = DeviceArray(np.asfortranarray(Xp).ravel(order="F"))
        out = DeviceArray(n=nops * r * r)
        g = _lib.GridV1(nx, ny, nz, h[0], h[1], h[2], irho.ptr.value)
        _lib.check(_lib.load().titus_galerkin(ctypes.byref(g), Xd.ptr, ld, r, out.ptr, None))
        res.append(out.get())
    assert np.array_equal(res[0], res[1])
