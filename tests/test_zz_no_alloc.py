"""titus_galerkin_noalloc (the allocation-free entry point) against
titus_galerkin on the same inputs: two calls in a row, the second with a
smaller grid and rank, each equal bit for bit (the same kernels and the same
partial-sum layout), and both against the oracle's stencils.  The second call
reuses the device array the first one filled, so a stale partial sum would
show as a difference."""
import numpy as np
import pytest

from titus_amd import _lib
from test_gpu_primitives import device, galerkin, oracle

pytestmark = pytest.mark.gpu


def test_noalloc_galerkin_equals_the_allocating_one_twice():
    lib = _lib.load()
    h = (0.1, 0.07, 0.05)
    for shape, r in (((96, 80, 1), 24), ((20, 17, 1), 5), ((12, 10, 11), 30), ((6, 5, 7), 2)):
        nx, ny, nz = shape
        C = nx * ny * nz
        print(f"\n[case] noalloc galerkin {nx}x{ny}x{nz} r={r}", flush=True)
        rng = np.random.default_rng(C * 7 + r)
        rho = rng.uniform(0.3, 1.6, C)
        X = rng.standard_normal((C, r))
        irho = device(1.0 / rho)
        a = galerkin(lib.titus_galerkin, nx, ny, nz, h, irho, X, r)
        b = galerkin(lib.titus_galerkin_noalloc, nx, ny, nz, h, irho, X, r)
        ref, scale = oracle(nx, ny, nz, h, rho, X)
        err = np.max(np.abs(b - ref) / scale)
        print(f"[noalloc {nx}x{ny}x{nz} r={r}] bitwise equal {np.array_equal(a, b)}, "
              f"oracle error {err:.2e}", flush=True)
        assert np.array_equal(a, b)
        assert err < 1e-13


def test_noalloc_refuses_sizes_beyond_its_limits():
    import ctypes
    lib = _lib.load()
    irho = device(np.ones(64))
    g = _lib.GridV1(8, 8, 1, 0.1, 0.1, 1.0, irho.ptr.value)
    X = device(np.zeros(64 * 97))
    code = lib.titus_galerkin_noalloc(ctypes.byref(g), X.ptr, 64, 97, X.ptr, None)
    assert code == -1 and b"rank" in lib.titus_last_error()
