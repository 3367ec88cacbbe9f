"""Host dense LA of the library (csrc/smallla.cpp) against numpy and the oracle's
basis rules (host build, no GPU)."""
import ctypes
import os

import numpy as np
import pytest

from oracle import dlra

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "titus_amd",
                   "_smallla_host.so")
P = ctypes.POINTER(ctypes.c_double)
PI = ctypes.POINTER(ctypes.c_int)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("host LA library not built")
    L = ctypes.CDLL(LIB)
    L.smallla_svd.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P]
    L.smallla_svd_left.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P]
    L.smallla_range_basis.argtypes = [ctypes.c_int, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, P,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      PI, P, PI]
    L.smallla_svd_completed.argtypes = [ctypes.c_int, P, ctypes.c_double, P, P, P]
    return L


def f(a):
    return np.asfortranarray(a, dtype=np.float64)


def dp(a):
    return a.ctypes.data_as(P)


def csvd(lib, A):
    m, n = A.shape
    k = min(m, n)
    A = f(A)
    U, s, V = f(np.zeros((m, k))), np.zeros(k), f(np.zeros((n, k)))
    assert lib.smallla_svd(m, n, dp(A), dp(U), dp(s), dp(V)) == 0
    return U, s, V


@pytest.mark.parametrize("shape", [(6, 6), (40, 7), (253, 18), (5, 9)])
def test_svd_matches_numpy(lib, shape):
    rng = np.random.default_rng(sum(shape))
    A = rng.standard_normal(shape) * np.logspace(0, -8, shape[1])[None, :]
    U, s, V = csvd(lib, A)
    s_ref = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(s, s_ref, rtol=1e-12, atol=1e-15 * s_ref[0])
    np.testing.assert_allclose(U @ np.diag(s) @ V.T, A, atol=1e-13 * s_ref[0])
    np.testing.assert_allclose(V.T @ V, np.eye(V.shape[1]), atol=1e-13)


def exactly_deficient(rng, m, n, rank):
    """Exact zeros: zero columns, a repeated column and a zero block."""
    A = rng.standard_normal((m, rank)) @ rng.standard_normal((rank, n))
    A[:, 1] = 0.0
    A[:, -1] = A[:, 0]
    A[:, 2:4] = 0.0
    return A


@pytest.mark.parametrize("shape", [(40, 7), (136, 30), (12, 12), (5, 9)])
def test_svd_left_vectors_do_not_depend_on_want_v(lib, shape):
    """The adjoint replays a segment with right vectors requested; its bases
    must equal the forward's (computed without them) bit for bit."""
    rng = np.random.default_rng(shape[0] + shape[1])
    m, n = shape
    k = min(m, n)
    for A in (rng.standard_normal(shape) * np.logspace(0, -14, n)[None, :],
              exactly_deficient(rng, m, n, min(3, k)), np.zeros(shape)):
        U, s, V = csvd(lib, A)
        U2, s2 = f(np.zeros((m, k))), np.zeros(k)
        assert lib.smallla_svd_left(m, n, dp(f(A)), dp(U2), dp(s2)) == 0
        assert np.array_equal(U, U2) and np.array_equal(s, s2)
        assert np.all(np.isfinite(V))
        np.testing.assert_allclose(V.T @ V, np.eye(k), atol=1e-12)
        np.testing.assert_allclose(U @ np.diag(s) @ V.T, A, atol=1e-13 * max(s[0], 1.0))


def crange(lib, A, ncols, against=None, scale=-1.0, greedy=False, cand=None, tol=1e-7):
    m, k = A.shape
    q = 0 if against is None else against.shape[1]
    out = f(np.zeros((m, ncols)))
    rank = ctypes.c_int()
    Ag = f(against) if q else f(np.zeros((m, 1)))
    c = np.ascontiguousarray(cand if cand is not None else [], dtype=np.int32)
    assert lib.smallla_range_basis(m, k, dp(f(A)), ncols, q, dp(Ag), scale, tol, int(greedy),
                                   len(c), c.ctypes.data_as(PI) if cand is not None else None,
                                   dp(out), ctypes.byref(rank)) == 0
    return out, rank.value


@pytest.mark.parametrize("greedy", [False, True])
def test_range_basis_same_rule_as_oracle(lib, greedy):
    rng = np.random.default_rng(11)
    m, k, ncols = 60, 6, 6
    A = rng.standard_normal((m, 3)) @ rng.standard_normal((3, k))      # rank 3
    A[:5] = 0.0
    cand = list(range(5, m))
    B, p = crange(lib, A, ncols, greedy=greedy, cand=cand)
    Bo = dlra.range_basis(A, ncols, cand, greedy=greedy)
    assert p == 3
    np.testing.assert_allclose(B.T @ B, np.eye(ncols), atol=1e-13)
    # same range part (up to rotation) and the same completion columns
    np.testing.assert_allclose(B[:, :3] @ B[:, :3].T, Bo[:, :3] @ Bo[:, :3].T, atol=1e-12)
    np.testing.assert_allclose(B[:, 3:], Bo[:, 3:], atol=1e-12)


def test_range_basis_against_block(lib):
    rng = np.random.default_rng(12)
    m, r = 40, 4
    X, _ = np.linalg.qr(rng.standard_normal((m, r)))
    K = X @ rng.standard_normal((r, r)) + rng.standard_normal((m, 2)) @ rng.standard_normal((2, r))
    scale = np.linalg.norm(K)
    B, p = crange(lib, K, r, against=X, scale=scale, greedy=True)
    Bo = dlra.range_basis(K, r, range(m), against=X, scale=scale, greedy=True)
    assert p == 2
    np.testing.assert_allclose(X.T @ B, 0, atol=1e-13)
    np.testing.assert_allclose(B @ B.T, Bo @ Bo.T, atol=1e-12)
    np.testing.assert_allclose(B[:, 2:], Bo[:, 2:], atol=1e-12)


def test_zero_block_gives_identity_columns(lib):
    B, p = crange(lib, np.zeros((10, 4)), 4, greedy=True)
    assert p == 0
    np.testing.assert_allclose(B, np.eye(10)[:, :4])


def test_svd_completed_matches_oracle(lib):
    rng = np.random.default_rng(13)
    n = 8
    S = rng.standard_normal((n, 3)) @ rng.standard_normal((3, n))
    U, d, V = f(np.zeros((n, n))), np.zeros(n), f(np.zeros((n, n)))
    assert lib.smallla_svd_completed(n, dp(f(S)), 1e-7, dp(U), dp(d), dp(V)) == 0
    Uo, do, Vo = dlra.svd_completed(S)
    np.testing.assert_allclose(d, do, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(U @ np.diag(d) @ V.T, S, atol=1e-12)
    np.testing.assert_allclose(U[:, 3:] @ U[:, 3:].T, Uo[:, 3:] @ Uo[:, 3:].T, atol=1e-12)
    np.testing.assert_allclose(U.T @ U, np.eye(n), atol=1e-13)


def _csvd_completed(lib, S, tol=1e-7):
    n = S.shape[0]
    U, d, V = f(np.zeros((n, n))), np.zeros(n), f(np.zeros((n, n)))
    assert lib.smallla_svd_completed(n, dp(f(S)), tol, dp(U), dp(d), dp(V)) == 0
    return U, d, V


def _orthogonal(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


@pytest.mark.parametrize("kind", ["random", "graded", "rank_deficient", "zero_rows_cols", "repeated",
                                  "zero", "tiny_sizes"])
def test_truncation_svd_against_numpy(lib, kind):
    """The truncation's SVD (bidiagonalization and QR sweeps) on the shapes a
    rank-adaptive solve hands it: kept singular values to 1e-13 of sigma_1,
    U D V^T equal to numpy's truncation of S at the same kept count,
    orthonormal U and V, descending values, zeros beyond the kept count, and
    the null directions completed as the oracle completes them."""
    rng = np.random.default_rng(["random", "graded", "rank_deficient", "zero_rows_cols", "repeated",
                                 "zero", "tiny_sizes"].index(kind) + 40)
    sizes = [1, 2, 3] if kind == "tiny_sizes" else [4, 9, 18, 31, 40]
    for n in sizes:
        if kind in ("random", "tiny_sizes"):
            S = rng.standard_normal((n, n))
        elif kind == "graded":
            # singular values 1 .. 1e-11 over n, never within 1 % of the 1e-7 cut
            sv = 10.0 ** -np.linspace(0.0, 11.0, n)
            sv = np.where(np.abs(np.log10(sv) + 7.0) < 0.005, sv * 1.5, sv)
            S = _orthogonal(rng, n) @ np.diag(sv) @ _orthogonal(rng, n).T
        elif kind == "rank_deficient":
            k = max(1, n // 3)
            S = rng.standard_normal((n, k)) @ rng.standard_normal((k, n))
        elif kind == "zero_rows_cols":
            S = rng.standard_normal((n, n))
            S[:, : n // 3] = 0.0
            S[n // 2 + 1:, :] = 0.0
        elif kind == "repeated":
            sv = np.repeat([3.0, 1.0, 0.25], -(-n // 3))[:n]
            S = _orthogonal(rng, n) @ np.diag(sv) @ _orthogonal(rng, n).T
        else:
            S = np.zeros((n, n))
        U, d, V = _csvd_completed(lib, S)
        Uo, so, Vto = np.linalg.svd(S)
        p = int(np.count_nonzero(so > 1e-7 * so[0])) if so[0] > 0 else 0
        assert int(np.count_nonzero(d)) == p, (kind, n)
        assert np.all(np.diff(d) <= 0.0), (kind, n)
        np.testing.assert_allclose(U.T @ U, np.eye(n), atol=1e-13)
        np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-13)
        if p:
            np.testing.assert_allclose(d[:p], so[:p], rtol=0, atol=1e-13 * so[0])
            St = (Uo[:, :p] * so[:p]) @ Vto[:p]
            np.testing.assert_allclose(U @ np.diag(d) @ V.T, St, atol=1e-13 * so[0])
        Uc, dc, Vc = dlra.svd_completed(S)
        # the completion columns span what the oracle's span, to what the
        # SVD determines them (sin of the subspace angle <= eps ||S|| / gap
        # at the cut, Wedin's bound); for a zero block they are the oracle's
        # identity columns
        gap = (so[p - 1] - (so[p] if p < n else 0.0)) if p else 1.0
        tol = max(1e-11, 50.0 * np.finfo(float).eps * so[0] / gap) if p else 1e-11
        np.testing.assert_allclose(U[:, p:] @ U[:, p:].T, Uc[:, p:] @ Uc[:, p:].T, atol=tol)
        np.testing.assert_allclose(V[:, p:] @ V[:, p:].T, Vc[:, p:] @ Vc[:, p:].T, atol=tol)
        if kind == "zero":
            np.testing.assert_array_equal(U, np.eye(n))
            np.testing.assert_array_equal(V, np.eye(n))
