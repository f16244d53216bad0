"""GPU unit tests of the batched complex GEMM, LU and triangular solves of
libqbhip against NumPy/LAPACK (through tests/native/libqbt_linalg.so)."""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.linalg as sla

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def probe():
    import qbhip
    L = C.CDLL(os.path.join(HERE, "native", "libqbt_linalg.so"))
    ctx = qbhip.default_context()
    return L, ctx.ptr


def _c(a):
    return a.ctypes.data


def _rand(rng, shape, dt):
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return a.astype(dt)


@pytest.mark.parametrize("dt,tol", [(np.complex128, 1e-13), (np.complex64, 3e-5)])
@pytest.mark.parametrize("M,N,K,b", [(16, 16, 4, 1), (64, 64, 16, 1), (100, 70, 33, 2),
                                      (300, 129, 128, 3)])
def test_gemm_sub(probe, dt, tol, M, N, K, b):
    L, ctx = probe
    rng = np.random.default_rng(M + N + K)
    # column-major batches: store as [b][col][row] C-order arrays
    A = _rand(rng, (b, K, M), dt)
    B = _rand(rng, (b, N, K), dt)
    Cm = _rand(rng, (b, N, M), dt)
    ref = np.stack([(Cm[i].T - A[i].T @ B[i].T).T for i in range(b)])
    f = L.qbt_gemm_sub_c128 if dt == np.complex128 else L.qbt_gemm_sub_c64
    rc = f(C.c_void_p(ctx.value), M, N, K, C.c_void_p(_c(A)), C.c_void_p(_c(B)),
           C.c_void_p(_c(Cm)), b)
    assert rc == 0
    err = np.max(np.abs(Cm - ref)) / np.max(np.abs(ref))
    print(dt.__name__, M, N, K, b, err)
    assert err < tol


@pytest.mark.parametrize("dt,tol", [(np.complex128, 1e-12), (np.complex64, 1e-4)])
# 4160: panel width 512 (n >= 4096) with a last panel of 64 columns
@pytest.mark.parametrize("n,b", [(8, 1), (40, 2), (160, 2), (300, 1), (2200, 2), (4160, 1)])
def test_getrf_and_solve(probe, dt, tol, n, b):
    L, ctx = probe
    rng = np.random.default_rng(n)
    T = _rand(rng, (b, n, n), dt)          # [b][col][row]
    A = T.copy()
    ipiv = np.zeros((b, n), dtype=np.int32)
    f = L.qbt_getrf_c128 if dt == np.complex128 else L.qbt_getrf_c64
    assert f(C.c_void_p(ctx.value), n, C.c_void_p(_c(A)), C.c_void_p(_c(ipiv)), b) == 0
    for i in range(b):
        Ti = T[i].T.astype(np.complex128)
        LU = A[i].T.astype(np.complex128)
        Lm = np.tril(LU, -1) + np.eye(n)
        Um = np.triu(LU)
        perm = np.arange(n)
        for s in range(n):
            p = ipiv[i, s]
            assert s <= p < n
            perm[s], perm[p] = perm[p], perm[s]
        err = np.max(np.abs(Ti[perm] - Lm @ Um)) / np.max(np.abs(Ti))
        print("getrf", dt.__name__, n, i, err)
        assert err < tol
        # partial pivoting on |re| + |im| (izamax): below the diagonal, the
        # unscaled entries l_ik u_kk never exceed the pivot u_kk in that measure
        a1 = lambda x: np.abs(x.real) + np.abs(x.imag)
        lm = np.tril(LU, -1) * np.diag(LU)[None, :]
        assert np.all(a1(lm) <= a1(np.diag(LU))[None, :] * (1 + 1e-5))
    # triangular solves with the device factors
    nr = 7
    Bm = _rand(rng, (b, nr, n), dt)
    X = Bm.copy()
    g = L.qbt_trsm_c128 if dt == np.complex128 else L.qbt_trsm_c64
    assert g(C.c_void_p(ctx.value), n, C.c_void_p(_c(A)), C.c_void_p(_c(X)), nr, b) == 0
    for i in range(b):
        LU = A[i].T.astype(np.complex128)
        ref = sla.solve_triangular(np.triu(LU), sla.solve_triangular(
            np.tril(LU, -1) + np.eye(n), Bm[i].T.astype(np.complex128), lower=True,
            unit_diagonal=True))
        err = np.max(np.abs(X[i].T - ref)) / np.max(np.abs(ref))
        print("trsm", dt.__name__, n, i, err)
        assert err < tol * 10


@pytest.mark.parametrize("variant", range(13))
def test_every_tile_variant(probe, variant):
    """Each MFMA tile variant the autotuner may pick computes the same GEMM,
    LU and triangular solves.  The 1300 x 1000 shape gives tile counts that
    are multiples of 8 with a partial group of m-tiles, the case in which the
    XCD-aware variants (5, 6, 9, complex64 11, complex128 12) remap tiles;
    variants 7 and 8 (and complex64 9-11, complex128 10-12) form complex
    products from three real ones (Gauss); complex128 9, 10 and 12 stage
    k-tiles two ahead (K = 33, 40, 128 cover one, two and an odd number of
    k-tiles)."""
    L, ctx = probe
    c = C.c_void_p(ctx.value)
    try:
        for ty in (0, 1):
            L.qbt_force_variant(c, ty, variant)
        for dt, tol in ((np.complex128, 1e-13), (np.complex64, 3e-5)):
            for (M, N, K, b) in [(100, 70, 33, 2), (300, 129, 128, 3), (64, 300, 40, 1),
                                 (1300, 1000, 40, 2)]:
                test_gemm_sub(probe, dt, tol, M, N, K, b)
        test_getrf_and_solve(probe, np.complex128, 1e-12, 300, 1)
        test_getrf_and_solve(probe, np.complex64, 1e-4, 160, 2)
    finally:
        for ty in (0, 1):
            L.qbt_force_variant(c, ty, -1)


def test_autotune_report_and_results():
    import qbhip
    ctx = qbhip.Context(0)
    rep = ctx.autotune(512, 4, 64, "mixed")
    assert "c64" in rep and "c128" in rep
    ctx.close()
