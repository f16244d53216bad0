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


@pytest.mark.parametrize("variant", range(23))
def test_every_tile_variant(probe, variant):
    """Each MFMA tile variant the autotuner may pick computes the same GEMM,
    LU and triangular solves.  The 1300 x 1000 shape gives tile counts that
    are multiples of 8 with a partial group of m-tiles, the case in which the
    XCD-aware variants (5, 6, 9, complex64 11 and 14, complex128 12) remap
    tiles; variants 7 and 8 (and complex64 9-11, complex128 10-12) form
    complex products from three real ones (Gauss); complex128 9, 10 and 12
    stage k-tiles two ahead (K = 33, 40, 128 cover one, two and an odd number
    of k-tiles); complex64 13-16 run on the bf16 matrix cores with three-way
    bf16 splits of the FP32 operands (32-deep k-tiles: K = 33 and 40 have a
    partial second k-tile), within the FP32 tolerance; complex128 13 is the
    lean 32 x 32 tile with 8-deep k-tiles of the LU panel that runs beside
    the trailing update, and complex128 14-16 are 64 x 64 Gauss tiles with
    8-deep k-tiles at 3-4 waves per SIMD (prefetch 2 for 14); K = 33 and 40
    end in partial k-tiles; complex128 17-19 stream the operands from global
    memory straight into the MFMA registers (no LDS, no barrier) with 7, 8 or
    6 k-steps of 4 in flight (128 x 32, 64 x 64 and 32 x 128 (20) workgroup tiles of 32 x 32
    wave tiles), so K = 33 ends in a partial step and K = 40 in a partial
    ring; 1300 x 1000 x 40 has interior waves, which skip the masks;
    complex64 17-22 are the same register-streamed tiles on FP32 MFMA."""
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


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_register_leaf_gives_the_bits_of_the_lds_leaf(probe, dt):
    """The triangular-solve leaf with each lane's column in registers
    (k_trsm_leaf_reg, one and two waves per workgroup) does the operations of
    the one-wave LDS leaf in the same order, so the LU factors and both
    triangular solves agree bit for bit.  n = 80 gives leaves of 32, 32 and 16
    rows; 32 matrices x 4100 right-hand sides take the two-wave form, with a
    partial last workgroup."""
    L, ctx = probe
    c = C.c_void_p(ctx.value)
    rng = np.random.default_rng(5)
    getrf = L.qbt_getrf_c128 if dt == np.complex128 else L.qbt_getrf_c64
    trsm = L.qbt_trsm_c128 if dt == np.complex128 else L.qbt_trsm_c64
    n, b, nr = 80, 32, 4100
    T = _rand(rng, (b, n, n), dt) + 4 * np.eye(n, dtype=dt)[None]
    B0 = _rand(rng, (b, nr, n), dt)
    T2 = _rand(rng, (2, 300, 300), dt)
    out = []
    try:
        for mode in (1, 0):
            L.qbt_set_trsm_leaf_lds(mode)
            A = T.copy()
            ipiv = np.zeros((b, n), dtype=np.int32)
            assert getrf(c, n, C.c_void_p(_c(A)), C.c_void_p(_c(ipiv)), b) == 0
            X = B0.copy()
            assert trsm(c, n, C.c_void_p(_c(A)), C.c_void_p(_c(X)), nr, b) == 0
            A2 = T2.copy()
            ip2 = np.zeros((2, 300), dtype=np.int32)
            assert getrf(c, 300, C.c_void_p(_c(A2)), C.c_void_p(_c(ip2)), 2) == 0
            out.append((A, ipiv, X, A2, ip2))
    finally:
        L.qbt_set_trsm_leaf_lds(-1)
    for u, v in zip(out[0], out[1]):
        assert np.array_equal(u, v)
    # and the solve is a solve
    Lm = np.tril(out[1][0][0].T, -1) + np.eye(n)
    Um = np.triu(out[1][0][0].T)
    perm_rhs = B0[0].T.astype(np.complex128)
    ref = sla.solve_triangular(Um, sla.solve_triangular(Lm, perm_rhs, lower=True,
                                                        unit_diagonal=True))
    tol = 1e-12 if dt == np.complex128 else 1e-4
    assert np.max(np.abs(out[1][2][0].T - ref)) / np.max(np.abs(ref)) < tol


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
@pytest.mark.parametrize("reg_variant", [17, 18, 19, 20, 21, 22])
def test_register_tiles_give_the_bits_of_the_lds_tile(probe, reg_variant, dt):
    """The register-streamed tiles (17-22 of each type) multiply the same
    operands in the same k order with the same MFMA and Gauss form as the
    64 x 64 Gauss LDS tile (complex128 16, complex64 10), so C - A B agrees bit
    for bit (edge tiles in M, N and a partial k-step included)."""
    L, ctx = probe
    c = C.c_void_p(ctx.value)
    ty, lds = (1, 16) if dt == np.complex128 else (0, 10)
    f = L.qbt_gemm_sub_c128 if dt == np.complex128 else L.qbt_gemm_sub_c64
    rng = np.random.default_rng(11)
    M, N, K, b = 300, 129, 131, 2
    A = _rand(rng, (b, K, M), dt)
    B = _rand(rng, (b, N, K), dt)
    C0 = _rand(rng, (b, N, M), dt)
    out = []
    try:
        for v in (lds, reg_variant):
            L.qbt_force_variant(c, ty, v)
            Cm = C0.copy()
            assert f(c, M, N, K, C.c_void_p(_c(A)), C.c_void_p(_c(B)), C.c_void_p(_c(Cm)), b) == 0
            out.append(Cm)
    finally:
        L.qbt_force_variant(c, ty, -1)
    assert np.array_equal(out[0], out[1])


def test_autotune_report_and_results():
    import qbhip
    ctx = qbhip.Context(0)
    rep = ctx.autotune(512, 4, 64, "mixed")
    assert "c64" in rep and "c128" in rep
    ctx.close()


def _oz(L, ctx, T, X, V):
    """qb_oz_residual through the harness: T [nb][N][N], X [nb][N][r] (numpy
    row/column indexing), V [N][r]; returns R [nb][N][r]."""
    nb, N, _ = T.shape
    r = X.shape[2]
    Tc = np.ascontiguousarray(T.transpose(0, 2, 1))  # column-major per node
    Xc = np.ascontiguousarray(X.transpose(0, 2, 1))
    Vc = np.ascontiguousarray(V.T)
    Rc = np.empty_like(Xc)
    P = lambda a: C.c_void_p(_c(a))
    rc = L.qbt_oz_residual(C.c_void_p(ctx.value), N, r, nb, P(Tc), P(Xc), P(Vc), P(Rc))
    assert rc == 0
    return Rc.transpose(0, 2, 1)


@pytest.mark.parametrize("N,r,nb", [(256, 8, 2), (200, 13, 3)])
def test_ozaki_product_is_exact_on_integers(probe, N, r, nb):
    """Integer-valued T and X whose row and column maxima lie in [2^52, 2^53)
    (scale exponents 0): with V = 0, -R is the integer product T X rounded
    once to FP64 (Garner's mixed-radix digits are exact; Horner's rule adds
    at most 16 roundings).  Compared with Python integer arithmetic.  A wrong
    lane layout of the int8 MFMA, a wrong modulus inverse or a wrong
    residue sign would give errors of order the entries themselves."""
    L, ctx = probe
    rng = np.random.default_rng(N + r)
    lim = 2 ** 52

    def ints(nb, n, m):
        """nb random complex integer matrices n x m (entries below 2^52 in
        magnitude) with one real part per row of magnitude in [2^52, 2^53)."""
        z = (rng.integers(-lim, lim, (nb, n, m)).astype(np.float64)
             + 1j * rng.integers(-lim, lim, (nb, n, m)).astype(np.float64))
        for bb in range(nb):
            for i in range(n):
                j = int(rng.integers(0, m))
                sgn = 1.0 if rng.integers(0, 2) else -1.0
                z[bb, i, j] = sgn * float(lim + int(rng.integers(0, lim))) + 1j * z[bb, i, j].imag
        return z

    T = ints(nb, N, N)
    X = ints(nb, r, N).transpose(0, 2, 1)  # one large entry per column
    V = np.zeros((N, r), complex)
    R = _oz(L, ctx, T, X, V)
    for b in range(nb):
        Ti = np.vectorize(int, otypes=[object])
        Tr, Tim = Ti(T[b].real), Ti(T[b].imag)
        Xr, Xim = Ti(X[b].real), Ti(X[b].imag)
        Pre = Tr.dot(Xr) - Tim.dot(Xim)
        Pim = Tr.dot(Xim) + Tim.dot(Xr)
        got = -R[b]
        for P, g in ((Pre, got.real), (Pim, got.imag)):
            ex = np.array([[float(v) for v in row] for row in P])
            err = np.abs(g - ex)
            assert np.all(err <= 2.0 ** -48 * np.abs(ex) + 1.0), float(np.max(err / np.abs(ex)))


@pytest.mark.parametrize("N,r,nb", [(256, 8, 2), (333, 21, 2), (1024, 40, 1)])
def test_ozaki_residual_matches_extended_precision(probe, N, r, nb):
    """R = V - T X for complex T with rows of widely different scales and X
    with columns of different scales (so the per-row and per-column
    exponents differ), against the product in 80-bit extended precision:
    every entry within 4 * 2^-53 * N * max|T_i.| * max|X_.c| (the truncation
    of T and X to 53 significant bits per row / column) plus 2^-52 (|V| +
    |T X|) (the FP64 roundings of the reconstructed product and of
    V - T X), and no worse than 8x the error of the FP64 GEMM of NumPy on
    the same bound."""
    L, ctx = probe
    rng = np.random.default_rng(7 * N + r)
    T = (rng.standard_normal((nb, N, N)) + 1j * rng.standard_normal((nb, N, N)))
    T *= np.exp(rng.uniform(-6, 6, (nb, N, 1)))
    X = (rng.standard_normal((nb, N, r)) + 1j * rng.standard_normal((nb, N, r)))
    X *= np.exp(rng.uniform(-4, 4, (nb, 1, r)))
    V = rng.standard_normal((N, r)) + 1j * rng.standard_normal((N, r))
    R = _oz(L, ctx, T, X, V)
    for b in range(nb):
        Tl, Xl = T[b].astype(np.clongdouble), X[b].astype(np.clongdouble)
        ref = V.astype(np.clongdouble) - Tl @ Xl
        tm = np.max(np.maximum(np.abs(T[b].real), np.abs(T[b].imag)), axis=1)[:, None]
        xm = np.max(np.maximum(np.abs(X[b].real), np.abs(X[b].imag)), axis=0)[None, :]
        TX = np.abs((Tl @ Xl).astype(np.complex128))
        bound = 4 * 2.0 ** -53 * N * tm * xm + 2.0 ** -52 * (np.abs(V) + TX)
        err = np.abs((R[b] - ref).astype(np.complex128))
        print(N, r, "max err / bound", float(np.max(err / bound)))
        assert np.all(err <= bound)
        e64 = np.abs((V - T[b] @ X[b]) - ref.astype(np.complex128))
        assert np.max(err / bound) <= max(8 * float(np.max(e64 / bound)), 0.05)


@pytest.mark.parametrize("variant", [13, 14, 15, 16])
def test_bf16_split_gemm_is_fp32_accurate(probe, variant):
    """The complex64 GEMM on bf16 MFMA (three-way bf16 splits of every FP32
    operand, six products per real product, FP32 accumulation) has the
    error of the FP32-MFMA kernel on the same operands, within 2x; a
    two-way split (or a dropped middle product) would be ~2^8 times worse."""
    L, ctx = probe
    c = C.c_void_p(ctx.value)
    rng = np.random.default_rng(variant)
    M, N, K, b = 515, 260, 700, 2
    A = _rand(rng, (b, K, M), np.complex64)
    B = _rand(rng, (b, N, K), np.complex64)
    C0 = _rand(rng, (b, N, M), np.complex64)
    ref = np.stack([(C0[i].T.astype(np.complex128) - A[i].T.astype(np.complex128)
                     @ B[i].T.astype(np.complex128)).T for i in range(b)])
    errs = {}
    try:
        for v in (0, variant):
            L.qbt_force_variant(c, 0, v)
            Cm = C0.copy()
            assert L.qbt_gemm_sub_c64(c, M, N, K, C.c_void_p(_c(A)), C.c_void_p(_c(B)),
                                      C.c_void_p(_c(Cm)), b) == 0
            errs[v] = float(np.max(np.abs(Cm - ref)) / np.max(np.abs(ref)))
    finally:
        L.qbt_force_variant(c, 0, -1)
    print(errs)
    assert errs[variant] <= 2 * errs[0] + 1e-7


def test_copy_kernel_reference_for_traffic_counters(probe):
    """A plain copy kernel of known size (k_copy16: 1 GiB read + 1 GiB
    written per launch, two launches): the reference against which the
    FETCH_SIZE / WRITE_SIZE counter runs over the library's kernels are
    scaled (the counters run over this test and the bench in the same
    rocprofv3 pass).  Here only the bandwidth is sanity-checked."""
    L, ctx = probe
    L.qbt_copy_bench.restype = C.c_double
    gbs = L.qbt_copy_bench(C.c_void_p(ctx.value), C.c_int64(1 << 30), 2)
    print("copy GB/s", gbs)
    assert gbs > 100.0
