"""Pins the oracle (oracle/oracle.py) against closed forms, brute force and
independent library routines.  The reference's own tests hold no golden
vectors for this path (its Beyn/Kress code is untested upstream), so each
check below is chosen so that a dropped term, wrong sign or transposed index
in the oracle fails it."""
import math

import mpmath as mp
import numpy as np
import pytest
import scipy.special as sp

from oracle import oracle as O


def test_kress_R_matches_fft_construction_and_brute_force():
    # the reference builds column 1 by ifft of a = [0, 1, 1/2, ..., 1/(n-1), 0, 1/(n-1), ..., 1]
    for N in (8, 9, 16, 64, 255, 256):
        n = N // 2 if N % 2 == 0 else (N - 1) // 2
        a = np.zeros(N, complex)
        mmax = n - 1 if N % 2 == 0 else n
        for m in range(1, mmax + 1):
            a[m] = 1 / m
            a[N - m] = 1 / m
        rjn = np.real(np.fft.ifft(a))
        col = -2 * math.pi * rjn
        if N % 2 == 0:
            col -= (4 * math.pi / N ** 2) * (-1.0) ** np.arange(N)
        assert np.max(np.abs(O.kress_R_col(N) - col)) < 1e-14
        # brute force in mpmath for one entry
        d = 3
        s = mp.mpf(0)
        for m in range(1, mmax + 1):
            s += mp.cos(2 * mp.pi * m * d / N) / m
        v = -(4 * mp.pi / N) * s - ((4 * mp.pi / N ** 2) * (-1) ** d if N % 2 == 0 else 0)
        assert abs(O.kress_R_col(N)[d] - float(v)) < 1e-14
    R = O.kress_R(16)
    assert np.allclose(R, R.T) and np.allclose(R[3, 1], O.kress_R_col(16)[2])


def test_chebfit_is_exact_interpolation_and_clenshaw_matches_chebval():
    rng = np.random.default_rng(0)
    for M in (4, 5, 7, 9):
        c = rng.standard_normal(M + 1) + 1j * rng.standard_normal(M + 1)
        t = O.lobatto_t(M)
        f = np.polynomial.chebyshev.chebval(t, c)
        # _chebfit! (chebyshev_core.jl:67-80) halves c_0 but not c_M: its
        # coefficients are the Lobatto interpolant's with the top one doubled
        expect = c.copy()
        expect[M] *= 2
        assert np.max(np.abs(O.chebfit(f) - expect)) < 1e-14
        tt = rng.uniform(-1, 1, 50)
        ref = np.polynomial.chebyshev.chebval(tt, c)
        got = O.clenshaw(np.broadcast_to(c, (50, M + 1)), tt)
        assert np.max(np.abs(got - ref)) < 1e-13


def test_breaks_and_panel_lookup():
    br = O.breaks_uniform(0.1, 2.3, 7)
    assert br[0] == 0.1 and br[-1] == 2.3 and len(br) == 8
    P = O.Panels(0.1, 2.3, 7, 5)
    r = np.array([0.1, 0.1 + 1e-12, 0.5, 2.3, 2.3 - 1e-12])
    p, t = P.locate(r)
    assert list(p) == [0, 0, 1, 6, 6]
    assert np.all(np.abs(t) <= 1 + 1e-12)
    assert abs(t[0] + 1) < 1e-14 and abs(t[3] - 1) < 1e-14


def test_scipy_amos_against_mpmath():
    # the oracle's table samples are scipy.special.hankel1/jv (AMOS)
    for z in (0.7 + 0.01j, 5.3 - 0.2j, 31.0 + 0.4j, 812.25 - 0.5j):
        for nu in (0, 1):
            h = complex(mp.hankel1(nu, mp.mpc(z)))
            j = complex(mp.besselj(nu, mp.mpc(z)))
            assert abs(sp.hankel1(nu, z) - h) / abs(h) < 1e-14
            assert abs(sp.jv(nu, z) - j) / max(abs(j), 1e-3) < 1e-13


def test_cfie_table_route_matches_exact_hankel_twin_at_real_k():
    """At real k the Chebyshev assembly must agree with the direct twin
    (cfie_kress.jl:416-510) up to the table interpolation error."""
    comps = O.make_geometry("star", [160])
    k = 17.3
    T = O.cfie_assemble(comps, [k], O.ChebParams(20000, 9, 8000, 9))[0]
    Td = O.cfie_direct_real(comps, k)
    assert np.max(np.abs(T - Td)) / np.max(np.abs(Td)) < 1e-12


def _circulant_spectrum(T):
    # eigenvalues of a circulant matrix = FFT of its first column
    return np.fft.fft(T[:, 0])


def test_circle_cfie_is_circulant_with_bessel_multipliers():
    """On the unit circle T(z) is circulant.  With the reference's kernel
    scaling D = 2K, S-part = 2S (Phi = (i/4) H_0), the jump relation
    2K e^{im t} = (1 + i pi z J_m H_m') e^{im t} and 2S e^{im t} =
    (i pi / 2)... give the multiplier pi z J_m(z) (H_m(z) - i H_m'(z)):
    T(z) is singular exactly where J_m(z) = 0."""
    N = 128
    comps = O.make_geometry("circle", [N])
    z = 9.0 + 0.2j
    T = O.cfie_assemble(comps, [z], O.ChebParams(6000, 7, 3000, 7))[0]
    # circulant: T[i+1, j+1] == T[i, j]
    assert np.max(np.abs(np.roll(np.roll(T, 1, 0), 1, 1) - T)) < 1e-12
    lam = _circulant_spectrum(T)
    m = np.fft.fftfreq(N, 1.0 / N).astype(int)
    Jm, Hm = sp.jv(m, z), sp.hankel1(m, z)
    Hpm = sp.h1vp(m, z)
    mult = math.pi * z * Jm * (Hm - 1j * Hpm)
    sel = np.abs(m) <= 20
    assert np.max(np.abs(lam[sel] - mult[sel])) < 1e-10


def test_circle_dlp_multipliers():
    N = 128
    comps = O.make_geometry("circle", [N])
    z = 11.0 - 0.3j
    F = O.dlp_assemble(comps[0], [z], O.ChebParams(6000, 7, 3000, 7))[0]
    lam = np.fft.fft(F[:, 0])
    m = np.fft.fftfreq(N, 1.0 / N).astype(int)
    mult = -1j * math.pi * z * sp.jv(m, z) * sp.h1vp(m, z)
    sel = np.abs(m) <= 20
    assert np.max(np.abs(lam[sel] - mult[sel])) < 1e-10


@pytest.mark.slow
def test_oracle_beyn_circle_gives_bessel_zeros():
    """cfg1 end to end: the Beyn eigenvalues inside |k-20| <= 0.5 are the
    zeros of J_m (multiplicity 2 for m > 0)."""
    comps = O.make_geometry("circle", [256])
    rng = np.random.default_rng(1)
    V = (rng.standard_normal((256, 24)) + 1j * rng.standard_normal((256, 24))) * np.sqrt(0.5)
    lam, S, rk, _, _ = O.beyn_solve(lambda zs: O.cfie_assemble(comps, zs), 20.0, 0.5, 32, V, 1e-10)
    inside = np.sort(lam[np.abs(lam - 20) <= 0.5].real)
    ex = O.jm_zeros_in_disk(20.0, 0.5)
    assert len(inside) == len(ex) == 10
    assert np.max(np.abs(inside - ex) / ex) < 1e-12


def test_jm_zero_fixture():
    ex = O.jm_zeros_in_disk(20.0, 0.5)
    np.testing.assert_allclose(np.unique(ex), [19.554536430997054, 19.615858510468243,
                                               19.615966903966921, 19.994430629816385,
                                               20.320789213566506], rtol=1e-14)


@pytest.mark.parametrize("kind,Ns,k", [("circle_hole", [60, 20], 12.0 + 0.1j),
                                       ("star", [96], 15.0 - 0.2j)])
def test_cfie_derivative_formulas_match_cauchy_derivatives(kind, Ns, k):
    """The wavenumber derivatives of the CFIE matrix (the reference's analytic
    formulas, restated in cfie_assemble_deriv) equal the derivatives of the
    plain assembly taken by Cauchy's integral formula on a circle around k;
    the undifferentiated output equals the plain assembly."""
    comps = O.make_geometry(kind, Ns)
    cp = O.ChebParams(3000, 7, 1500, 7)
    A, A1, A2 = O.cfie_assemble_deriv(comps, [k], cp)
    P = O.cfie_assemble(comps, [k], cp)[0]
    assert np.max(np.abs(A[0] - P)) < 1e-14 * np.max(np.abs(P)) * 10
    d1, d2 = O.cauchy_derivatives(lambda zs: O.cfie_assemble(comps, zs, cp), k)
    e1 = np.max(np.abs(A1[0] - d1)) / np.max(np.abs(d1))
    e2 = np.max(np.abs(A2[0] - d2)) / np.max(np.abs(d2))
    print(kind, e1, e2)
    assert e1 < 1e-11 and e2 < 1e-10


def test_dlp_derivative_formulas_match_cauchy_derivatives():
    comp = O.make_geometry("ellipse", [64])[0]
    k = 9.0 - 0.2j
    cp = O.ChebParams(3000, 7, 1500, 7)
    F, F1, F2 = O.dlp_assemble_deriv(comp, [k], cp)
    P = O.dlp_assemble(comp, [k], cp)[0]
    assert np.max(np.abs(F[0] - P)) < 1e-13
    d1, d2 = O.cauchy_derivatives(lambda zs: O.dlp_assemble(comp, zs, cp), k)
    e1 = np.max(np.abs(F1[0] - d1)) / np.max(np.abs(d1))
    e2 = np.max(np.abs(F2[0] - d2)) / np.max(np.abs(d2))
    print(e1, e2)
    assert e1 < 1e-11 and e2 < 1e-10


def test_cfie_derivative_reference_sign_differs_only_between_components():
    """reference_offblock_sign flips exactly the two double-layer derivative
    terms of the between-component blocks: the same-component blocks are
    untouched and the difference there is 2 x (i/2) w_j inn k H0 for dT/dk."""
    comps = O.make_geometry("circle_hole", [60, 20])
    cp = O.ChebParams(3000, 7, 1500, 7)
    k = 12.0 + 0.1j
    A, A1, A2 = O.cfie_assemble_deriv(comps, [k], cp)
    B, B1, B2 = O.cfie_assemble_deriv(comps, [k], cp, reference_offblock_sign=True)
    assert np.array_equal(A, B)
    for X, Y in ((A1, B1), (A2, B2)):
        assert np.array_equal(X[0, :60, :60], Y[0, :60, :60])
        assert np.array_equal(X[0, 60:, 60:], Y[0, 60:, 60:])
    ca, cb = comps[1], comps[0]   # rows of the hole, columns of the outer boundary
    DX = ca.x[:, None] - cb.x[None, :]
    DY = ca.y[:, None] - cb.y[None, :]
    inn = cb.dy[None, :] * DX - cb.dx[None, :] * DY
    r = np.hypot(DX, DY)
    h0 = sp.hankel1(0, k * r)
    expect = -2 * (0.5j * cb.ws[None, :] * inn * k * h0)   # A1 - B1 = -(d1 - (-d1))
    got = A1[0, 60:, :60] - B1[0, 60:, :60]
    assert np.max(np.abs(got - expect)) < 1e-11 * np.max(np.abs(expect))
