"""Pins of the oracle's kernel away from the unit circle.

The unit-circle checks (tests/test_oracle.py) cannot see mistakes that vanish
when the speed |gamma'|, the curvature and the quadrature weights are
constant.  The checks here do:

  * radius 0.7: speed 0.7 and curvature 1/0.7 enter the multipliers;
  * the concentric annulus 0.3 < |x| < 1: two components with different
    weights (2 pi / N_c), speeds and orientations, and the off-component
    blocks, against the zeros of J_m(k) Y_m(0.3 k) - J_m(0.3 k) Y_m(k);
  * Green's representation on the star, the ellipse and the disk with an
    off-centre hole: with T phi = -2 u_inc on the boundary for a field u_inc
    radiated from outside the domain, the combined potential of phi
    reproduces u_inc inside with spectral accuracy.  Every term of T (Kress
    split, log term, diagonal limits, off-component blocks) enters phi.

Each test names the slips it catches; the Green tests also run the same
check on deliberately slipped operators (built by _cfie_exact below, a
compact exact-Hankel builder that reproduces the oracle's matrix without a
slip) and require them to fail by orders of magnitude."""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.special as sp

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return json.load(open(os.path.join(GOLD, name)))


def _c(v):
    return complex(v[0], v[1])


# --------------------------------------------------------------------------
# circle of radius 0.7
# --------------------------------------------------------------------------


def test_circle_radius_07_multipliers_match_fixture():
    """CFIE and DLP on the circle of radius a = 0.7 (speed a, curvature 1/a):
    circulant with the multipliers pi a z J_m(a z) (H_m(a z) - i H_m'(a z))
    and -i pi a z J_m(a z) H_m'(a z) (mpmath fixture).  Catches a wrong
    power of the speed anywhere (in the S-term weights s_j, in the diagonal
    log(k^2 s^2 / 4), in the curvature term kappa = x' x y'' / |x'|^2 / 2pi,
    all of which the unit circle sets to 1) and a radius-dependent slip in the
    Kress split."""
    g = _load("circle_r07_multipliers.json")
    N = 128
    comps = O.make_geometry("circle", [N], R=0.7)
    cp = O.ChebParams(6000, 7, 3000, 7)
    m_all = np.fft.fftfreq(N, 1.0 / N).astype(int)
    for key, ent in g.items():
        assert ent["radius"] == 0.7
        z = _c(ent["z"])
        T = O.cfie_assemble(comps, [z], cp)[0]
        F = O.dlp_assemble(comps[0], [z], cp)[0]
        assert np.max(np.abs(np.roll(np.roll(T, 1, 0), 1, 1) - T)) < 1e-12
        lt, lf = np.fft.fft(T[:, 0]), np.fft.fft(F[:, 0])
        for e in ent["multipliers"]:
            i = int(np.flatnonzero(m_all == e["m"])[0])
            assert abs(lt[i] - _c(e["cfie"])) < 1e-10, (key, e["m"])
            assert abs(lf[i] - _c(e["dlp"])) < 1e-10, (key, e["m"])


def test_circle_radius_07_speed_slip_is_visible():
    """The same multipliers with the speed exponent slipped (a -> 1 in the
    closed form, i.e. the unit-circle answer) differ by far more than the
    tolerance, so the fixture does see the speed."""
    g = _load("circle_r07_multipliers.json")
    ent = next(iter(g.values()))
    z = _c(ent["z"])
    m = np.array([e["m"] for e in ent["multipliers"]])
    got = np.array([_c(e["cfie"]) for e in ent["multipliers"]])
    unit = math.pi * z * sp.jv(m, z) * (sp.hankel1(m, z) - 1j * sp.h1vp(m, z))
    assert np.max(np.abs(got - unit)) > 1e-2


# --------------------------------------------------------------------------
# concentric annulus
# --------------------------------------------------------------------------


def test_annulus_fixture_zeros_are_roots():
    """The fixture's zeros are roots of J_m(k) Y_m(0.3k) - J_m(0.3k) Y_m(k)
    (SciPy, independent of mpmath), and there is no other root of m < 60 in
    the disk."""
    g = _load("annulus_zeros.json")
    f = lambda m, k: sp.jv(m, k) * sp.yv(m, 0.3 * k) - sp.jv(m, 0.3 * k) * sp.yv(m, k)
    for e in g["zeros"]:
        k = e["k"]
        d = 1e-9 * k
        assert np.sign(f(e["m"], k - d)) != np.sign(f(e["m"], k + d)), e
    ks = np.linspace(13.45, 14.55, 20001)
    count = 0
    for m in range(60):
        v = f(m, ks)
        s = np.flatnonzero(np.sign(v[:-1]) != np.sign(v[1:]))
        count += sum(1 for i in s if 13.5 <= ks[i] <= 14.5)
    assert count == len(g["zeros"])


@pytest.mark.slow
def test_oracle_beyn_on_concentric_annulus_gives_cross_product_zeros():
    """Beyn with the oracle's CFIE on the annulus 0.3 < |x| < 1 (outer circle
    256 nodes, reversed hole 128 nodes; 32 contour nodes on |k - 14| = 0.5):
    the ten eigenvalues in the disk are the zeros of
    J_m(k) Y_m(0.3k) - J_m(0.3k) Y_m(k) with multiplicity 2 (m = 1, 2, 3, 6,
    10), to 1e-11.  The hole has speed 0.3 against 1 and weights 2 pi/128
    against 2 pi/256, so this catches: s_i <-> s_j and w_i <-> w_j in the
    off-component blocks (:632-648), a dropped or mis-signed off-component
    double-layer term, the hole's orientation reversal (cfie.jl:461-475),
    and the speed in the hole's diagonal limits.  Low-m modes feel the hole
    (j_{1,5} = 16.47 of the full disk is absent)."""
    g = _load("annulus_zeros.json")
    comps = O.make_geometry("circle_hole", [256, 128], hole=(0.3, 0.0, 0.0))
    rng = np.random.default_rng(2)
    V = (rng.standard_normal((384, 24)) + 1j * rng.standard_normal((384, 24))) * np.sqrt(0.5)
    cp = O.ChebParams(6000, 7, 3000, 7)
    lam, S, rk, _, _ = O.beyn_solve(lambda zs: O.cfie_assemble(comps, zs, cp), 14.0, 0.5, 32, V,
                                    1e-10)
    inside = np.sort(lam[np.abs(lam - 14.0) <= 0.5].real)
    ex = np.sort([e["k"] for e in g["zeros"] for _ in range(e["mult"])])
    print(inside, ex)
    assert len(inside) == len(ex) == 10
    assert np.max(np.abs(inside - ex) / ex) < 1e-11


# --------------------------------------------------------------------------
# Green's representation
# --------------------------------------------------------------------------


def _phi(k, x, y, src):
    """Phi(x, src) = (i/4) H0(k |x - src|)."""
    return 0.25j * sp.hankel1(0, k * np.hypot(x - src[0], y - src[1]))


def _potential(comps, k, phi, px, py):
    """u(x) = sum_j w_j [dPhi(x, y_j)/dnu_j + i k Phi(x, y_j)] |gamma'_j| phi_j
    with nu |gamma'| = (y', -x') (the combined field the CFIE matrix is minus
    twice the interior trace of); plain trapezoid rule, accurate for points
    away from the boundary."""
    X = np.concatenate([c.x for c in comps])
    Y = np.concatenate([c.y for c in comps])
    DX = np.concatenate([c.dx for c in comps])
    DY = np.concatenate([c.dy for c in comps])
    W = np.concatenate([c.ws for c in comps])
    S = np.hypot(DX, DY)
    u = np.zeros(len(px), complex)
    for q, (x, y) in enumerate(zip(px, py)):
        dx, dy = x - X, y - Y
        r = np.hypot(dx, dy)
        inn = DY * dx - DX * dy
        kern = (0.25j * k) * sp.hankel1(1, k * r) * inn / r + 1j * k * (0.25j * sp.hankel1(0, k * r)) * S
        u[q] = np.sum(W * kern * phi)
    return u


def _dl_potential(comp, k, phi, px, py):
    """u(x) = sum_j w_j dPhi(x, y_j)/dnu_j |gamma'_j| phi_j (the double layer
    the DLP matrix F = I - D is minus twice the interior trace of)."""
    u = np.zeros(len(px), complex)
    for q, (x, y) in enumerate(zip(px, py)):
        dx, dy = x - comp.x, y - comp.y
        r = np.hypot(dx, dy)
        inn = comp.dy * dx - comp.dx * dy
        u[q] = np.sum(comp.ws * (0.25j * k) * sp.hankel1(1, k * r) * inn / r * phi)
    return u


def _green_error(T, comps, k, srcs, px, py, potential=_potential):
    """max over sources of |u - u_inc| / max |u_inc| at the interior points,
    with T phi = -2 u_inc on the boundary."""
    X = np.concatenate([c.x for c in comps])
    Y = np.concatenate([c.y for c in comps])
    lu = sla.lu_factor(T)
    worst = 0.0
    for src in srcs:
        f = _phi(k, X, Y, src)
        phi = sla.lu_solve(lu, -2.0 * f)
        u = potential(comps, k, phi, px, py) if potential is _potential else \
            potential(comps[0], k, phi, px, py)
        ref = _phi(k, px, py, src)
        worst = max(worst, float(np.max(np.abs(u - ref)) / np.max(np.abs(ref))))
    return worst


def _cfie_exact(comps, k, slip=None):
    """The CFIE-Kress matrix with exact Hankels (any complex k; J_n from
    scipy), with one of the slips the Green check must catch:
      'sij'     the S-term takes the target's speed s_i instead of s_j
      'wij'     the source weight w_j is replaced by the target's w_i
      'kappa'   the diagonal curvature term divides by |gamma'|^3
      'speed'   the diagonal log term takes log(k^2 s^4 / 4)
      'offsign' the off-component double-layer term has the wrong sign.
    Without a slip it equals the oracle's matrix (checked in the test)."""
    offs = np.concatenate([[0], np.cumsum([len(c.x) for c in comps])])
    N = int(offs[-1])
    A = np.zeros((N, N), complex)
    for a, ca in enumerate(comps):
        for b, cb in enumerate(comps):
            DX = ca.x[:, None] - cb.x[None, :]
            DY = ca.y[:, None] - cb.y[None, :]
            r = np.hypot(DX, DY)
            sa = np.hypot(ca.dx, ca.dy)
            sb = np.hypot(cb.dx, cb.dy)
            inn = cb.dy[None, :] * DX - cb.dx[None, :] * DY
            wj = (ca.ws[:, None] if slip == "wij" else cb.ws[None, :]) * np.ones_like(r)
            sj = (sa[:, None] if slip == "sij" else sb[None, :]) * np.ones_like(r)
            if a == b:
                n = len(ca.x)
                np.fill_diagonal(r, 1.0)
                h0, h1 = sp.hankel1(0, k * r), sp.hankel1(1, k * r)
                j0, j1 = sp.jv(0, k * r), sp.jv(1, k * r)
                Rk = O.kress_R(n)
                with np.errstate(divide="ignore"):
                    lt = np.log(4.0 * np.sin((ca.ts[:, None] - ca.ts[None, :]) / 2.0) ** 2)
                np.fill_diagonal(lt, 0.0)
                cD = inn / r
                L1 = (-k / (2 * np.pi)) * j1
                D = Rk * cD * L1 + wj * cD * (0.5j * k * h1) - wj * lt * cD * L1
                M1 = (-1 / (2 * np.pi)) * j0
                S = Rk * sj * M1 + wj * sj * (0.5j * h0) - wj * sj * lt * M1
                blk = -(D + 1j * k * S)
                p = 3 if slip == "kappa" else 2
                kap = (-(ca.dx * ca.ddy - ca.dy * ca.ddx) / sa ** p) / (2 * np.pi)
                sl = sa ** 2 if slip == "speed" else sa
                m2 = ((0.5j - O.EULER_OVER_PI) - np.log((k * k / 4.0) * sl * sl) / (2 * np.pi)) * sa
                sval = np.diag(Rk) * (-sa / (2 * np.pi)) + ca.ws * m2
                blk[np.arange(n), np.arange(n)] = 1.0 - (ca.ws * kap + 1j * k * sval)
            else:
                h0, h1 = sp.hankel1(0, k * r), sp.hankel1(1, k * r)
                sgn = -1.0 if slip == "offsign" else 1.0
                blk = -(sgn * wj * inn / r * (0.5j * k * h1) + 1j * k * (wj * sj * 0.5j * h0))
            A[offs[a]:offs[a + 1], offs[b]:offs[b + 1]] = blk
    return A


_CASES = {
    # geometry, node counts, sources outside the domain, interior points
    "star": (dict(kind="star", Ns=[400]), [(1.8, 0.9), (-0.4, -1.6)],
             [(0.0, 0.0), (0.4, 0.2), (-0.3, 0.45), (0.1, -0.55)]),
    "ellipse": (dict(kind="ellipse", Ns=[320]), [(1.5, 0.8), (-1.4, -0.3)],
                [(0.0, 0.0), (0.5, 0.1), (-0.35, -0.25)]),
    "circle_hole": (dict(kind="circle_hole", Ns=[300, 120]), [(0.35, 0.05), (1.6, 0.4)],
                    [(-0.5, 0.2), (0.35, 0.6), (-0.2, -0.6)]),
}


@pytest.mark.parametrize("case", sorted(_CASES))
@pytest.mark.parametrize("k", [10.3, 10.3 + 0.2j, 9.7 - 0.15j])
def test_cfie_green_representation(case, k):
    """T(k) phi = -2 u_inc on the boundary, u_inc = (i/4) H0(k |x - x_src|)
    radiated from outside the domain (inside the hole, or outside the outer
    curve); the combined potential of phi reproduces u_inc at interior
    points to 1e-11 relative, with the oracle's Chebyshev-table assembly at
    real and complex k.  The star and the ellipse have non-constant speed
    and curvature, the disk with an off-centre hole two components with
    different weights and an off-component block with no symmetry.  Catches
    (each run below on a slipped operator, failing by > 1e-4):
    s_i <-> s_j in the S-term, w_i <-> w_j, the power of |gamma'| in the
    curvature term, the power of the speed in the diagonal log term, and the
    sign of the off-component double-layer term."""
    spec, srcs, ev = _CASES[case]
    comps = O.make_geometry(spec["kind"], spec["Ns"])
    T = O.cfie_assemble(comps, [k], O.ChebParams(8000, 9, 3000, 9))[0]
    px = np.array([p[0] for p in ev])
    py = np.array([p[1] for p in ev])
    err = _green_error(T, comps, k, srcs, px, py)
    print(case, k, "oracle", err)
    assert err < 1e-11
    # the slip-free exact builder is the oracle's operator ...
    E = _cfie_exact(comps, k)
    assert np.max(np.abs(E - T)) < 1e-10 * np.max(np.abs(T))
    # ... and every slip breaks the representation
    slips = ["sij", "kappa", "speed"] + (["wij", "offsign"] if len(comps) > 1 else [])
    for slip in slips:
        e = _green_error(_cfie_exact(comps, k, slip), comps, k, srcs, px, py)
        print(case, k, slip, e)
        assert e > 1e-4, slip


@pytest.mark.parametrize("kind,Ns", [("ellipse", [320]), ("star", [400])])
def test_dlp_green_representation_and_direct_twin(kind, Ns):
    """DLP-Kress: F = I - D with F phi = -2 u_inc, the double-layer potential
    of phi reproduces u_inc inside (1e-11), for the oracle's table assembly
    at real and complex k and for its exact-Hankel twin at real k
    (construct_fredholm_matrix!, dlp_kress.jl:906-932), which equals the
    table route to the tables' tolerance and I minus construct_dlp_matrix!
    (:802-828).  Catches the curvature power in D_ii = w_i kappa_i, a
    dropped log-term correction, and the Kress R weights on a non-circle."""
    comp = O.make_geometry(kind, Ns)[0]
    srcs = _CASES[kind][1]
    ev = _CASES[kind][2]
    px = np.array([p[0] for p in ev])
    py = np.array([p[1] for p in ev])
    cp = O.ChebParams(8000, 9, 3000, 9)
    for k in (10.3, 10.3 + 0.2j):
        F = O.dlp_assemble(comp, [k], cp)[0]
        err = _green_error(F, [comp], k, srcs, px, py, potential=_dl_potential)
        print(kind, k, "dlp", err)
        assert err < 1e-11
    k = 10.3
    Fd = O.dlp_direct_real(comp, k)
    Ft = O.dlp_assemble(comp, [k], cp)[0]
    assert np.max(np.abs(Fd - Ft)) < 1e-12 * np.max(np.abs(Fd))
    assert np.array_equal(Fd, np.eye(len(comp.x)) - O.dlp_matrix_direct_real(comp, k))
    assert _green_error(Fd, [comp], k, srcs, px, py, potential=_dl_potential) < 1e-9
    # slip: the curvature term with |gamma'|^3 (a speed power the circle hides)
    G = O.geom_cache(comp)
    s = G.speed
    Fs = Fd.copy()
    idx = np.arange(len(s))
    Fs[idx, idx] = 1.0 - comp.ws * G.kappa / s
    assert _green_error(Fs, [comp], k, srcs, px, py, potential=_dl_potential) > 1e-4


# --------------------------------------------------------------------------
# the column restatements used as checkers at full size
# --------------------------------------------------------------------------


def test_column_restatements_equal_the_full_assemblies():
    """dlp_assemble_cols (the checker of the DLP full-size GPU test),
    cfie_assemble_cols and the C restatement cfie_cols_c (the checker of the
    star N = 8192 GPU test and the CPU baseline) equal the full oracle
    assemblies column for column, with the same tables."""
    cp = O.ChebParams(3000, 7, 1500, 7)
    comp = O.make_geometry("ellipse", [256])[0]
    z = 30.0 + 0.3j
    F = O.dlp_assemble(comp, [z], cp)[0]
    rmin, rmax = O._rrange_single_fast(comp)
    assert np.max(np.abs(O.dlp_assemble_cols(comp, z, cp, 40, 72, rmin, rmax) - F[:, 40:72])) \
        < 1e-14 * np.max(np.abs(F))
    comp = O.make_geometry("star", [300])[0]
    for z in (25.0 + 0.2j, 25.0 - 0.45j, 3.0 + 0.1j):
        T = O.cfie_assemble([comp], [z], cp)[0]
        rmin, rmax = O._rrange_single_fast(comp)
        a = O.cfie_assemble_cols(comp, z, cp, 100, 140, rmin, rmax)
        assert np.max(np.abs(a - T[:, 100:140])) < 1e-14 * np.max(np.abs(T)), z
        if abs(z) * np.min(np.hypot(np.diff(comp.x), np.diff(comp.y))) >= 0.2:
            c = O.cfie_cols_c(comp, z, cp, 100, 140, rmin, rmax)
            assert np.max(np.abs(c - T[:, 100:140])) < 1e-13 * np.max(np.abs(T)), z
