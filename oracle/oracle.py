"""CPU oracle for the boundary-integral hot path of QuantumBilliards.jl.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module, and only as the checker
(or the timed CPU baseline).  The product path (quantumbilliards.jl_amd/) never
imports it.

It restates, in NumPy/SciPy, the reference's algorithm for:
  * geometry inputs and the Kress node rule
      src/solvers/sweepmethods/cfie/cfie.jl:424, :461-475, :535-561, :822-832
      src/solvers/sweepmethods/dlp/dlp_kress.jl:359-386
      src/billiards/{circle.jl:69-75, ellipse.jl:69-79, star_with_hole.jl:16-28,
                     circle_with_circle_holes.jl:43-54}
  * geometry cache and the Kress R matrix
      src/billiards/geometry/boundarypoints.jl:310-365, :823-884
  * block caches / panel lookup
      src/chebyshev/chebyshev_euclidian_helmholtz_cfie_kress.jl:362-448
      src/chebyshev/chebyshev_euclidian_helmholtz_dlp_kress.jl:119-182
      src/chebyshev/chebyshev_bessels.jl:574-577
  * Chebyshev tables (breaks, Lobatto samples, DCT-I fit, Clenshaw)
      src/chebyshev/chebyshev_core.jl:67-103, :114-314
      src/chebyshev/chebyshev_bessels.jl:331-372, :461-504
    with the samples taken from scipy.special.hankel1 / jv, which wrap the
    same AMOS routines (zbesh/zbesj) that SpecialFunctions.besselh/besselj wrap.
  * CFIE-Kress and DLP-Kress assembly
      src/chebyshev/chebyshev_euclidian_helmholtz_cfie_kress.jl:543-665, :1431-1450
      src/chebyshev/chebyshev_euclidian_helmholtz_dlp_kress.jl:702-769, :1086-1124
  * the exact-Hankel (direct, real k) twins
      src/solvers/sweepmethods/cfie/cfie_kress.jl:416-510
      src/solvers/sweepmethods/dlp/dlp_kress.jl:802-828, :906-932
  * the Beyn contour loop
      src/solvers/acceleratedmethods/beyn.jl:257-322, :364-397
    with LAPACK (scipy.linalg) for LU / solve / SVD / eig, as the reference
    uses LAPACK through Julia's LinearAlgebra.

How each part is pinned (besides the reference's own sources) is stated in
tests/test_oracle.py: closed forms (circle: circulant T(z) whose spectrum is
given by J_m H_m products; Beyn eigenvalues = zeros of J_m), brute-force
Kress R, DCT-I exactness, mpmath Bessel values, and LAPACK residuals.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla
import scipy.special as sp

TWO_PI = 2.0 * math.pi
INV_TWO_PI = 1.0 / TWO_PI
EULER = 0.57721566490153286061
EULER_OVER_PI = EULER / math.pi
H_CUT_SMALL = 1e-3   # chebyshev_bessels.jl:238
H_CUT = 0.2          # chebyshev_bessels.jl:239

# --------------------------------------------------------------------------
# Geometry (curves parametrised by t in [0, 2pi); derivatives w.r.t. t)
# --------------------------------------------------------------------------


def curve_circle(R=1.0, cx=0.0, cy=0.0):
    """make_circle / CircleSegment (circle.jl:69-75, circlesegment.jl:16-18,175-182)."""
    def f(t):
        c, s = np.cos(t), np.sin(t)
        return (R * c + cx, R * s + cy, -R * s, R * c, -R * c, -R * s)
    return f, TWO_PI * R


def curve_ellipse(a=1.0, b=0.6):
    """make_full_ellipse (ellipse.jl:69-79): r(u) = (a cos 2pi u, b sin 2pi u)."""
    def f(t):
        c, s = np.cos(t), np.sin(t)
        return (a * c, b * s, -a * s, b * c, -a * c, -b * s)
    L = 4.0 * a * sp.ellipe(1.0 - (b / a) ** 2) if a >= b else 4.0 * b * sp.ellipe(1.0 - (a / b) ** 2)
    return f, L


def curve_star(R=1.0, eps=0.2, m=5):
    """make_star_polar_boundary (star_with_hole.jl:16-28): r = R(1 + eps cos(m t))."""
    def f(t):
        c, s = np.cos(t), np.sin(t)
        r = R * (1.0 + eps * np.cos(m * t))
        r1 = -R * eps * m * np.sin(m * t)
        r2 = -R * eps * m * m * np.cos(m * t)
        x, y = r * c, r * s
        dx = r1 * c - r * s
        dy = r1 * s + r * c
        ddx = r2 * c - 2.0 * r1 * s - r * c
        ddy = r2 * s + 2.0 * r1 * c - r * s
        return (x, y, dx, dy, ddx, ddy)
    # perimeter by high-order periodic trapezoid (spectrally exact)
    tt = np.arange(4096) * (TWO_PI / 4096)
    _, _, dx, dy, _, _ = f(tt)
    L = float(np.sum(np.hypot(dx, dy)) * (TWO_PI / 4096))
    return f, L


@dataclass
class Comp:
    """One boundary component, the fields of BoundaryPointsCFIE the path reads
    (boundarypoints.jl:487-502)."""
    x: np.ndarray
    y: np.ndarray
    dx: np.ndarray
    dy: np.ndarray
    ddx: np.ndarray
    ddy: np.ndarray
    ts: np.ndarray
    ws: np.ndarray


def kress_nodes(fcurve, N: int) -> Comp:
    """_evaluate_points(::CFIE_kress) node rule (cfie.jl:424, :535-561):
    t_j = 2pi (j - 1/2)/N, ws = 2pi/N, derivatives rescaled to d/dt."""
    j = np.arange(1, N + 1, dtype=np.float64)
    ts = TWO_PI * (j - 0.5) / N
    x, y, dx, dy, ddx, ddy = fcurve(ts)
    return Comp(x, y, dx, dy, ddx, ddy, ts, np.full(N, TWO_PI / N))


def reverse_comp(c: Comp) -> Comp:
    """_reverse_component_orientation (cfie.jl:461-475): reverse xy, tangent
    reversed and negated, tangent_2 reversed, ts and ws reversed."""
    return Comp(c.x[::-1].copy(), c.y[::-1].copy(), -c.dx[::-1], -c.dy[::-1],
                c.ddx[::-1].copy(), c.ddy[::-1].copy(), c.ts[::-1].copy(), c.ws[::-1].copy())


def n_points(k: float, L: float, b: float, min_pts: int = 20) -> int:
    """N = max(min_pts, round(k L b / 2pi)) bumped to even (cfie.jl:537-549)."""
    N = max(min_pts, int(round(k * L * b / TWO_PI)))
    return N + (N % 2)


def make_geometry(kind: str, Ns, **kw):
    """Components for the workloads: 'circle', 'ellipse', 'star', 'circle_hole'."""
    if kind == "circle":
        return [kress_nodes(curve_circle(kw.get("R", 1.0))[0], Ns[0])]
    if kind == "ellipse":
        return [kress_nodes(curve_ellipse(kw.get("a", 1.0), kw.get("b", 0.6))[0], Ns[0])]
    if kind == "star":
        return [kress_nodes(curve_star(kw.get("R", 1.0), kw.get("eps", 0.2), kw.get("m", 5))[0], Ns[0])]
    if kind == "circle_hole":
        outer = kress_nodes(curve_circle(kw.get("R", 1.0))[0], Ns[0])
        rh, hx, hy = kw.get("hole", (0.3, 0.35, 0.0))
        hole = reverse_comp(kress_nodes(curve_circle(rh, hx, hy)[0], Ns[1]))
        return [outer, hole]
    raise ValueError(kind)


# --------------------------------------------------------------------------
# Kress R and the geometry cache
# --------------------------------------------------------------------------


def kress_R_col(N: int) -> np.ndarray:
    """First column of the Kress circulant (boundarypoints.jl:310-365):
    even N=2n: R_d = -(4pi/N) sum_{m=1}^{n-1} cos(2pi m d/N)/m - (4pi/N^2)(-1)^d,
    odd N=2n+1: R_d = -(4pi/N) sum_{m=1}^{n} cos(2pi m d/N)/m.
    R[i,j] = R_col[(i-j) mod N]."""
    d = np.arange(N)
    n = N // 2 if N % 2 == 0 else (N - 1) // 2
    mmax = n - 1 if N % 2 == 0 else n
    acc = np.zeros(N)
    for m in range(1, mmax + 1):
        acc += np.cos(TWO_PI * ((m * d) % N) / N) / m
    R = -(2.0 * TWO_PI / N) * acc
    if N % 2 == 0:
        R -= (2.0 * TWO_PI / (N * N)) * np.where(d % 2 == 0, 1.0, -1.0)
    return R


def kress_R(N: int) -> np.ndarray:
    col = kress_R_col(N)
    i = np.arange(N)
    return col[(i[:, None] - i[None, :]) % N]


@dataclass
class GeomCache:
    """cfie_geom_cache (boundarypoints.jl:854-884)."""
    R: np.ndarray
    invR: np.ndarray
    inner: np.ndarray
    logterm: np.ndarray
    speed: np.ndarray
    kappa: np.ndarray


def geom_cache(c: Comp) -> GeomCache:
    DX = c.x[:, None] - c.x[None, :]
    DY = c.y[:, None] - c.y[None, :]
    R = np.hypot(DX, DY)
    np.fill_diagonal(R, 1.0)
    invR = 1.0 / R
    np.fill_diagonal(invR, 0.0)
    inner = c.dy[None, :] * DX - c.dx[None, :] * DY
    DT = c.ts[:, None] - c.ts[None, :]
    with np.errstate(divide="ignore"):
        lt = np.log(4.0 * np.sin(DT / 2.0) ** 2)
    np.fill_diagonal(lt, 0.0)
    speed = np.sqrt(c.dx ** 2 + c.dy ** 2)
    kappa = INV_TWO_PI * (-(c.dx * c.ddy - c.dy * c.ddx) / (c.dx ** 2 + c.dy ** 2))
    return GeomCache(R, invR, inner, lt, speed, kappa)


# --------------------------------------------------------------------------
# Chebyshev tables
# --------------------------------------------------------------------------


def breaks_uniform(rmin: float, rmax: float, npan: int) -> np.ndarray:
    """_breaks_uniform (chebyshev_core.jl:95-103)."""
    h = (rmax - rmin) / npan
    b = np.arange(npan + 1) * h + rmin
    b[-1] = rmax
    return b


def lobatto_t(M: int) -> np.ndarray:
    """cospi(j/M), j = 0..M."""
    j = np.arange(M + 1)
    t = np.cos(np.pi * j / M)
    t[0] = 1.0
    t[-1] = -1.0
    if M % 2 == 0:
        t[M // 2] = 0.0
    return t


def chebfit(f: np.ndarray) -> np.ndarray:
    """_chebfit! (chebyshev_core.jl:67-80) on the last axis: DCT-I with
    endpoint half weights, c_0 halved."""
    M = f.shape[-1] - 1
    j = np.arange(M + 1)
    C = np.cos(np.pi * np.outer(j, j) / M)   # [m, j]
    w = np.ones(M + 1)
    w[0] = w[-1] = 0.5
    c = (2.0 / M) * np.einsum("...j,mj->...m", f * w, C)
    c[..., 0] *= 0.5
    return c


def build_table(kind: str, nu: int, k: complex, rmin: float, rmax: float, npan: int, M: int):
    """plan_h (bessels:461-471) / plan_j (bessels:494-504): coefficient array
    [npan, M+1] of H_nu^(1)(k r) or J_nu(k r) on uniform panels."""
    br = breaks_uniform(rmin, rmax, npan)
    a, b = br[:-1, None], br[1:, None]
    t = lobatto_t(M)[None, :]
    r = ((b + a) + (b - a) * t) / 2.0
    z = complex(k) * r
    f = sp.hankel1(nu, z) if kind == "h" else sp.jv(nu, z)
    return chebfit(f)


def clenshaw(c: np.ndarray, t: np.ndarray) -> np.ndarray:
    """_cheb_clenshaw (chebyshev_core.jl:287-314); c[..., m] per point."""
    M = c.shape[-1] - 1
    b1 = np.zeros(t.shape, np.complex128)
    b2 = np.zeros(t.shape, np.complex128)
    u = 2.0 * t
    for m in range(M, 0, -1):
        b0 = (u * b1 - b2) + c[..., m]
        b2 = b1
        b1 = b0
    return (t * b1 - b2) + c[..., 0]


@dataclass
class Panels:
    rmin: float
    rmax: float
    npan: int
    M: int

    @property
    def dr(self):
        return (self.rmax - self.rmin) / self.npan

    def locate(self, r: np.ndarray):
        """_find_panel_uniform (bessels:574-577) + t = (2r-(b+a))/(b-a)
        (cfie-kress cheb :437-444).  Returns 0-based panel index."""
        invdr = 1.0 / self.dr
        p = np.floor((r - self.rmin) * invdr).astype(np.int64) + 1
        p = np.clip(p, 1, self.npan)
        h = (self.rmax - self.rmin) / self.npan
        a = (p - 1) * h + self.rmin
        b = np.where(p == self.npan, self.rmax, p * h + self.rmin)
        t = (2.0 * r - (b + a)) / (b - a)
        return p - 1, t


@dataclass
class ChebParams:
    n_panels_h: int = 15000
    M_h: int = 5
    n_panels_j: int = 3000
    M_j: int = 5


def _rrange_blocks(comps):
    """Global padded [rmin, rmax] over all component blocks
    (cfie-kress cheb :374-418; dlp-kress cheb :135-147)."""
    rmin, rmax = np.inf, 0.0
    allx = [c.x for c in comps]
    ally = [c.y for c in comps]
    eps = np.finfo(np.float64).eps
    for a in range(len(comps)):
        for b in range(len(comps)):
            R = np.hypot(allx[a][:, None] - allx[b][None, :], ally[a][:, None] - ally[b][None, :])
            if a == b:
                np.fill_diagonal(R, 0.0)
            m = R > eps
            rmin = min(rmin, float(np.min(R[m])) * 0.95)
            rmax = max(rmax, float(np.max(R[m])) * 1.05)
    return rmin, rmax


def _rrange_single_fast(c: Comp):
    """Padded [rmin, rmax] of one component, computed in column blocks."""
    rmin, rmax = np.inf, 0.0
    n = len(c.x)
    for j0 in range(0, n, 1024):
        js = np.arange(j0, min(n, j0 + 1024))
        R = np.hypot(c.x[:, None] - c.x[None, js], c.y[:, None] - c.y[None, js])
        R[js, np.arange(len(js))] = 0.0
        m = R > np.finfo(float).eps
        rmin = min(rmin, float(np.min(R[m])))
        rmax = max(rmax, float(np.max(R[m])))
    return rmin * 0.95, rmax * 1.05


def _hankel_direct(z):
    """Direct branch of h0_h1_multi_ks_at_r! (bessels:1190-1206): |z|<1e-3
    small series, else AMOS."""
    return sp.hankel1(0, z), sp.hankel1(1, z)


# --------------------------------------------------------------------------
# Assembly
# --------------------------------------------------------------------------


def cfie_assemble_cols(comp: Comp, z: complex, cp: ChebParams, j0: int, j1: int, rmin: float,
                       rmax: float, exact: bool = False):
    """Columns [j0, j1) of the single-component CFIE matrix at one z (the
    same-block formulas of cheb-CFIE :579-615), for bounded CPU timing at
    large N without N x N caches.  exact=True: H0, H1, J0, J1 from AMOS at
    every distance instead of the tables."""
    n = len(comp.x)
    if not exact:
        ph = Panels(rmin, rmax, cp.n_panels_h, cp.M_h)
        pj = Panels(0.0, rmax, cp.n_panels_j, cp.M_j)
        H0 = build_table("h", 0, z, rmin, rmax, cp.n_panels_h, cp.M_h)
        H1 = build_table("h", 1, z, rmin, rmax, cp.n_panels_h, cp.M_h)
        J0 = build_table("j", 0, z, 0.0, rmax, cp.n_panels_j, cp.M_j)
        J1 = build_table("j", 1, z, 0.0, rmax, cp.n_panels_j, cp.M_j)
    js = np.arange(j0, j1)
    DX = comp.x[:, None] - comp.x[None, js]
    DY = comp.y[:, None] - comp.y[None, js]
    r = np.hypot(DX, DY)
    diag = np.arange(n)[:, None] == js[None, :]
    r[diag] = 1.0
    invr = 1.0 / r
    inner = comp.dy[None, js] * DX - comp.dx[None, js] * DY
    d = (np.arange(n)[:, None] - js[None, :]) % n
    Rk = kress_R_col(n)[d]
    with np.errstate(divide="ignore"):
        lt = np.log(4.0 * np.sin((comp.ts[:, None] - comp.ts[None, js]) / 2.0) ** 2)
    lt[diag] = 0.0
    if exact:
        h0, h1 = sp.hankel1(0, z * r), sp.hankel1(1, z * r)
        j0v, j1v = sp.jv(0, z * r), sp.jv(1, z * r)
    else:
        pidx, th = ph.locate(r)
        pjx, tj = pj.locate(r)
        h0, h1 = clenshaw(H0[pidx], th), clenshaw(H1[pidx], th)
        j0v, j1v = clenshaw(J0[pjx], tj), clenshaw(J1[pjx], tj)
        # |k r| < 0.2: direct evaluation (h0_h1_j0_j1_multi_ks_at_r!, bessels:1155-1160)
        dm = (np.abs(z * r) < H_CUT) & ~diag
        if np.any(dm):
            h0[dm], h1[dm] = _hankel_direct(z * r[dm])
    sj = np.sqrt(comp.dx[js] ** 2 + comp.dy[js] ** 2)[None, :]
    wj = comp.ws[js][None, :]
    cD = inner * invr
    aL1, aL2, ik = -z * INV_TWO_PI, 0.5j * z, 1j * z
    D = (Rk * cD) * (aL1 * j1v) + (wj * cD) * (aL2 * h1) - (wj * lt * cD) * (aL1 * j1v)
    S = (Rk * sj) * (-INV_TWO_PI * j0v) + (wj * sj) * (0.5j * h0) - (wj * sj * lt) * (-INV_TWO_PI * j0v)
    A = -(D + ik * S)
    # diagonal limits (cheb-CFIE :585-591) for the diagonal entries in the block
    ii = js[(js >= 0) & (js < n)]
    s_ = np.sqrt(comp.dx[ii] ** 2 + comp.dy[ii] ** 2)
    kap = INV_TWO_PI * (-(comp.dx[ii] * comp.ddy[ii] - comp.dy[ii] * comp.ddx[ii]) / s_ ** 2)
    w_ = comp.ws[ii]
    m2 = ((0.5j - EULER_OVER_PI) - INV_TWO_PI * np.log((z * z / 4.0) * (s_ * s_))) * s_
    sval = kress_R_col(n)[0] * (-INV_TWO_PI * s_) + w_ * m2
    A[ii, ii - j0] = 1.0 - (w_ * kap + ik * sval)
    return A


def dlp_assemble_cols(comp: Comp, z: complex, cp: ChebParams, j0: int, j1: int, rmin: float,
                      rmax: float):
    """Columns [j0, j1) of F(z) = I - D(z) for DLP-Kress at one z
    (cheb-DLP :702-769, :1086-1095), without N x N caches."""
    n = len(comp.x)
    ph = Panels(rmin, rmax, cp.n_panels_h, cp.M_h)
    pj = Panels(0.0, rmax, cp.n_panels_j, cp.M_j)
    H1 = build_table("h", 1, z, rmin, rmax, cp.n_panels_h, cp.M_h)
    J1 = build_table("j", 1, z, 0.0, rmax, cp.n_panels_j, cp.M_j)
    js = np.arange(j0, j1)
    DX = comp.x[:, None] - comp.x[None, js]
    DY = comp.y[:, None] - comp.y[None, js]
    r = np.hypot(DX, DY)
    diag = np.arange(n)[:, None] == js[None, :]
    r[diag] = 1.0
    inner = comp.dy[None, js] * DX - comp.dx[None, js] * DY
    d = (np.arange(n)[:, None] - js[None, :]) % n
    Rk = kress_R_col(n)[d]
    with np.errstate(divide="ignore"):
        lt = np.log(4.0 * np.sin((comp.ts[:, None] - comp.ts[None, js]) / 2.0) ** 2)
    lt[diag] = 0.0
    pidx, th = ph.locate(r)
    pjx, tj = pj.locate(r)
    h1, j1v = clenshaw(H1[pidx], th), clenshaw(J1[pjx], tj)
    cD = inner / r
    wj = comp.ws[js][None, :]
    l1, ht = (-z * INV_TWO_PI) * j1v, (0.5j * z) * h1
    D = (Rk * cD) * l1 + (wj * cD) * ht - (wj * lt * cD) * l1
    ii = js
    kap = INV_TWO_PI * (-(comp.dx[ii] * comp.ddy[ii] - comp.dy[ii] * comp.ddx[ii])
                        / (comp.dx[ii] ** 2 + comp.dy[ii] ** 2))
    D[ii, ii - j0] = comp.ws[ii] * kap
    F = -D
    F[ii, ii - j0] += 1.0
    return F


def _liboracle():
    import ctypes
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle.so")
    L = ctypes.CDLL(path)
    L.qbo_cfie_cols.restype = None
    return L


def cfie_cols_c(comp: Comp, z: complex, cp: ChebParams, j0: int, j1: int, rmin: float,
                rmax: float, tables=None):
    """Same as cfie_assemble_cols, through the C/OpenMP restatement in
    oracle/qb_oracle.c (tables built here with AMOS).  Returns (N, j1-j0)."""
    import ctypes as C
    n = len(comp.x)
    if tables is None:
        H0 = build_table("h", 0, z, rmin, rmax, cp.n_panels_h, cp.M_h)
        H1 = build_table("h", 1, z, rmin, rmax, cp.n_panels_h, cp.M_h)
        J0 = build_table("j", 0, z, 0.0, rmax, cp.n_panels_j, cp.M_j)
        J1 = build_table("j", 1, z, 0.0, rmax, cp.n_panels_j, cp.M_j)
        tables = (np.ascontiguousarray(np.stack([H0, H1], axis=-1)),
                  np.ascontiguousarray(np.stack([J0, J1], axis=-1)))
    H, J = tables
    out = np.empty((j1 - j0, n), np.complex128)
    rk = np.ascontiguousarray(kress_R_col(n))
    lt = np.zeros(n)
    d = np.arange(1, n)
    lt[1:] = np.log(4.0 * np.sin(np.pi * d / n) ** 2)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
            (comp.x, comp.y, comp.dx, comp.dy, comp.ddx, comp.ddy, comp.ws)]
    P = lambda a: C.c_void_p(a.ctypes.data)
    _liboracle().qbo_cfie_cols(C.c_int64(n), *[P(a) for a in arrs], P(rk), P(lt),
                               C.c_int64(j0), C.c_int64(j1), C.c_double(z.real),
                               C.c_double(z.imag), P(H), C.c_int(cp.n_panels_h), C.c_int(cp.M_h),
                               C.c_double(rmin), C.c_double(rmax), P(J),
                               C.c_int(cp.n_panels_j), C.c_int(cp.M_j), P(out))
    return out.T


def cfie_assemble(comps, zs, cp: ChebParams = ChebParams(), tables=None, cols=None,
                  exact: bool = False):
    """construct_boundary_matrices! for CFIE_kress (cheb-CFIE:1431-1450 ->
    _all_k_nosymm_CFIE_chebyshev! :543-665).  Returns [nz, N, N] with
    A(z) = I - (D + i z S), row = target, column = source.  exact=True takes
    H0, H1, J0, J1 from AMOS at every distance instead of the Chebyshev
    tables (the same formulas; the tables' interpolation error removed) --
    the exact-Hankel route of cfie_kress.jl:416-510 carried to complex z."""
    zs = np.asarray(zs, dtype=np.complex128)
    if exact:
        return _cfie_assemble_exact(comps, zs)
    if cols is not None:
        assert len(comps) == 1
        c = comps[0]
        rmin, rmax = _rrange_single_fast(c)
        return np.stack([cfie_assemble_cols(c, z, cp, cols[0], cols[1], rmin, rmax) for z in zs])
    rmin, rmax = _rrange_blocks(comps)
    ph = Panels(rmin, rmax, cp.n_panels_h, cp.M_h)
    pj = Panels(0.0, rmax, cp.n_panels_j, cp.M_j)
    offs = np.concatenate([[0], np.cumsum([len(c.x) for c in comps])])
    N = int(offs[-1])
    out = np.zeros((len(zs), N, N), np.complex128)
    caches = [geom_cache(c) for c in comps]
    for q, k in enumerate(zs):
        if tables is None:
            H0 = build_table("h", 0, k, rmin, rmax, cp.n_panels_h, cp.M_h)
            H1 = build_table("h", 1, k, rmin, rmax, cp.n_panels_h, cp.M_h)
            J0 = build_table("j", 0, k, 0.0, rmax, cp.n_panels_j, cp.M_j)
            J1 = build_table("j", 1, k, 0.0, rmax, cp.n_panels_j, cp.M_j)
        else:
            H0, H1, J0, J1 = tables[q]
        aL1, aL2 = -k * INV_TWO_PI, 0.5j * k
        aM1, aM2 = -INV_TWO_PI, 0.5j
        ik = 1j * k
        for a, ca in enumerate(comps):
            for b, cb in enumerate(comps):
                ro, co = offs[a], offs[b]
                if a == b:
                    G = caches[a]
                    n = len(ca.x)
                    Rk = kress_R(n)
                    r = G.R
                    pidx, th = ph.locate(r)
                    pjx, tj = pj.locate(r)
                    h0 = clenshaw(H0[pidx], th)
                    h1 = clenshaw(H1[pidx], th)
                    j0 = clenshaw(J0[pjx], tj)
                    j1 = clenshaw(J1[pjx], tj)
                    z = k * r
                    az = np.abs(z)
                    dm = (az < H_CUT)
                    if np.any(dm & ~np.eye(n, dtype=bool)):
                        hh0, hh1 = _hankel_direct(z[dm])
                        h0[dm], h1[dm] = hh0, hh1
                    sj = G.speed[None, :]
                    wj = cb.ws[None, :]
                    cD = G.inner * G.invR
                    L1 = aL1 * j1
                    L2 = aL2 * h1
                    M1 = aM1 * j0
                    M2 = aM2 * h0
                    lt = G.logterm
                    D = (Rk * cD) * L1 + (wj * cD) * L2 - (wj * lt * cD) * L1
                    S = (Rk * sj) * M1 + (wj * sj) * M2 - (wj * sj * lt) * M1
                    A = -(D + ik * S)
                    # diagonal (cheb-CFIE :585-591)
                    s = G.speed
                    w = ca.ws
                    dval = w * G.kappa
                    m1 = aM1 * s
                    m2 = ((0.5j - EULER_OVER_PI) - INV_TWO_PI * np.log((k * k / 4.0) * (s * s))) * s
                    sval = np.diag(Rk) * m1 + w * m2
                    A[np.arange(n), np.arange(n)] = 1.0 - (dval + ik * sval)
                    out[q, ro:ro + n, co:co + n] = A
                else:
                    DX = ca.x[:, None] - cb.x[None, :]
                    DY = ca.y[:, None] - cb.y[None, :]
                    r = np.hypot(DX, DY)
                    invr = np.where(r > np.finfo(float).eps, 1.0 / r, 0.0)
                    inn = cb.dy[None, :] * DX - cb.dx[None, :] * DY
                    pidx, th = ph.locate(r)
                    h0 = clenshaw(H0[pidx], th)
                    h1 = clenshaw(H1[pidx], th)
                    z = k * r
                    dm = np.abs(z) < H_CUT
                    if np.any(dm):
                        hh0, hh1 = _hankel_direct(z[dm])
                        h0[dm], h1[dm] = hh0, hh1
                    sjb = np.sqrt(cb.dx ** 2 + cb.dy ** 2)[None, :]
                    wj = cb.ws[None, :]
                    cD = wj * inn * invr
                    cS = wj * sjb
                    A = -(cD * aL2 * h1 + ik * (cS * aM2 * h0))
                    out[q, ro:ro + len(ca.x), co:co + len(cb.x)] = A
    return out


def _cfie_assemble_exact(comps, zs):
    offs = np.concatenate([[0], np.cumsum([len(c.x) for c in comps])])
    N = int(offs[-1])
    out = np.zeros((len(zs), N, N), np.complex128)
    caches = [geom_cache(c) for c in comps]
    for q, k in enumerate(zs):
        aL1, aL2, aM1, aM2, ik = -k * INV_TWO_PI, 0.5j * k, -INV_TWO_PI, 0.5j, 1j * k
        for a, ca in enumerate(comps):
            for b, cb in enumerate(comps):
                ro, co = offs[a], offs[b]
                if a == b:
                    G = caches[a]
                    n = len(ca.x)
                    Rk = kress_R(n)
                    z = k * G.R
                    h0, h1 = sp.hankel1(0, z), sp.hankel1(1, z)
                    j0, j1 = sp.jv(0, z), sp.jv(1, z)
                    sj = G.speed[None, :]
                    wj = cb.ws[None, :]
                    cD = G.inner * G.invR
                    lt = G.logterm
                    D = (Rk * cD) * (aL1 * j1) + (wj * cD) * (aL2 * h1) - (wj * lt * cD) * (aL1 * j1)
                    S = (Rk * sj) * (aM1 * j0) + (wj * sj) * (aM2 * h0) - (wj * sj * lt) * (aM1 * j0)
                    A = -(D + ik * S)
                    s_ = G.speed
                    m2 = ((0.5j - EULER_OVER_PI) - INV_TWO_PI * np.log((k * k / 4.0) * (s_ * s_))) * s_
                    sval = np.diag(Rk) * (aM1 * s_) + ca.ws * m2
                    A[np.arange(n), np.arange(n)] = 1.0 - (ca.ws * G.kappa + ik * sval)
                    out[q, ro:ro + n, co:co + n] = A
                else:
                    DX = ca.x[:, None] - cb.x[None, :]
                    DY = ca.y[:, None] - cb.y[None, :]
                    r = np.hypot(DX, DY)
                    inn = cb.dy[None, :] * DX - cb.dx[None, :] * DY
                    h0, h1 = sp.hankel1(0, k * r), sp.hankel1(1, k * r)
                    sjb = np.sqrt(cb.dx ** 2 + cb.dy ** 2)[None, :]
                    wj = cb.ws[None, :]
                    A = -((wj * inn / r) * aL2 * h1 + ik * ((wj * sjb) * aM2 * h0))
                    out[q, ro:ro + len(ca.x), co:co + len(cb.x)] = A
    return out


def dlp_assemble_exact(comp: Comp, zs):
    """dlp_assemble with H1, J1 from AMOS at every distance (no tables)."""
    zs = np.asarray(zs, dtype=np.complex128)
    G = geom_cache(comp)
    n = len(comp.x)
    Rk = kress_R(n)
    out = np.zeros((len(zs), n, n), np.complex128)
    cD = G.inner * G.invR
    w = comp.ws
    for q, k in enumerate(zs):
        z = k * G.R
        h1, j1 = sp.hankel1(1, z), sp.jv(1, z)
        l1 = (-k * INV_TWO_PI) * j1
        D = (Rk * cD) * l1 + (w[None, :] * cD) * ((0.5j * k) * h1) - (w[None, :] * G.logterm * cD) * l1
        D[np.arange(n), np.arange(n)] = w * G.kappa
        out[q] = np.eye(n) - D
    return out


def dlp_assemble(comp: Comp, zs, cp: ChebParams = ChebParams(), tables=None):
    """construct_boundary_matrices! for DLP_kress (cheb-DLP:1110-1124 ->
    :702-769 then F = I - D, :1086-1095)."""
    zs = np.asarray(zs, dtype=np.complex128)
    G = geom_cache(comp)
    n = len(comp.x)
    Rk = kress_R(n)
    R = G.R.copy()
    np.fill_diagonal(R, 0.0)
    m = R > np.finfo(float).eps
    rmin = float(np.min(R[m])) * 0.95
    rmax = float(np.max(R[m])) * 1.05
    ph = Panels(rmin, rmax, cp.n_panels_h, cp.M_h)
    pj = Panels(0.0, rmax, cp.n_panels_j, cp.M_j)
    pidx, th = ph.locate(G.R)
    pjx, tj = pj.locate(G.R)
    out = np.zeros((len(zs), n, n), np.complex128)
    cD = G.inner * G.invR
    w = comp.ws
    for q, k in enumerate(zs):
        if tables is None:
            H1 = build_table("h", 1, k, rmin, rmax, cp.n_panels_h, cp.M_h)
            J1 = build_table("j", 1, k, 0.0, rmax, cp.n_panels_j, cp.M_j)
        else:
            H1, J1 = tables[q]
        h1 = clenshaw(H1[pidx], th)
        j1 = clenshaw(J1[pjx], tj)
        l1 = (-k * INV_TWO_PI) * j1
        ht = (0.5j * k) * h1
        D = (Rk * cD) * l1 + (w[None, :] * cD) * ht - (w[None, :] * G.logterm * cD) * l1
        D[np.arange(n), np.arange(n)] = w * G.kappa
        out[q] = np.eye(n) - D
    return out


def _tables4(k, rmin, rmax, cp):
    return (build_table("h", 0, k, rmin, rmax, cp.n_panels_h, cp.M_h),
            build_table("h", 1, k, rmin, rmax, cp.n_panels_h, cp.M_h),
            build_table("j", 0, k, 0.0, rmax, cp.n_panels_j, cp.M_j),
            build_table("j", 1, k, 0.0, rmax, cp.n_panels_j, cp.M_j))


def _h_cut(k, r, h0, h1):
    """|kr| < 0.2: direct values (h0_h1_j0_j1_multi_ks_at_r!, bessels:1155-1160);
    r == 0 (the diagonal, handled by its limit formulas) is left alone."""
    z = k * r
    dm = (np.abs(z) < H_CUT) & (r > 0)
    if np.any(dm):
        hh0, hh1 = _hankel_direct(z[dm])
        h0[dm], h1[dm] = hh0, hh1
    return h0, h1


def cfie_assemble_deriv(comps, zs, cp: ChebParams = ChebParams(),
                        reference_offblock_sign: bool = False):
    """construct_boundary_matrices_with_derivatives! for CFIE_kress
    (cheb-CFIE:1507-1530 -> _all_k_nosymm_CFIE_chebyshev_deriv!, :881-1059).
    Returns (A, A', A'') each [nz, N, N], derivatives in the wavenumber.

    Between components the reference writes the double-layer derivative
    terms with a minus sign, dval1 = -(i/2) w_j inn k H0 and
    dval2 = -(i/2) w_j inn (H0 - k r H1) (cheb-CFIE:1031-1032), while the
    derivative of its own entry dval = w_j (inn/r)(i k/2) H1(kr) (:641) is
    +(i/2) w_j inn k H0.  The default here is the derivative of the matrix;
    reference_offblock_sign=True reproduces the reference's sign."""
    zs = np.asarray(zs, dtype=np.complex128)
    rmin, rmax = _rrange_blocks(comps)
    ph = Panels(rmin, rmax, cp.n_panels_h, cp.M_h)
    pj = Panels(0.0, rmax, cp.n_panels_j, cp.M_j)
    offs = np.concatenate([[0], np.cumsum([len(c.x) for c in comps])])
    N = int(offs[-1])
    out = [np.zeros((len(zs), N, N), np.complex128) for _ in range(3)]
    caches = [geom_cache(c) for c in comps]
    for q, k in enumerate(zs):
        H0, H1, J0, J1 = _tables4(k, rmin, rmax, cp)
        aL1, aL2, aM1, aM2, ik = -k * INV_TWO_PI, 0.5j * k, -INV_TWO_PI, 0.5j, 1j * k
        for a, ca in enumerate(comps):
            for b, cb in enumerate(comps):
                ro, co = offs[a], offs[b]
                na, nb = len(ca.x), len(cb.x)
                if a == b:
                    G = caches[a]
                    Rk = kress_R(na)
                    r = G.R
                    pidx, th = ph.locate(r)
                    pjx, tj = pj.locate(r)
                    h0, h1 = _h_cut(k, r, clenshaw(H0[pidx], th), clenshaw(H1[pidx], th))
                    j0, j1 = clenshaw(J0[pjx], tj), clenshaw(J1[pjx], tj)
                    sj = G.speed[None, :]
                    wj = cb.ws[None, :]
                    lt = G.logterm
                    inn, invr = G.inner, G.invR
                    kr = k * r
                    L1, L2, M1, M2 = aL1 * j1, aL2 * h1, aM1 * j0, aM2 * h0
                    cR, cW = Rk * inn * INV_TWO_PI, wj * inn * INV_TWO_PI
                    dval = (Rk * inn * invr) * L1 + (wj * inn * invr) * L2 - (wj * lt * inn * invr) * L1
                    dval1 = -cR * k * j0 + cW * k * (lt * j0 + 1j * np.pi * h0)
                    dval2 = cR * (kr * j1 - j0) + cW * (lt * (j0 - kr * j1) + 1j * np.pi * (h0 - kr * h1))
                    sval = (Rk * sj) * M1 + (wj * sj) * M2 - (wj * sj * lt) * M1
                    sval1 = (r * sj * INV_TWO_PI) * (Rk * j1 - wj * (lt * j1 + 1j * np.pi * h1))
                    sval2 = (r * sj * INV_TWO_PI / k) * (Rk * (kr * j0 - j1) + wj * (
                        lt * (j1 - kr * j0) + 1j * np.pi * (h1 - kr * h0)))
                    A = -(dval + ik * sval)
                    A1 = -(dval1 + 1j * sval + ik * sval1)
                    A2 = -(dval2 + 2j * sval1 + ik * sval2)
                    # diagonal (cheb-CFIE:922-941)
                    s, w = G.speed, ca.ws
                    m2 = ((0.5j - EULER_OVER_PI) - INV_TWO_PI * np.log((k * k / 4.0) * (s * s))) * s
                    sv = np.diag(Rk) * (aM1 * s) + w * m2
                    sv1 = w * (-s / (np.pi * k))
                    sv2 = w * (s / (np.pi * k * k))
                    ii = np.arange(na)
                    A[ii, ii] = 1.0 - (w * G.kappa + ik * sv)
                    A1[ii, ii] = -(1j * sv + ik * sv1)
                    A2[ii, ii] = -(2j * sv1 + ik * sv2)
                else:
                    DX = ca.x[:, None] - cb.x[None, :]
                    DY = ca.y[:, None] - cb.y[None, :]
                    r = np.hypot(DX, DY)
                    invr = 1.0 / r
                    inn = cb.dy[None, :] * DX - cb.dx[None, :] * DY
                    pidx, th = ph.locate(r)
                    h0, h1 = _h_cut(k, r, clenshaw(H0[pidx], th), clenshaw(H1[pidx], th))
                    sjb = np.sqrt(cb.dx ** 2 + cb.dy ** 2)[None, :]
                    wj = cb.ws[None, :]
                    cD, cD1, cS = wj * inn * invr, wj * inn, wj * sjb
                    sg = -1.0 if reference_offblock_sign else 1.0
                    dval = cD * aL2 * h1
                    dval1 = sg * 0.5j * cD1 * k * h0
                    dval2 = sg * 0.5j * cD1 * (h0 - k * r * h1)
                    sval = cS * aM2 * h0
                    sval1 = -0.5j * cS * r * h1
                    sval2 = 0.5j * cS * r * (h1 - k * r * h0) / k
                    A = -(dval + ik * sval)
                    A1 = -(dval1 + 1j * sval + ik * sval1)
                    A2 = -(dval2 + 2j * sval1 + ik * sval2)
                for o, X in zip(out, (A, A1, A2)):
                    o[q, ro:ro + na, co:co + nb] = X
    return tuple(out)


def dlp_assemble_deriv(comp: Comp, zs, cp: ChebParams = ChebParams()):
    """construct_boundary_matrices_with_derivatives! for DLP_kress
    (cheb-DLP:1126-1144 -> :881-952, then F = I - D, F' = -D', F'' = -D'',
    :1097-1108).  H0, H1, J0, J1 with the |kr| < 0.2 direct branch."""
    zs = np.asarray(zs, dtype=np.complex128)
    G = geom_cache(comp)
    n = len(comp.x)
    Rk = kress_R(n)
    R = G.R.copy()
    np.fill_diagonal(R, 0.0)
    m = R > np.finfo(float).eps
    rmin, rmax = float(np.min(R[m])) * 0.95, float(np.max(R[m])) * 1.05
    ph = Panels(rmin, rmax, cp.n_panels_h, cp.M_h)
    pj = Panels(0.0, rmax, cp.n_panels_j, cp.M_j)
    pidx, th = ph.locate(G.R)
    pjx, tj = pj.locate(G.R)
    inn, invr, lt, r = G.inner, G.invR, G.logterm, G.R
    wj = comp.ws[None, :]
    out = [np.zeros((len(zs), n, n), np.complex128) for _ in range(3)]
    ii = np.arange(n)
    for q, k in enumerate(zs):
        H0, H1, J0, J1 = _tables4(k, rmin, rmax, cp)
        h0, h1 = _h_cut(k, r, clenshaw(H0[pidx], th), clenshaw(H1[pidx], th))
        j0, j1 = clenshaw(J0[pjx], tj), clenshaw(J1[pjx], tj)
        cD = inn * invr
        l1 = (-k * INV_TWO_PI) * j1
        D = (Rk * cD) * l1 + (wj * cD) * ((0.5j * k) * h1) - (wj * lt * cD) * l1
        D1 = (k * INV_TWO_PI) * (Rk * (-inn) * j0 + wj * (lt * inn * j0 + 1j * np.pi * inn * h0))
        u, v = j0 - k * r * j1, h0 - k * r * h1
        D2 = (INV_TWO_PI * inn) * (Rk * (-u) + wj * (lt * u + 1j * np.pi * v))
        D[ii, ii] = comp.ws * G.kappa
        D1[ii, ii] = 0.0
        D2[ii, ii] = 0.0
        out[0][q] = np.eye(n) - D
        out[1][q] = -D1
        out[2][q] = -D2
    return tuple(out)


# --------------------------------------------------------------------------
# Symmetry reduction (the reduced Kress workspaces)
# --------------------------------------------------------------------------
def sym_scales(kind: str, n: int = 0, m: int = 0, parity=1) -> np.ndarray:
    """Orbit weights: [1, p] for one reflection, [1, px, py, px py] for both
    (apply_symmetry_orbit_from_maps, symmetry.jl:776-787), and the rotation
    characters cis(2 pi (m mod n) l / n) (_rotation_tables, symmetry.jl:64-82)."""
    if kind == "rot":
        return np.array([np.exp(2j * np.pi * ((m % n) * l) / n) for l in range(n)])
    if kind == "origin":
        px, py = parity
        return np.array([1, px, py, px * py], dtype=np.complex128)
    return np.array([1, parity], dtype=np.complex128)


def orbits_index(N: int, kind: str, n: int = 0, m: int = 0, parity=1):
    """periodic_symmetry_index_orbits (symmetry.jl:214-298), the DLP_kress
    orbits, as (img [m x G] 0-based with img[:, 0] = Ifund, scale [m x G])."""
    rows = []
    if kind == "rot":
        for q1 in range(1, N // n + 1):
            rows.append([((q1 + l * (N // n)) - 1) % N for l in range(n)])
    elif kind in ("x_axis", "y_axis"):
        for q1 in range(1, N // 2 + 1):
            rows.append([q1 - 1, ((N - q1 + 1) - 1) % N])
    else:
        for q1 in range(1, N // 4 + 1):
            rows.append([q1 - 1, ((N // 2 - q1 + 1) - 1) % N, ((N - q1 + 1) - 1) % N,
                         ((q1 + N // 2) - 1) % N])
    img = np.array(rows, dtype=np.int64)
    return img, np.tile(sym_scales(kind, n, m, parity), (len(rows), 1))


def orbits_coords(comps, kind: str, n: int = 0, m: int = 0, parity=1, center=(0.0, 0.0),
                  tol: float = 1e-10):
    """The CFIE_kress orbits (build_CFIE_kress_reduced_workspace,
    cheb-CFIE:475-505): node maps by coordinates (build_symmetry_maps,
    symmetry.jl:746-766: match_index for reflections, cyclic shifts per
    component for rotations, :648-730), nodes visited in order.  Brute force
    over all node pairs."""
    x = np.concatenate([c.x for c in comps])
    y = np.concatenate([c.y for c in comps])
    N = len(x)

    def match(tx, ty, tl):
        d = (x[None, :] - tx[:, None]) ** 2 + (y[None, :] - ty[:, None]) ** 2
        hit = d < tl * tl
        if not np.all(hit.sum(axis=1) == 1):
            raise ValueError("no unique match")
        return np.argmax(hit, axis=1)

    if kind == "rot":
        cx, cy = center
        maps = [np.arange(N)]
        for l in range(1, n):
            th = 2 * np.pi * l / n
            c, s = math.cos(th), math.sin(th)
            maps.append(match(cx + c * (x - cx) - s * (y - cy), cy + s * (x - cx) + c * (y - cy),
                              max(1e-8, tol)))
    elif kind == "y_axis":
        maps = [np.arange(N), match(-x, y, tol)]
    elif kind == "x_axis":
        maps = [np.arange(N), match(x, -y, tol)]
    else:
        maps = [np.arange(N), match(-x, y, tol), match(x, -y, tol), match(-x, -y, tol)]
    M = np.stack(maps, axis=1)
    seen = np.zeros(N, bool)
    rows = []
    for g in range(N):
        if not seen[g]:
            rows.append(M[g])
            seen[M[g]] = True
    img = np.array(rows, dtype=np.int64)
    return img, np.tile(sym_scales(kind, n, m, parity), (len(rows), 1))


def reduce_matrix(A: np.ndarray, img: np.ndarray, scale: np.ndarray) -> np.ndarray:
    """Reduced matrix from the full one, entry by entry as the reduced
    assembly sums it (cheb-CFIE:745-820, cheb-DLP:806-841):
    A_red(a, b) = sum_l scale[b, l] A(img[a, 0], img[b, l])."""
    rep = img[:, 0]
    return np.einsum("bl,alb->ab", scale, A[rep][:, img.T])


def cauchy_derivatives(assemble, z: complex, rho: float = 0.25, n: int = 24):
    """dT/dz and d2T/dz2 of an analytic matrix function from its values on
    the circle z + rho e^{i theta_m} (trapezoid rule of Cauchy's integral
    formula, exponentially accurate).  `assemble(zs)` returns [len(zs), N, N]."""
    th = 2 * np.pi * np.arange(n) / n
    e = np.exp(1j * th)
    F = assemble(z + rho * e)
    d1 = np.tensordot(np.exp(-1j * th), F, axes=(0, 0)) / (n * rho)
    d2 = 2 * np.tensordot(np.exp(-2j * th), F, axes=(0, 0)) / (n * rho * rho)
    return d1, d2


def dlp_matrix_direct_real(comp: Comp, k: float):
    """construct_dlp_matrix! (dlp_kress.jl:802-828): the Kress-split
    double-layer matrix D at real k with exact Hankels, J_1 = real(H_1);
    D_ii = w_i kappa_i."""
    G = geom_cache(comp)
    n = len(comp.x)
    Rk = kress_R(n)
    h1 = sp.hankel1(1, k * G.R)
    j1 = h1.real
    cD = G.inner * G.invR
    w = comp.ws[None, :]
    l1 = (-k * INV_TWO_PI) * cD * j1
    l2 = (0.5j * k) * cD * h1 - l1 * G.logterm
    D = Rk * l1 + w * l2
    D[np.arange(n), np.arange(n)] = comp.ws * G.kappa
    return D


def dlp_direct_real(comp: Comp, k: float):
    """construct_fredholm_matrix! (dlp_kress.jl:906-932): F = I - D with D as
    in dlp_matrix_direct_real, at real k with exact Hankels."""
    return np.eye(len(comp.x)) - dlp_matrix_direct_real(comp, k)


def cfie_direct_real(comps, k: float):
    """Exact-Hankel CFIE at real k (cfie_kress.jl:416-510), one component
    (the twin of the Chebyshev path, with J_n = real(H_n) for real k)."""
    assert len(comps) == 1
    c = comps[0]
    G = geom_cache(c)
    n = len(c.x)
    Rk = kress_R(n)
    r = G.R
    h0 = sp.hankel1(0, k * r)
    h1 = sp.hankel1(1, k * r)
    j0, j1 = h0.real, h1.real
    cD = G.inner * G.invR
    wj = c.ws[None, :]
    sj = G.speed[None, :]
    aL1, aL2 = -k * INV_TWO_PI, 0.5j * k
    D = (Rk * cD) * (aL1 * j1) + (wj * cD) * (aL2 * h1) - (wj * G.logterm * cD) * (aL1 * j1)
    S = (Rk * sj) * (-INV_TWO_PI * j0) + (wj * sj) * (0.5j * h0) - (wj * sj * G.logterm) * (-INV_TWO_PI * j0)
    A = -(D + 1j * k * S)
    s = G.speed
    dval = c.ws * G.kappa
    m2 = ((0.5j - EULER_OVER_PI) - INV_TWO_PI * np.log((k * k / 4.0) * (s * s))) * s
    sval = np.diag(Rk) * (-INV_TWO_PI * s) + c.ws * m2
    A[np.arange(n), np.arange(n)] = 1.0 - (dval + 1j * k * sval)
    return A


# --------------------------------------------------------------------------
# Beyn contour loop
# --------------------------------------------------------------------------


def contour(k0: complex, R: float, nq: int):
    """Nodes and weights (beyn.jl:258-261)."""
    th = TWO_PI * np.arange(nq) / nq
    e = np.exp(1j * th)
    return k0 + R * e, (R / nq) * e


def beyn_moments(Ts, zj, wj, V):
    """LU + solve + axpy loop (beyn.jl:270-285)."""
    A0 = np.zeros(V.shape, np.complex128)
    A1 = np.zeros(V.shape, np.complex128)
    for T, z, w in zip(Ts, zj, wj):
        lu = sla.lu_factor(T, check_finite=False)
        X = sla.lu_solve(lu, V, check_finite=False)
        A0 += w * X
        A1 += (w * z) * X
    return A0, A1


def beyn_B(A0, A1, svd_tol):
    """Thin SVD, absolute rank threshold, B = U_k^H A1 W_k Sigma_k^{-1}
    (beyn.jl:286-321)."""
    U, S, Wh = np.linalg.svd(A0, full_matrices=False)
    rk = int(np.count_nonzero(S >= svd_tol))
    if rk == 0:
        return np.zeros((0, 0), np.complex128), U[:, :0], S, rk
    Uk = U[:, :rk]
    Wk = Wh.conj().T[:, :rk]
    B = Uk.conj().T @ ((A1 @ Wk) / S[:rk][None, :])
    return B, Uk, S, rk


def beyn_solve(assemble, k0: complex, R: float, nq: int, V: np.ndarray, svd_tol: float):
    """construct_B_matrix + eigen!(B) (beyn.jl:257-322, :386-397), without the
    rank-saturation retry.  `assemble(zs)` returns [nq, N, N]."""
    zj, wj = contour(k0, R, nq)
    Ts = assemble(zj)
    A0, A1 = beyn_moments(Ts, zj, wj, V)
    B, Uk, S, rk = beyn_B(A0, A1, svd_tol)
    lam = np.linalg.eigvals(B) if rk > 0 else np.zeros(0, np.complex128)
    return lam, S, rk, A0, A1


def jm_zeros_in_disk(k0: float, R: float, mmax: int = 200):
    """Dirichlet circle eigenvalues (unit radius) inside |k-k0|<=R: zeros of
    J_m, with multiplicity 2 for m>0."""
    out = []
    for m in range(mmax):
        zs = sp.jn_zeros(m, 60)
        for z in zs:
            if abs(z - k0) <= R:
                out += [z] * (1 if m == 0 else 2)
        if zs[0] > k0 + R:
            break
    return np.sort(np.array(out))


# --------------------------------------------------------------------------
# Residual check of Beyn eigenpairs
# --------------------------------------------------------------------------


def residual_select(T_of, lam, Uk, Y, k0: complex, R: float, res_tol: float, epss: float = 1e-15,
                    auto_discard_spurious: bool = True):
    """residual_and_norm_select with matnorm=:one (beyn.jl:557-593): for every
    lam_j inside the disk, phi_j = Uk Y[:, j], y = T(lam_j) phi_j,
    tens = ||y||_2, tensN = ||y||_1 / (||T||_1 (||phi||_1 + epss) + epss);
    kept if not (auto_discard_spurious and tens >= res_tol).  `T_of(z)`
    returns the N x N matrix T(z).  Returns (idx, tens[idx], tensN[idx]) and
    the per-candidate (y norms, operator norms) for every inside index."""
    idx, tens, tensN, detail = [], [], [], {}
    for j, lj in enumerate(lam):
        if abs(lj - k0) > R:
            continue
        T = T_of(lj)
        phi = Uk @ Y[:, j]
        y = T @ phi
        r2 = float(np.linalg.norm(y))
        nA = float(np.max(np.sum(np.abs(T), axis=0)))
        ninf = float(np.max(np.sum(np.abs(T), axis=1)))
        tn = float(np.sum(np.abs(y))) / (nA * (float(np.sum(np.abs(phi))) + epss) + epss)
        detail[j] = (r2, float(np.sum(np.abs(y))), float(np.max(np.abs(y))), nA, ninf)
        if auto_discard_spurious and r2 >= res_tol:
            continue
        idx.append(j)
        tens.append(r2)
        tensN.append(tn)
    return np.array(idx, int), np.array(tens), np.array(tensN), detail
