"""Billiard geometries and the Kress boundary discretisation (host side).

Restates, for the geometries the hot path is run on, what the reference's
`evaluate_points` produces for the Kress solvers:
  * node rule t_j = 2pi (j - 1/2)/N, ws = 2pi/N, N = max(min_pts, round(k L b / 2pi))
    bumped to even (src/solvers/sweepmethods/cfie/cfie.jl:424, :535-561;
    dlp_kress.jl:359-386);
  * derivatives rescaled to the periodic parameter t in [0, 2pi)
    (cfie.jl:553-554);
  * holes (components after the first) orientation-reversed
    (_reverse_component_orientation, cfie.jl:461-475; evaluate_points,
    cfie.jl:822-832);
  * composite smooth components (the stadium) through the global arc-length
    parameter (_evaluate_points_smooth_composite, cfie.jl:682-714;
    _eval_composite_geom_global_t, boundarypoints.jl:727-767).
Curves are closed forms (the reference differentiates its PolarSegment
r-functions with ForwardDiff, polarsegment.jl:98-156, which is exact).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

TWO_PI = 2.0 * math.pi


@dataclass
class Curve:
    """A smooth closed curve; `geom(t)` returns x, y, x', y', x'', y'' with
    derivatives taken w.r.t. t in [0, 2pi)."""
    geom: Callable
    length: float
    name: str = "curve"


@dataclass
class Segment:
    """A smooth open segment of a composite component, parametrised by u in
    [0, 1]; `geom(u)` returns point, d/du and d2/du2 (6 arrays)."""
    geom: Callable
    length: float


@dataclass
class Billiard:
    """full_boundary: list of components; each component is [Curve] (smooth
    closed) or a list of Segments forming one C1 closed curve."""
    full_boundary: list
    area: float
    length: float
    name: str = "billiard"
    area_fundamental: Optional[float] = None  # Weyl area of a symmetry sector


def _circle_curve(R, cx=0.0, cy=0.0):
    def g(t):
        c, s = np.cos(t), np.sin(t)
        return (R * c + cx, R * s + cy, -R * s, R * c, -R * c, -R * s)
    return Curve(g, TWO_PI * R, "circle")


def make_circle(R: float = 1.0, x0: float = 0.0, y0: float = 0.0) -> Billiard:
    """make_circle (src/billiards/circle.jl:69-75)."""
    return Billiard([[_circle_curve(R, x0, y0)]], math.pi * R * R, TWO_PI * R, "circle")


def _ellipse_perimeter(a, b):
    # Ramanujan-free: spectrally exact periodic trapezoid of |gamma'|
    t = np.arange(8192) * (TWO_PI / 8192)
    return float(np.sum(np.hypot(a * np.sin(t), b * np.cos(t))) * (TWO_PI / 8192))


def make_full_ellipse(a: float = 1.0, b: float = 0.6) -> Billiard:
    """make_full_ellipse (src/billiards/ellipse.jl:69-79): (a cos t, b sin t)."""
    def g(t):
        c, s = np.cos(t), np.sin(t)
        return (a * c, b * s, -a * s, b * c, -a * c, -b * s)
    L = _ellipse_perimeter(a, b)
    return Billiard([[Curve(g, L, "ellipse")]], math.pi * a * b, L, "ellipse")


def make_star_polar_boundary(R: float = 1.0, eps: float = 0.2, m: int = 5) -> Billiard:
    """make_star_polar_boundary (src/billiards/star_with_hole.jl:16-28):
    r(t) = R (1 + eps cos(m t))."""
    def g(t):
        c, s = np.cos(t), np.sin(t)
        r = R * (1.0 + eps * np.cos(m * t))
        r1 = -R * eps * m * np.sin(m * t)
        r2 = -R * eps * m * m * np.cos(m * t)
        return (r * c, r * s, r1 * c - r * s, r1 * s + r * c,
                r2 * c - 2.0 * r1 * s - r * c, r2 * s + 2.0 * r1 * c - r * s)
    t = np.arange(8192) * (TWO_PI / 8192)
    _, _, dx, dy, _, _ = g(t)
    L = float(np.sum(np.hypot(dx, dy)) * (TWO_PI / 8192))
    area = math.pi * R * R * (1.0 + 0.5 * eps * eps)
    return Billiard([[Curve(g, L, "star")]], area, L, "star")


def CircularHoleBilliard(R: float = 1.0, holes: Sequence = ((0.3, 0.35, 0.0),)) -> Billiard:
    """CircularHoleBilliard / make_circle_with_holes
    (src/billiards/circle_with_circle_holes.jl:43-54, :135-167); holes are
    (r, hx, hy)."""
    comps = [[_circle_curve(R)]]
    area = math.pi * R * R
    L = TWO_PI * R
    for (r, hx, hy) in holes:
        comps.append([_circle_curve(r, hx, hy)])
        area -= math.pi * r * r
        L += TWO_PI * r
    return Billiard(comps, area, L, "circle_hole")


def make_full_stadium(half_width: float = 1.0, radius: float = 1.0) -> Billiard:
    """Full stadium (src/billiards/stadium.jl:23-41): one C1 closed component
    of five segments starting at (half_width + radius, 0), counter-clockwise:
    quarter arc, top edge (right to left), half arc, bottom edge (left to
    right), quarter arc."""
    a, r = half_width, radius

    def line(p0, p1):
        p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
        d = p1 - p0
        def g(u):
            u = np.asarray(u, float)
            z = np.zeros_like(u)
            return (p0[0] + d[0] * u, p0[1] + d[1] * u, d[0] + z, d[1] + z, z, z)
        return Segment(g, float(np.hypot(*d)))

    def arc(cx, cy, arc_angle, shift):
        # CircleSegment(radius, arc_angle, shift_angle, x0, y0) (circlesegment.jl:16-18)
        def g(u):
            u = np.asarray(u, float)
            th = arc_angle * u + shift
            c, s = np.cos(th), np.sin(th)
            return (cx + r * c, cy + r * s, -arc_angle * r * s, arc_angle * r * c,
                    -arc_angle ** 2 * r * c, -arc_angle ** 2 * r * s)
        return Segment(g, arc_angle * r)

    segs = [arc(a, 0.0, math.pi / 2, 0.0), line((a, r), (-a, r)),
            arc(-a, 0.0, math.pi, math.pi / 2), line((-a, -r), (a, -r)),
            arc(a, 0.0, math.pi / 2, 3 * math.pi / 2)]
    L = sum(s.length for s in segs)
    return Billiard([segs], math.pi * r * r + 4 * a * r, L, "stadium")


@dataclass
class BoundaryPointsCFIE:
    """Fields of BoundaryPointsCFIE (boundarypoints.jl:487-502) that the
    Kress hot path reads."""
    xy: np.ndarray          # (N, 2)
    tangent: np.ndarray     # (N, 2)
    tangent_2: np.ndarray   # (N, 2)
    ts: np.ndarray
    ws: np.ndarray
    compid: int = 1

    def __len__(self):
        return len(self.ts)


def _component_length(comp) -> float:
    return float(sum(c.length for c in comp))


def _nodes(comp, N: int, idx: int) -> BoundaryPointsCFIE:
    j = np.arange(1, N + 1, dtype=np.float64)
    ts = TWO_PI * (j - 0.5) / N
    if len(comp) == 1 and isinstance(comp[0], Curve):
        x, y, dx, dy, ddx, ddy = comp[0].geom(ts)
    else:
        # composite: global arc-length parameter (boundarypoints.jl:727-767)
        lens = np.array([s.length for s in comp])
        cum = np.concatenate([[0.0], np.cumsum(lens)])
        Ltot = cum[-1]
        s = (ts / TWO_PI) * Ltot
        seg = np.clip(np.searchsorted(cum, s, side="right") - 1, 0, len(comp) - 1)
        x = np.empty(N); y = np.empty(N); dx = np.empty(N); dy = np.empty(N)
        ddx = np.empty(N); ddy = np.empty(N)
        for q, sg in enumerate(comp):
            m = seg == q
            if not np.any(m):
                continue
            u = np.clip((s[m] - cum[q]) / lens[q], 0.0, 1.0)
            gx, gy, gdx, gdy, gddx, gddy = sg.geom(u)
            du = Ltot / (TWO_PI * lens[q])
            x[m], y[m] = gx, gy
            dx[m], dy[m] = gdx * du, gdy * du
            ddx[m], ddy[m] = gddx * du * du, gddy * du * du
    return BoundaryPointsCFIE(np.column_stack([x, y]), np.column_stack([dx, dy]),
                              np.column_stack([ddx, ddy]), ts, np.full(N, TWO_PI / N), idx)


def reverse_orientation(p: BoundaryPointsCFIE) -> BoundaryPointsCFIE:
    """_reverse_component_orientation (cfie.jl:461-475)."""
    return BoundaryPointsCFIE(p.xy[::-1].copy(), -p.tangent[::-1], p.tangent_2[::-1].copy(),
                              p.ts[::-1].copy(), p.ws[::-1].copy(), p.compid)


def n_kress(k: float, L: float, b: float, min_pts: int = 20, divisor: int = 2) -> int:
    """max(min_pts, round(k L b / 2pi)), bumped up to a multiple of `divisor`:
    2 without symmetry, 4 with a reflection, lcm(2, n) with a C_n rotation
    (cfie.jl:537-549, dlp_kress.jl:361-372)."""
    N = max(min_pts, int(round(k * L * b / TWO_PI)))
    rem = N % divisor
    return N + (divisor - rem if rem else 0)


def evaluate_points(solver, billiard: Billiard, k: float,
                    N: Optional[Sequence[int]] = None) -> List[BoundaryPointsCFIE]:
    """evaluate_points for CFIE_kress / DLP_kress (cfie.jl:822-832,
    dlp_kress.jl:563-606).  `N` (one count per component) overrides the
    points-per-wavelength rule, for workloads that fix N."""
    from .symmetry import node_divisor
    comps = billiard.full_boundary
    div = node_divisor(getattr(solver, "symmetry", None))
    out = []
    for idx, comp in enumerate(comps):
        n = int(N[idx]) if N is not None else n_kress(k, _component_length(comp),
                                                     solver.pts_scaling_factor, solver.min_pts,
                                                     div)
        p = _nodes(comp, n, idx + 1)
        out.append(p if idx == 0 else reverse_orientation(p))
    return out
