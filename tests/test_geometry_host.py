"""The stadium's Kress nodes on the CPU, pinned by closed forms.

make_full_stadium (src/billiards/stadium.jl:23-41) is one C1 component of five
segments walked through the global arc-length parameter
(boundarypoints.jl:727-767, cfie.jl:682-714). The oracle's geometry list has no
stadium, so the nodes are checked against what the mathematics fixes:
  * every node lies on the stadium: |y| = r where |x| <= a, and
    (|x| - a)^2 + y^2 = r^2 on the caps;
  * the parameter is proportional to arc length: |gamma'(t)| = L / 2pi at
    every node, with L = 4a + 2 pi r (a wrong du scale breaks it);
  * the signed curvature (x'y'' - y'x'') / |gamma'|^3 is 1/r on the caps and 0
    on the edges (a wrong second-derivative scale or sign breaks it);
  * sum ws |gamma'| = L and 1/2 sum ws (x y' - y x') = pi r^2 + 4 a r, with a
    positive sign for counter-clockwise order (the trapezoid rule converges as
    h^2 here: the curvature jumps at the four joints);
  * the node set is invariant under x -> -x and y -> -y for N a multiple of 4
    (the node rule with a reflection), node j mapping to N+1-j and
    N/2+1-j (mod N); a wrong segment order or start point breaks it.
"""
import math

import numpy as np
import pytest

from qbhip import geometry as G


def _stadium_nodes(a, r, N):
    b = G.make_full_stadium(a, r)
    p = G.evaluate_points(None, b, 1.0, N=[N])[0]
    return b, p


@pytest.mark.parametrize("a,r,N", [(1.0, 1.0, 4000), (0.5, 1.3, 2400), (2.0, 0.4, 3200)])
def test_stadium_nodes_closed_forms(a, r, N):
    b, p = _stadium_nodes(a, r, N)
    L = 4 * a + 2 * math.pi * r
    assert abs(b.length - L) < 1e-13 * L and abs(b.area - (math.pi * r * r + 4 * a * r)) < 1e-13
    x, y = p.xy[:, 0], p.xy[:, 1]
    dx, dy = p.tangent[:, 0], p.tangent[:, 1]
    ddx, ddy = p.tangent_2[:, 0], p.tangent_2[:, 1]
    edge = np.abs(x) <= a
    assert np.max(np.abs(np.abs(y[edge]) - r)) < 1e-13
    assert np.max(np.abs((np.abs(x[~edge]) - a) ** 2 + y[~edge] ** 2 - r * r)) < 1e-13
    # arc-length parametrisation
    speed = np.hypot(dx, dy)
    assert np.max(np.abs(speed - L / (2 * math.pi))) < 1e-12 * L
    # curvature: 1/r on the caps, 0 on the edges (nodes away from the joints)
    kappa = (dx * ddy - dy * ddx) / speed ** 3
    inner = np.abs(x) < a * (1 - 1e-9)
    assert np.max(np.abs(kappa[~edge] - 1.0 / r)) < 1e-12 / r
    assert np.max(np.abs(kappa[inner])) < 1e-12 / r
    # perimeter, area and orientation from the quadrature
    assert abs(np.sum(p.ws * speed) - L) < 1e-12 * L
    area = 0.5 * np.sum(p.ws * (x * dy - y * dx))
    assert area > 0 and abs(area - b.area) < 10 * (L / N) ** 2 * b.area
    # first node just above the start point (a + r, 0)
    assert x[0] > a and 0 < y[0] < 2 * L / N


@pytest.mark.parametrize("N", [400, 1000])
def test_stadium_nodes_mirror_under_both_reflections(N):
    _, p = _stadium_nodes(1.0, 1.0, N)
    xy = p.xy
    j = np.arange(N)
    # y -> -y: node j <-> N-1-j (0-based)
    assert np.max(np.abs(xy[N - 1 - j] - xy * [1, -1])) < 1e-13
    # x -> -x: node j <-> N/2-1-j (mod N)
    assert np.max(np.abs(xy[(N // 2 - 1 - j) % N] - xy * [-1, 1])) < 1e-13
    # a plausible mistake -- the segments walked from the top edge, so the
    # start point is (a, r) -- breaks the y mirror
    b = G.make_full_stadium(1.0, 1.0)
    segs = b.full_boundary[0]
    bad = G.Billiard([segs[1:] + segs[:1]], b.area, b.length, "stadium")
    q = G.evaluate_points(None, bad, 1.0, N=[N])[0].xy
    assert np.max(np.abs(q[N - 1 - j] - q * [1, -1])) > 0.1
