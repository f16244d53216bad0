"""Host-side logic of the serving loop (no GPU): request streams and node counts."""
import math

import numpy as np

from qbhip import geometry as G
from qbhip.serve import node_counts, poisson_stream


def test_poisson_stream_is_seeded_and_in_range():
    a = poisson_stream(50, rate=2.0, seed=3)
    b = poisson_stream(50, rate=2.0, seed=3)
    assert [(r.shape, r.k0, r.R, r.tol, r.arrival) for r in a] == \
           [(r.shape, r.k0, r.R, r.tol, r.arrival) for r in b]
    assert all(50 <= r.k0 <= 1000 and 0.25 <= r.R <= 0.5 and r.tol in (1e-8, 1e-12) for r in a)
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    assert {r.shape for r in a} <= {"ellipse", "stadium", "star"}


def test_node_counts_follow_points_per_wavelength():
    b = G.CircularHoleBilliard(1.0, holes=[(0.3, 0.35, 0.0)])
    n = node_counts(b, 100.0, 10.0)
    assert n == [int(math.ceil(1000.0)) + (-int(math.ceil(1000.0))) % 2,
                 int(math.ceil(300.0)) + (-int(math.ceil(300.0))) % 2]
    assert all(x % 2 == 0 for x in n)
