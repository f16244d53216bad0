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


def test_uneven_workload_node_classes():
    """The uneven batched-window workload: N = 10 k L / 2 pi with L = 2.6 pi,
    rounded up to a multiple of 256 (1536 at k = 100.5, 10496 at k = 800.5),
    split 10 : 3 between the outer circle and the hole, both even."""
    from qbhip.serve import hole_node_counts, uneven_windows
    assert sum(hole_node_counts(100.5)) == 1536 and sum(hole_node_counts(800.5)) == 10496
    for k in np.linspace(100.5, 800.5, 57):
        no, nh = hole_node_counts(k)
        n = no + nh
        assert n % 256 == 0 and no % 2 == 0 and nh % 2 == 0
        assert 13 * k <= n < 13 * k + 256
        assert abs(no / n - 10 / 13) < 2 / n
    a = uneven_windows(128, seed=4)
    b = uneven_windows(128, seed=4)
    assert [(r.k0, r.N) for r in a] == [(r.k0, r.N) for r in b]
    assert all(100 <= r.k0 <= 800 and r.R == 0.5 and r.nq == 32 for r in a)
    assert all(sum(r.N) == sum(hole_node_counts(r.k0 + r.R)) for r in a)
    # windows share node counts: they can be batched
    classes = {tuple(r.N) for r in a}
    assert len(classes) < len(a)


def test_real_roots_filter():
    """summarize() counts the candidates with |Im k| <= REAL_TOL |k|."""
    from qbhip.serve import REAL_TOL, real_roots
    lam = np.array([500.0 + 1e-12j, 500.1 - 4e-5j, 499.9 + 0.02j, 500.2 + 0j,
                    500.3 + 0.9 * REAL_TOL * 500.3j, 500.4 - 1.1 * REAL_TOL * 500.4j])
    got = real_roots(lam)
    assert np.array_equal(got, lam[[0, 3, 4]])
    assert real_roots(np.zeros(0, complex)).size == 0


def test_rank_threshold_follows_the_solve_tolerance():
    """The rank threshold of a moment block is the absolute svd_tol for
    tol 1e-12 at sigma_max <= 1 and rises to RANK_NOISE tol sigma_max for a
    looser tolerance, so solve noise at 1e-8 does not count as rank: a block
    with singular values 1, 1e-3 and a noise floor of 1e-7 has rank 2 at
    tol 1e-8 and rank 4 (saturated) at tol 1e-12."""
    from qbhip.serve import RANK_NOISE, _saturated, rank_threshold
    rng = np.random.default_rng(3)
    Q1, _ = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    Q2, _ = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    S = np.array([1.0, 1e-3, 2e-7, 1e-7])
    Rb = np.triu(np.linalg.qr((Q1 * S) @ Q2)[1])
    assert rank_threshold(Rb, 1e-10, 1e-12) == 1e-10
    assert math.isclose(rank_threshold(Rb, 1e-10, 1e-8), RANK_NOISE * 1e-8, rel_tol=1e-9)
    assert _saturated(Rb[None], 1e-10, 1e-12)
    assert not _saturated(Rb[None], 1e-10, 1e-8)
    assert not _saturated((Rb * 1e-4)[None], 1e-10, 1e-12)


def test_b_from_qr_relative_threshold_counts_the_rank_threshold():
    """B_from_qr with rel_tol = RANK_NOISE tol keeps exactly the singular
    values at or above rank_threshold: one SVD instead of two in the serving
    loop's host stage."""
    from qbhip.beyn import B_from_qr
    from qbhip.serve import RANK_NOISE, rank_threshold
    rng = np.random.default_rng(4)
    Q1, _ = np.linalg.qr(rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6)))
    Q2, _ = np.linalg.qr(rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6)))
    S = np.array([2.0, 0.5, 1e-3, 3e-6, 1e-7, 1e-12])
    Rb = np.triu(np.linalg.qr((Q1 * S) @ Q2)[1])
    Cm = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    for tol in (1e-12, 1e-8):
        thr = rank_threshold(Rb, 1e-10, tol)
        B, Ur, Sv, rk = B_from_qr(Rb, Cm, 1e-10, RANK_NOISE * tol)
        assert rk == int(np.count_nonzero(Sv >= thr)) and B.shape == (rk, rk)
    assert B_from_qr(Rb, Cm, 1e-10, RANK_NOISE * 1e-12)[3] == 5
    assert B_from_qr(Rb, Cm, 1e-10, RANK_NOISE * 1e-8)[3] == 4
