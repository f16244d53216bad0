"""Several disks on one geometry as one window (qb_window_create_multi): the
contour nodes of all disks are assembled, factored and solved as one batch,
and every disk's moments, singular values and eigenvalues are those of a
window of its own."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CHEB = dict(n_panels_h=6000, M_h=7, n_panels_j=2000, M_j=7)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_multi_disk_window_matches_separate_windows(prec):
    import qbhip as q
    b = q.CircularHoleBilliard(1.0, holes=[(0.3, 0.35, 0.0)])
    solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(solver, b, 25.0, N=[200, 60])
    geom = q.Geometry(pts, solver=solver)
    k0s, Rs, nq, r = [23.5, 25.0, 26.2], [0.35, 0.4, 0.3], 24, 48
    V = q.probe_block(np.random.default_rng(5), geom.dim, r)
    w = q.Window(solver, pts, k0s, Rs, nq, precision=prec, geom=geom, **CHEB)
    assert w.nwin == 3 and w.nblocks == 3
    A0, A1 = w.moments(V)
    assert A0.shape == (3, geom.dim, r)
    Rm, Cm = w.moments_qr(V)
    assert Rm.shape == (3, r, r)
    found = 0
    for c, (k0, R) in enumerate(zip(k0s, Rs)):
        w1 = q.Window(solver, pts, k0, R, nq, precision=prec, geom=geom, **CHEB)
        B0, B1 = w1.moments(V)
        s0 = np.max(np.abs(B0))
        assert np.max(np.abs(A0[c] - B0)) <= 1e-11 * s0, c
        assert np.max(np.abs(A1[c] - B1)) <= 1e-11 * np.max(np.abs(B1)), c
        R1, C1 = w1.moments_qr(V)
        w1.close()
        B, Ur, S, rk = q.B_from_qr(Rm[c], Cm[c], 1e-10)
        Bs, Urs, S1, rk1 = q.B_from_qr(R1, C1, 1e-10)
        assert rk == rk1 and np.max(np.abs(S - S1)) <= 1e-12 * S1[0], c
        lam, lam1 = np.linalg.eigvals(B), np.linalg.eigvals(Bs)
        lam = np.sort_complex(lam[np.abs(lam - k0) <= R])
        lam1 = np.sort_complex(lam1[np.abs(lam1 - k0) <= R])
        assert len(lam) == len(lam1), c
        if len(lam):
            assert np.max(np.abs(lam - lam1)) <= 1e-10 * k0, c
        found += len(lam)
        # the basis of each block is orthonormal and spans its own moments
        U = w.basis(Ur, irrep=c)
        assert np.max(np.abs(U.conj().T @ U - np.eye(rk))) < 1e-12
    w.close()
    assert found > 0


def test_multi_disk_window_on_the_circle_gives_the_bessel_zeros():
    """Each disk of a multi-disk window gives the zeros of J_m.  The disks at
    19.0 and 20.0 (R = 0.5) share the real node 19.5: the conjugate-pair
    assembly must assemble both copies (each node pairs with itself or with
    one unused conjugate)."""
    import qbhip as q
    b = q.make_circle(1.0)
    solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(solver, b, 20.0, N=[256])
    k0s = [19.0, 20.0, 21.3]
    w = q.Window(solver, pts, k0s, 0.5, 32, precision="mixed",
                 n_panels_h=4000, M_h=5, n_panels_j=1500, M_j=5)
    V = q.probe_block(np.random.default_rng(2), 256, 40)
    Rm, Cm = w.moments_qr(V)
    w.close()
    for c, k0 in enumerate(k0s):
        B, _, _, _ = q.B_from_qr(Rm[c], Cm[c], 1e-10)
        lam = np.linalg.eigvals(B)
        got = np.sort(lam[np.abs(lam - k0) <= 0.5].real)
        ex = O.jm_zeros_in_disk(k0, 0.5)
        assert len(got) == len(ex) and np.max(np.abs(got - ex) / ex) < 1e-10, (k0, got, ex)


def test_multi_disk_window_arguments():
    import qbhip as q
    b = q.make_circle(1.0)
    solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(solver, b, 20.0, N=[128])
    with pytest.raises(Exception, match="radius"):
        q.Window(solver, pts, [20.0, 21.0], [0.5, 0.0], 8)
    with pytest.raises(Exception, match="4096"):
        q.Window(solver, pts, [20.0] * 200, 0.5, 32)
