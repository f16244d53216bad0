"""GPU tests of the spectrum driver: the device residual check of Beyn
eigenpairs against the oracle, the spurious-root filters and the Weyl-planned
sweep on the unit circle (eigenvalues = zeros of J_m)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def qb():
    import qbhip
    return qbhip


CP = dict(n_panels_h=3000, M_h=5, n_panels_j=1000, M_j=5)


@pytest.mark.parametrize("kind,Ns,k0", [("circle", [256], 20.0), ("circle_hole", [200, 60], 25.0)])
def test_residual_norms_match_oracle(kind, Ns, k0):
    """y = T(lam) phi and the operator norms for random complex lam and phi,
    against the oracle's T(lam) (same tables) and NumPy norms."""
    q = qb()
    if kind == "circle":
        b = q.make_circle(1.0)
    else:
        b = q.CircularHoleBilliard(1.0, holes=[(0.3, 0.35, 0.0)])
    solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(solver, b, k0, N=Ns)
    N = sum(Ns)
    rng = np.random.default_rng(7)
    lam = k0 + rng.uniform(-0.4, 0.4, 5) + 1j * rng.uniform(-0.05, 0.05, 5)
    Phi = rng.standard_normal((N, 5)) + 1j * rng.standard_normal((N, 5))
    r2, r1, rinf, n1, ninf = q.residual_norms(solver, pts, lam, Phi, batch=2, **CP)
    comps = O.make_geometry(kind, Ns)
    Tref = O.cfie_assemble(comps, lam, O.ChebParams(CP["n_panels_h"], CP["M_h"],
                                                     CP["n_panels_j"], CP["M_j"]))
    for qi in range(5):
        T = Tref[qi]
        y = T @ Phi[:, qi]
        ref = (np.linalg.norm(y), np.sum(np.abs(y)), np.max(np.abs(y)),
               np.linalg.norm(T, 1), np.linalg.norm(T, np.inf))
        got = (r2[qi], r1[qi], rinf[qi], n1[qi], ninf[qi])
        print(kind, qi, [abs(g - r) / r for g, r in zip(got, ref)])
        for g, r in zip(got, ref):
            assert abs(g - r) <= 1e-12 * r


def _circle_window(precision="fp64"):
    q = qb()
    b = q.make_circle(1.0)
    solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(solver, b, 20.0, N=[256])
    lam, Uk, Y, k0, R, _ = q.solve_vect(solver, None, pts, 20.0 + 0j, 0.5, nq=32, r=24,
                                        svd_tol=1e-10, rng=np.random.default_rng(1),
                                        precision=precision, **CP)
    return q, b, solver, pts, lam, Uk, Y


def test_residual_and_norm_select_circle():
    """cfg1 window: the kept eigenvalues are the J_m zeros in the disk; the
    device residuals and normalised tensions match the oracle's on the same
    eigenpairs."""
    q, b, solver, pts, lam, Uk, Y = _circle_window()
    idx, Phi, tens, tensN, logs = q.residual_and_norm_select(
        solver, lam, Uk, Y, 20.0, 0.5, pts, res_tol=1e-6, collect_logs=True, **CP)
    kept = np.sort(lam[idx].real)
    ex = O.jm_zeros_in_disk(20.0, 0.5)
    print("kept", kept, "tens", tens, "tensN", tensN)
    assert len(kept) == len(ex) and np.max(np.abs(kept - ex) / ex) < 1e-12
    comps = O.make_geometry("circle", [256])
    cp = O.ChebParams(CP["n_panels_h"], CP["M_h"], CP["n_panels_j"], CP["M_j"])
    T_of = lambda z: O.cfie_assemble(comps, [z], cp)[0]
    oidx, otens, otensN, _ = O.residual_select(T_of, lam, Uk, Y, 20.0, 0.5, 1e-6)
    assert list(oidx) == list(idx)
    # the residuals of true eigenpairs sit at the rounding level (~1e-10 for
    # ||T|| ~ 1, ||phi||_2 = 1), so they agree to an absolute tolerance
    assert np.all(np.abs(tens - otens) <= 1e-13)
    assert np.all(np.abs(tensN - otensN) <= 1e-14)
    assert len(logs) == len(idx) + 0  # every inside root was logged as kept
    # a res_tol below every residual drops everything inside
    idx0, _, _, _, _ = q.residual_and_norm_select(solver, lam, Uk, Y, 20.0, 0.5, pts,
                                                  res_tol=1e-300, **CP)
    assert len(idx0) == 0
    # without auto-discard every inside root is kept
    idx1, _, _, _, _ = q.residual_and_norm_select(solver, lam, Uk, Y, 20.0, 0.5, pts,
                                                  res_tol=1e-300, auto_discard_spurious=False,
                                                  **CP)
    assert len(idx1) == int(np.count_nonzero(np.abs(lam - 20.0) <= 0.5))


def test_imag_check_matches_full_select():
    """The |Im lambda|-ordered filter keeps the same roots as the full check
    when every candidate is examined (pad larger than the candidate count)."""
    q, b, solver, pts, lam, Uk, Y = _circle_window()
    idx_keep, resid = q.imag_k_check_EXPERIMENTAL(solver, [lam], [Uk], [Y], [20.0 + 0j], [0.5],
                                                  [pts], res_tol=1e-6, pad=1000, group_size=3,
                                                  **CP)
    idx, _, tens, _, _ = q.residual_and_norm_select(solver, lam, Uk, Y, 20.0, 0.5, pts,
                                                    res_tol=1e-6, **CP)
    assert list(idx_keep[0]) == list(idx)
    # residuals of true eigenpairs are rounding-level: absolute agreement
    assert np.all(np.abs(resid[0] - tens) <= 1e-13)


@pytest.mark.parametrize("imag_check", [False, True])
def test_compute_spectrum_beyn_circle(imag_check):
    """Weyl-planned sweep of [19.2, 20.8] on the unit circle (CFIE-Kress,
    b = 10), with the full residual check or the |Im lambda| filter: every
    kept wavenumber is a zero of J_m, and every zero of J_m in the interval
    is found with its multiplicity."""
    q = qb()
    b = q.make_circle(1.0)
    solver = q.CFIE_kress(10.0, b)
    ks, tens, us, pts, tensN = q.compute_spectrum_beyn(
        solver, b, 19.2, 20.8, m=10, Rmax=0.5, nq=32, r=40, svd_tol=1e-10, res_tol=1e-6,
        cheb_tol=1e-12, use_imag_check_EXPERIMENTAL=imag_check)
    ks = np.sort(ks)
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                    "jm_zeros.json")))
    assert g["interval"] == [19.2, 20.8]
    ex = np.sort([e["k"] for e in g["interval_zeros"] for _ in range(e["mult"])])
    print(len(ks), len(ex), ks, ex)
    assert len(ks) == len(ex)
    assert np.max(np.abs(ks - ex) / ex) < 1e-10
    assert len(us) == len(ks) and len(pts) == len(ks) and len(tensN) == len(ks)


def test_real_roots_are_those_the_residual_check_keeps():
    """On a large window (the disk with a hole, N = 6400, k0 ~ 490, 32
    nodes) A0 has directions above the absolute rank threshold that carry no
    eigenvalue; their roots have |Im k| far above 1e-8 |k|.  The reference's
    residual filter (imag_k_check_EXPERIMENTAL with every candidate checked)
    keeps exactly the real roots that the serving summaries count
    (serve.real_roots)."""
    q = qb()
    from qbhip import serve as S
    req = S.uneven_windows(128, seed=4)[9]
    b = S.SHAPES["circle_hole"]()
    solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(solver, b, req.k0, N=req.N)
    assert sum(req.N) == 6400
    zj = req.k0 + req.R * np.exp(2j * np.pi * np.arange(req.nq) / req.nq)
    nh, mh, nj, mj, _ = q.chebyshev_params(solver, pts, zj, tol=1e-12)
    cp = dict(n_panels_h=nh, M_h=mh, n_panels_j=nj, M_j=mj)
    r = q.probe_count(b, req.k0, req.R, req.nq, 1e-10)
    lam, Uk, Y, _, _, _ = q.solve_vect(solver, None, pts, req.k0 + 0j, req.R, nq=req.nq, r=r,
                                       svd_tol=1e-10, rng=np.random.default_rng(req.seed), **cp)
    ins = np.flatnonzero(np.abs(lam - req.k0) <= req.R)
    real = set(ins[np.abs(lam[ins].imag) <= S.REAL_TOL * np.abs(lam[ins])].tolist())
    print(len(ins), len(real), np.sort(np.abs(lam[ins].imag))[-5:])
    assert 0 < len(real) < len(ins)  # the window has spurious roots
    idx, resid = q.imag_k_check_EXPERIMENTAL(solver, [lam], [Uk], [Y], [req.k0 + 0j], [req.R],
                                             [pts], res_tol=1e-6, pad=10 ** 6, group_size=64,
                                             **cp)
    assert set(idx[0].tolist()) == real
    assert np.max(resid[0]) < 1e-8
