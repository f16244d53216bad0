"""GPU parity of the symmetry-reduced assembly (qb_geom_set_orbits) and of
the Beyn path on reduced matrices.

  * reduced T(z), and T, dT/dk, d2T/dk2, against the oracle's reduction of
    its full matrices (oracle.reduce_matrix, pinned in
    tests/test_symmetry_host.py), for CFIE with coordinate orbits and DLP with
    index orbits, reflections and rotations, one and two components;
  * Beyn on the reduced unit-disk operator: irrep m of C_4 and both parities
    of the y -> -y reflection give the zeros of J_m' they must
    (tests/golden/jm_zeros.json), in FP64 and mixed precision;
  * on the star r = 1 + 0.2 cos 5t the five C_5 irreps together give the
    eigenvalues of the unreduced window;
  * compute_spectrum_beyn with a symmetry; orbit-table errors.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def qb():
    import qbhip
    return qbhip


def _sym(kind, n, m, p):
    q = qb()
    if kind == "rot":
        return q.Rotation(n, m)
    if kind == "origin":
        return q.XYReflection(*p)
    return q.XReflection(p) if kind == "y_axis" else q.YReflection(p)


def _billiard(kind, hole=None):
    q = qb()
    if kind == "ellipse":
        return q.make_full_ellipse(1.0, 0.6)
    if kind == "star":
        return q.make_star_polar_boundary(1.0, 0.2, 5)
    if kind == "circle":
        return q.make_circle(1.0)
    return q.CircularHoleBilliard(1.0, holes=[hole])


def _setup(kind, Ns, op, sk, n, m, p, hole=None):
    q = qb()
    b = _billiard(kind, hole)
    cls = q.CFIE_kress if op == "cfie" else q.DLP_kress
    solver = cls(10.0, b, symmetry=_sym(sk, n, m, p))
    pts = q.evaluate_points(solver, b, 1.0, N=Ns)
    comps = O.make_geometry(kind, Ns, **({"hole": hole} if hole else {}))
    N = sum(Ns)
    if op == "dlp":
        img, sc = O.orbits_index(N, sk, n, m, p)
    else:
        img, sc = O.orbits_coords(comps, sk, n, m, p)
    return solver, pts, comps, img, sc


def _relerr(A, B):
    out = []
    for a, b in zip(A, B):
        s = np.max(np.abs(b))
        msk = np.abs(b) >= 1e-3 * s
        out.append((np.max(np.abs(a - b)) / s, np.max(np.abs(a - b)[msk] / np.abs(b)[msk])))
    return np.array(out)


HOLE = (0.3, 0.0, 0.0)
CASES = [
    ("ellipse", [480], "cfie", "origin", 0, 0, (-1, 1), None),
    ("ellipse", [480], "cfie", "y_axis", 0, 0, 1, None),
    ("star", [600], "cfie", "rot", 5, 2, 1, None),
    ("circle_hole", [320, 96], "cfie", "rot", 4, 1, 1, HOLE),
    ("circle_hole", [320, 96], "cfie", "x_axis", 0, 0, -1, HOLE),
    ("ellipse", [480], "dlp", "origin", 0, 0, (1, -1), None),
    ("star", [600], "dlp", "rot", 5, 4, 1, None),
    # index orbits of a rotation the ellipse does not have: the reference sums
    # them all the same; the distances do not mirror, so no pair sharing
    ("ellipse", [480], "dlp", "rot", 4, 1, 1, None),
]
# whether the orbit distances mirror (pair-shared kernel) for each case
PAIRED = [True, True, True, True, True, True, True, False]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_reduced_assembly_matches_oracle(case, capfd, monkeypatch):
    kind, Ns, op, sk, n, m, p, hole = CASES[case]
    print("case", CASES[case], flush=True)
    q = qb()
    monkeypatch.setenv("QB_SYM_TRACE", "1")
    solver, pts, comps, img, sc = _setup(kind, Ns, op, sk, n, m, p, hole)
    q.Geometry(pts, None, solver).close()
    err = capfd.readouterr().err
    assert ("pair=1" if PAIRED[case] else "pair=0") in err, err
    zj, _ = O.contour(30.0, 0.5, 4)
    cp = dict(n_panels_h=3000, M_h=5, n_panels_j=1000, M_j=5)
    T = q.assemble(solver, pts, zj, **cp)
    assert T.shape == (4, len(img), len(img))
    ocp = O.ChebParams(3000, 5, 1000, 5)
    full = O.cfie_assemble(comps, zj, ocp) if op == "cfie" else O.dlp_assemble(comps[0], zj, ocp)
    ref = np.stack([O.reduce_matrix(A, img, sc) for A in full])
    e = _relerr(T, ref)
    print(e.max(axis=0))
    assert e[:, 0].max() < 1e-12 and e[:, 1].max() < 1e-11


@pytest.mark.parametrize("case", [3, 5])
def test_reduced_derivatives_match_oracle(case):
    kind, Ns, op, sk, n, m, p, hole = CASES[case]
    q = qb()
    solver, pts, comps, img, sc = _setup(kind, Ns, op, sk, n, m, p, hole)
    zj = np.array([25.3 + 0.2j, 24.6 - 0.1j])
    cp = dict(n_panels_h=3000, M_h=7, n_panels_j=1000, M_j=7)
    got = q.assemble_with_derivatives(solver, pts, zj, **cp)
    ocp = O.ChebParams(3000, 7, 1000, 7)
    # the library's default is the reference's between-component sign
    # (cheb-CFIE:1031-1032), the same in the reduced kernels
    full = (O.cfie_assemble_deriv(comps, zj, ocp, reference_offblock_sign=True) if op == "cfie"
            else O.dlp_assemble_deriv(comps[0], zj, ocp))
    for g, f in zip(got, full):
        ref = np.stack([O.reduce_matrix(A, img, sc) for A in f])
        e = _relerr(g, ref)
        print(e.max(axis=0))
        assert e[:, 0].max() < 1e-12 and e[:, 1].max() < 1e-10
    # the value output equals the plain reduced assembly (for DLP up to the
    # |kr| < 0.2 direct Hankel branch, which the derivative path takes and the
    # plain DLP path does not, cheb-DLP:702-769 vs :881-952)
    P = q.assemble(solver, pts, zj, **cp)
    tol = 1e-14 if op == "cfie" else 1e-12
    assert np.max(np.abs(np.asarray(got[0]) - np.asarray(P))) <= tol * np.max(np.abs(P))


def _irrep_zeros(rule, key="zeros"):
    g = json.load(open(os.path.join(HERE, "golden", "jm_zeros.json")))
    out = []
    for e in g[key]:
        for mm in ({e["m"], -e["m"]} if e["m"] else {0}):
            if rule(mm):
                out.append(e["k"])
    return np.sort(out)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("sk,n,m,p", [("rot", 4, 0, 1), ("rot", 4, 1, 1), ("rot", 4, 2, 1),
                                      ("rot", 4, 3, 1), ("x_axis", 0, 0, 1),
                                      ("x_axis", 0, 0, -1)])
def test_reduced_beyn_on_the_disk(precision, sk, n, m, p):
    q = qb()
    b = q.make_circle(1.0)
    solver = q.CFIE_kress(10.0, b, symmetry=_sym(sk, n, m, p))
    pts = q.evaluate_points(solver, b, 20.0, N=[256])
    lam = q.solve(solver, None, pts, 20.0 + 0j, 0.5, nq=32, r=12, svd_tol=1e-10,
                  rng=np.random.default_rng(1), precision=precision)
    got = np.sort(lam[np.abs(lam - 20.0) <= 0.5].real)
    if sk == "rot":
        ex = _irrep_zeros(lambda mm: (mm - m) % n == 0)
    else:
        ex = _irrep_zeros(lambda mm: mm >= (0 if p == 1 else 1))
    print(sk, m, p, precision, got, ex)
    assert len(got) == len(ex)
    assert np.max(np.abs(got - ex) / ex) < 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_c5_irreps_together_give_the_full_window(precision):
    """Star r = 1 + 0.2 cos 5t, N = 400, k0 = 30, R = 0.3: the eigenvalues of
    the five reduced windows (Rotation(5, m), m = 0..4) together are those of
    the unreduced window."""
    q = qb()
    b = q.make_star_polar_boundary(1.0, 0.2, 5)
    kw = dict(nq=48, r=40, svd_tol=1e-10, precision=precision, n_panels_h=3000, M_h=7,
              n_panels_j=1000, M_j=7)
    full_solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(full_solver, b, 30.0, N=[400])
    lf = q.solve(full_solver, None, pts, 30.0 + 0j, 0.3, rng=np.random.default_rng(2), **kw)
    lf = np.sort_complex(lf[np.abs(lf - 30.0) <= 0.3])
    parts = []
    for m in range(5):
        s = q.CFIE_kress(10.0, b, symmetry=q.Rotation(5, m))
        lm = q.solve(s, None, pts, 30.0 + 0j, 0.3, rng=np.random.default_rng(2), **kw)
        parts.append(lm[np.abs(lm - 30.0) <= 0.3])
    lr = np.sort_complex(np.concatenate(parts))
    print(precision, [len(x) for x in parts], len(lf))
    assert len(lr) == len(lf) > 0
    assert np.max(np.abs(lr - lf)) < 1e-10 * 30.0


def test_compute_spectrum_beyn_with_symmetry():
    """Weyl-planned sweep of [19.2, 20.8] on the disk in irrep 1 of C_4: the
    kept wavenumbers are the zeros of J_|m'|, m' = 1 (mod 4), each once."""
    q = qb()
    b = q.make_circle(1.0)
    solver = q.CFIE_kress(10.0, b, symmetry=q.Rotation(4, 1))
    ks, tens, us, pts, tensN = q.compute_spectrum_beyn(
        solver, b, 19.2, 20.8, m=10, Rmax=0.5, nq=32, r=24, svd_tol=1e-10, res_tol=1e-6,
        cheb_tol=1e-12, use_imag_check_EXPERIMENTAL=False)
    ks = np.sort(ks)
    ex = _irrep_zeros(lambda mm: (mm - 1) % 4 == 0, key="interval_zeros")
    print(ks, ex)
    assert len(ks) == len(ex) and np.max(np.abs(ks - ex) / ex) < 1e-10
    N = len(pts[0][0])
    assert all(len(u) == N // 4 for u in us)


def test_orbit_table_errors():
    q = qb()
    ctx = q.default_context()
    L = q.lib()
    b = q.make_full_ellipse(1.0, 0.6)
    # N = 122: nodes 30 and 91 sit on the y axis; their orbit [30, 30, 91, 91]
    # repeats them and the table no longer partitions the nodes
    pts = q.evaluate_points(q.CFIE_kress(10.0, b), b, 1.0, N=[122])
    with pytest.raises(AssertionError, match="partition"):
        q.Geometry(pts, ctx, q.CFIE_kress(10.0, b, symmetry=q.XYReflection(1, 1)))
    g = q.Geometry(pts, ctx)
    img = np.arange(122, dtype=np.int64).reshape(61, 2)
    sc = np.ones((61, 2), np.complex128)
    dup = img.copy()
    dup[5, 1] = dup[5, 0]
    rc = L.qb_geom_set_orbits(ctx.ptr, g.ptr, 61, 2, dup.ctypes.data, sc.ctypes.data)
    assert rc == q._lib.QB_E_ARG and b"appears twice" in L.qb_last_error(ctx.ptr)
    sc[3, 0] = -1.0
    rc = L.qb_geom_set_orbits(ctx.ptr, g.ptr, 61, 2, img.ctypes.data, sc.ctypes.data)
    assert rc == q._lib.QB_E_ARG and b"weight 1" in L.qb_last_error(ctx.ptr)
    rc = L.qb_geom_set_orbits(ctx.ptr, g.ptr, 60, 2, img.ctypes.data, np.ones((61, 2), complex).ctypes.data)
    assert rc == q._lib.QB_E_ARG and b"partition" in L.qb_last_error(ctx.ptr)
    sc[3, 0] = 1.0
    assert L.qb_geom_set_orbits(ctx.ptr, g.ptr, 61, 2, img.ctypes.data, sc.ctypes.data) == 0
    assert L.qb_geom_dim(g.ptr) == 61 and L.qb_geom_size(g.ptr) == 122
    assert L.qb_geom_set_orbits(ctx.ptr, g.ptr, 0, 0, None, None) == 0
    assert L.qb_geom_dim(g.ptr) == 122
    g.close()


def _star_irreps(N=400):
    q = qb()
    b = q.make_star_polar_boundary(1.0, 0.2, 5)
    solvers = [q.CFIE_kress(10.0, b, symmetry=q.Rotation(5, m)) for m in range(5)]
    pts = q.evaluate_points(q.CFIE_kress(10.0, b), b, 30.0, N=[N])
    return b, solvers, pts


def test_multi_irrep_assembly_matches_single_irreps():
    """Assembling the five C_5 irreps together (one evaluation of every image
    term, qb_geom_set_orbits_multi) gives each irrep's reduced matrices."""
    q = qb()
    b, solvers, pts = _star_irreps(600)
    zj = np.array([30.2 + 0.1j, 29.7 - 0.2j, 30.4 + 0.0j])
    cp = dict(n_panels_h=3000, M_h=7, n_panels_j=1000, M_j=7)
    g = q.Geometry(pts, None, irreps=solvers)
    assert g.nirr == 5 and q.lib().qb_geom_irreps(g.ptr) == 5
    T = q.assemble(solvers[0], pts, zj, geom=g, **cp)
    assert T.shape == (15, 120, 120)
    for m, s in enumerate(solvers):
        ref = q.assemble(s, pts, zj, **cp)
        got = T[m * 3:(m + 1) * 3]
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(m, err)
        assert err < 1e-13
    # one irrep only for the derivative assembly and the residual check
    with pytest.raises(AssertionError, match="one irrep"):
        q.assemble_with_derivatives(solvers[0], pts, zj, geom=g, **cp)
    with pytest.raises(AssertionError, match="one irrep"):
        q.residual_norms(solvers[0], pts, zj[:1], np.zeros((120, 1), complex), geom=g, **cp)
    g.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_multi_irrep_window_matches_single_irrep_windows(precision):
    """One window over all five C_5 irreps (one batched LU of 5 x 48
    matrices, moments and QR per irrep) gives the eigenvalues and bases of
    the five single-irrep windows."""
    q = qb()
    b, solvers, pts = _star_irreps(400)
    cp = dict(n_panels_h=3000, M_h=7, n_panels_j=1000, M_j=7)
    g = q.Geometry(pts, None, irreps=solvers)
    V = q.probe_block(np.random.default_rng(7), 80, 40)
    w = q.Window(solvers[0], pts, 30.0 + 0j, 0.3, 48, precision=precision, geom=g, **cp)
    assert w.nirr == 5
    R, Cm = w.moments_qr(V, 1e-13)
    assert R.shape == (5, 40, 40)
    for m, s in enumerate(solvers):
        B, Ur, S, rk = q.B_from_qr(R[m], Cm[m], 1e-10)
        lam = np.sort_complex(np.linalg.eigvals(B))
        w1 = q.Window(s, pts, 30.0 + 0j, 0.3, 48, precision=precision, **cp)
        R1, C1 = w1.moments_qr(V, 1e-13)
        B1, Ur1, S1, rk1 = q.B_from_qr(R1, C1, 1e-10)
        lam1 = np.sort_complex(np.linalg.eigvals(B1))
        # leading directions only: those near the 1e-10 rank threshold are
        # ill-conditioned and turn under rounding-level changes of A0
        kl = int(np.count_nonzero(S1 >= 1e-6 * S1[0]))
        U1 = w1.basis(Ur1[:, :kl])
        w1.close()
        assert rk == rk1 and np.max(np.abs(S - S1)) <= 1e-12 * S1[0], m
        inside = np.abs(lam1 - 30.0) <= 0.3
        assert np.max(np.abs(lam[inside] - lam1[inside])) < 1e-10 * 30, m
        U = w.basis(Ur[:, :kl], irrep=m)
        P, P1 = U @ U.conj().T, U1 @ U1.conj().T
        assert np.max(np.abs(P - P1)) < 1e-9, m
    w.close()
    g.close()
