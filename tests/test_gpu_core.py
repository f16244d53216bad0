"""GPU parity of the hot path against the oracle (oracle/oracle.py)."""
import numpy as np
import pytest
import scipy.special as sp

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def qb():
    import qbhip
    return qbhip


def _pts(kind, Ns, **kw):
    q = qb()
    if kind == "circle":
        b = q.make_circle(kw.get("R", 1.0))
    elif kind == "ellipse":
        b = q.make_full_ellipse(1.0, 0.6)
    elif kind == "star":
        b = q.make_star_polar_boundary(1.0, 0.2, 5)
    elif kind == "circle_hole":
        b = q.CircularHoleBilliard(1.0, holes=[(0.3, 0.35, 0.0)])
    return b, q.evaluate_points(q.CFIE_kress(10.0, b), b, 1.0, N=Ns)


def _relerr(A, B):
    """max |A-B| / max |B| per matrix, and entrywise relative error over
    entries at least 1e-3 of the matrix maximum."""
    out = []
    for a, b in zip(A, B):
        s = np.max(np.abs(b))
        m = np.abs(b) >= 1e-3 * s
        out.append((np.max(np.abs(a - b)) / s, np.max(np.abs(a - b)[m] / np.abs(b)[m])))
    return np.array(out)


def test_bessel_device_matches_amos():
    rng = np.random.default_rng(3)
    rho = np.exp(rng.uniform(np.log(1e-3), np.log(2500), 20000))
    im = rng.uniform(-1.2, 1.2, rho.size)
    re = np.sqrt(np.maximum(rho ** 2 - im ** 2, 1e-6))
    z = re + 1j * im
    h0, h1, j0, j1 = qb().bessel_eval(z)
    for a, b in ((h0, sp.hankel1(0, z)), (h1, sp.hankel1(1, z))):
        assert np.max(np.abs(a - b) / np.abs(b)) < 2e-14
    for a, b in ((j0, sp.jv(0, z)), (j1, sp.jv(1, z))):
        scale = np.maximum(np.abs(b), np.minimum(1.0, np.exp(np.abs(z.imag)) / np.sqrt(np.abs(z))))
        assert np.max(np.abs(a - b) / scale) < 1e-14


@pytest.mark.parametrize("kind,Ns,k0", [("circle", [256], 20.0), ("star", [512], 60.0),
                                         ("circle_hole", [200, 60], 25.0)])
def test_cfie_assembly_matches_oracle(kind, Ns, k0):
    q = qb()
    b, pts = _pts(kind, Ns)
    zj, _ = O.contour(k0, 0.5, 8)
    cp = dict(n_panels_h=4000, M_h=5, n_panels_j=1500, M_j=5)
    T = q.assemble(q.CFIE_kress(10.0, b), pts, zj, **cp)
    comps = O.make_geometry(kind, Ns)
    Tref = O.cfie_assemble(comps, zj, O.ChebParams(4000, 5, 1500, 5))
    e = _relerr(T, Tref)
    print(kind, e.max(axis=0))
    assert e[:, 0].max() < 1e-12 and e[:, 1].max() < 1e-11


def test_cfie_stadium_assembly_matches_oracle():
    """The stadium (a C1 composite of five segments; its nodes are pinned by
    closed forms in tests/test_geometry_host.py) assembled on the device and
    by the oracle from the same nodes."""
    q = qb()
    b = q.make_full_stadium(1.0, 1.0)
    pts = q.evaluate_points(q.CFIE_kress(10.0, b), b, 1.0, N=[400])
    p = pts[0]
    comps = [O.Comp(p.xy[:, 0].copy(), p.xy[:, 1].copy(), p.tangent[:, 0].copy(),
                    p.tangent[:, 1].copy(), p.tangent_2[:, 0].copy(), p.tangent_2[:, 1].copy(),
                    p.ts.copy(), p.ws.copy())]
    zj, _ = O.contour(15.0, 0.3, 8)
    cp = dict(n_panels_h=4000, M_h=5, n_panels_j=1500, M_j=5)
    T = q.assemble(q.CFIE_kress(10.0, b), pts, zj, **cp)
    Tref = O.cfie_assemble(comps, zj, O.ChebParams(4000, 5, 1500, 5))
    e = _relerr(T, Tref)
    print("stadium", e.max(axis=0))
    assert e[:, 0].max() < 1e-12 and e[:, 1].max() < 1e-11


@pytest.mark.parametrize("kind,Ns,op", [("circle_hole", [200, 60], "cfie"),
                                        ("star", [1000], "dlp"), ("star", [4000], "cfie")])
def test_table_range_matches_oracle(kind, Ns, op):
    """The padded distance range of the tables that qb_assemble reports
    (rminmax) against the oracle's: 0.95 x the smallest and 1.05 x the
    largest pairwise distance above eps, i != j, over all component blocks
    (CFIE) or over the first component (DLP). Distances agree to the last
    bit of hypot."""
    import ctypes as C
    from qbhip import _lib
    from qbhip.solvers import _cheb
    q = qb()
    b, pts = _pts(kind, Ns)
    solver = (q.CFIE_kress if op == "cfie" else q.DLP_kress)(10.0, b)
    ctx = q.default_context()
    g = q.Geometry(pts, ctx, solver)
    zj = np.array([20.0 + 0.1j])
    raw = np.empty((1, g.dim, g.dim), np.complex128)
    rr = np.zeros(2)
    cp = _cheb(400, 5, 200, 5)
    rc = _lib.lib().qb_assemble(ctx.ptr, g.ptr, solver.op, zj.ctypes.data, 1, C.byref(cp),
                                raw.ctypes.data, 0, rr.ctypes.data)
    _lib.check(ctx.ptr, rc)
    g.close()
    comps = O.make_geometry(kind, Ns)
    ref = np.array(O._rrange_blocks(comps if op == "cfie" else comps[:1]))
    assert np.max(np.abs(rr - ref) / ref) < 1e-15, (rr, ref)


def test_dlp_assembly_matches_oracle():
    q = qb()
    b = q.make_full_ellipse(1.0, 0.6)
    pts = q.evaluate_points(q.DLP_kress(10.0, b), b, 1.0, N=[384])
    zj, _ = O.contour(40.0, 0.5, 6)
    T = q.assemble(q.DLP_kress(10.0, b), pts, zj, n_panels_h=3000, M_h=5, n_panels_j=1000, M_j=5)
    Tref = O.dlp_assemble(O.make_geometry("ellipse", [384])[0], zj, O.ChebParams(3000, 5, 1000, 5))
    e = _relerr(T, Tref)
    print(e.max(axis=0))
    assert e[:, 0].max() < 1e-12 and e[:, 1].max() < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_int8res"])
def test_window_moments_match_host_solve(precision):
    """LU + solves + moments on the device against LAPACK on the same T (the
    window's assembly: conjugate pairs from the envelope tables; mixed:
    refinement residual as a complex128 GEMM; mixed_int8res: residual on the
    int8 matrix cores), and every node's full-column backward error at exit
    below the refinement tolerance."""
    q = qb()
    b, pts = _pts("star", [320])
    solver = q.CFIE_kress(10.0, b)
    k0, R, nq = 30.0 + 0.0j, 0.4, 6
    zj, wj = O.contour(k0, R, nq)
    cp = dict(n_panels_h=3000, M_h=5, n_panels_j=1000, M_j=5)
    T = q.assemble(solver, pts, zj, conj_pairs=True, **cp)
    V = q.probe_block(np.random.default_rng(5), 320, 16)
    A0r, A1r = O.beyn_moments(np.array(T), zj, wj, V)
    win = q.Window(solver, pts, k0, R, nq, precision=precision, **cp)
    A0, A1 = win.moments(V, 1e-13)
    print(win.stats)
    for a, r in ((A0, A0r), (A1, A1r)):
        assert np.max(np.abs(a - r)) / np.max(np.abs(r)) < (1e-11 if precision == "fp64" else 1e-10)
    if precision != "fp64":
        assert 0.0 < win.stats["refine_resid_max"] <= 1e-13
    win.close()


def test_refinement_falls_back_to_fp64_factors():
    """A refinement tolerance no FP64 residual can certify (1e-20) sends every
    node of a mixed-precision window to the FP64 fallback: the nodes are
    refactored in FP64 as one batch and solved there, so the moments are the
    FP64 window's and the host solve's."""
    q = qb()
    b, pts = _pts("star", [320])
    solver = q.CFIE_kress(10.0, b)
    k0, R, nq = 30.0 + 0.0j, 0.4, 6
    zj, wj = O.contour(k0, R, nq)
    cp = dict(n_panels_h=3000, M_h=5, n_panels_j=1000, M_j=5)
    T = q.assemble(solver, pts, zj, conj_pairs=True, **cp)
    V = q.probe_block(np.random.default_rng(5), 320, 16)
    A0r, A1r = O.beyn_moments(np.array(T), zj, wj, V)
    w64 = q.Window(solver, pts, k0, R, nq, precision="fp64", **cp)
    B0, B1 = w64.moments(V)
    w64.close()
    win = q.Window(solver, pts, k0, R, nq, precision="mixed", **cp)
    A0, A1 = win.moments(V, 1e-20)
    st = win.stats
    win.close()
    print(st)
    assert st["nodes_fp64_fallback"] == nq
    for a, f, r in ((A0, B0, A0r), (A1, B1, A1r)):
        assert np.max(np.abs(a - f)) / np.max(np.abs(f)) < 1e-13
        assert np.max(np.abs(a - r)) / np.max(np.abs(r)) < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_beyn_circle_gives_bessel_zeros(precision):
    """cfg1: unit circle, N=256, k0=20, R=0.5, nq=32, 24 probes, svd_tol 1e-10."""
    q = qb()
    b = q.make_circle(1.0)
    solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(solver, b, 20.0, N=[256])
    lam = q.solve(solver, None, pts, 20.0 + 0j, 0.5, nq=32, r=24, svd_tol=1e-10,
                  rng=np.random.default_rng(1), precision=precision)
    inside = np.sort(lam[np.abs(lam - 20.0) <= 0.5].real)
    ex = O.jm_zeros_in_disk(20.0, 0.5)
    print(precision, inside - ex if len(inside) == len(ex) else (inside, ex))
    assert len(inside) == len(ex)
    assert np.max(np.abs(inside - ex) / ex) < 1e-12


def test_chebyshev_params_tables_meet_tol_against_amos():
    """The picks of the device tuner (qb_cheb_tune,
    chebyshev_optimal_panelization.jl:221-262), started from small tables so
    that it has to grow them, checked independently: the oracle builds the
    four tables at the picked (np, M) and evaluates them at random radii (not
    the tuner's uniform grid) against scipy's AMOS H0, H1, J0, J1. The
    absolute error stays within 10 tol; a table built for another z or range
    would miss by orders of magnitude."""
    q = qb()
    b, pts = _pts("star", [300])
    solver = q.CFIE_kress(10.0, b)
    zj, _ = O.contour(30.0, 0.5, 8)
    tol = 1e-10
    nh, mh, nj, mj, errs = q.chebyshev_params(solver, pts, zj, tol=tol, npanels_h_init=200,
                                              M_h_init=5, npanels_j_init=100, M_j_init=5)
    print("picks", (nh, mh, nj, mj), errs)
    assert max(errs) <= tol and (nh, mh, nj, mj) != (200, 5, 100, 5)
    rmin, rmax = O._rrange_blocks(O.make_geometry("star", [300]))
    # H tables start at max(min_j 0.2/|z_j|, rmin) (optimal_panelization.jl:222)
    rmin = max(rmin, float(np.min(0.2 / np.abs(zj))))
    rng = np.random.default_rng(5)
    rh, rj = rng.uniform(rmin, rmax, 20000), rng.uniform(0.0, rmax, 20000)
    ph, pj = O.Panels(rmin, rmax, nh, mh), O.Panels(0.0, rmax, nj, mj)
    worst = 0.0
    for z in zj:
        for kind, nu, P, r in (("h", 0, ph, rh), ("h", 1, ph, rh), ("j", 0, pj, rj),
                               ("j", 1, pj, rj)):
            c = O.build_table(kind, nu, z, P.rmin, P.rmax, P.npan, P.M)
            pidx, t = P.locate(r)
            f = O.clenshaw(c[pidx], t)
            d = sp.hankel1(nu, z * r) if kind == "h" else sp.jv(nu, z * r)
            worst = max(worst, float(np.max(np.abs(f - d))))
    assert worst <= 10 * tol, worst


def test_dlp_ellipse_full_size_sampled_columns():
    """Config 2 at full size: DLP-Kress on the ellipse (1, 0.6), N = 4096,
    z = 400 + 0.5 e^{i theta} (8 of the 64 contour nodes), tables tuned to
    1e-12 on the device; sampled columns against the oracle."""
    q = qb()
    b = q.make_full_ellipse(1.0, 0.6)
    solver = q.DLP_kress(10.0, b)
    pts = q.evaluate_points(solver, b, 400.0, N=[4096])
    zall, _ = O.contour(400.0, 0.5, 64)
    nh, mh, nj, mj, errs = q.chebyshev_params(solver, pts, zall, tol=1e-12)
    assert max(errs) < 1e-12
    zs = zall[::8]
    T = q.assemble(solver, pts, zs, n_panels_h=nh, M_h=mh, n_panels_j=nj, M_j=mj)
    comp = O.make_geometry("ellipse", [4096])[0]
    rmin, rmax = O._rrange_single_fast(comp)
    cp = O.ChebParams(nh, mh, nj, mj)
    rng = np.random.default_rng(0)
    worst = 0.0
    for qi, z in enumerate(zs):
        j0 = int(rng.integers(0, 4096 - 32))
        F = O.dlp_assemble_cols(comp, z, cp, j0, j0 + 32, rmin, rmax)
        G = np.asarray(T[qi][:, j0:j0 + 32])
        s = np.max(np.abs(F))
        worst = max(worst, np.max(np.abs(G - F)) / s)
    print("cfg2 worst normwise", worst, (nh, mh, nj, mj))
    assert worst < 1e-12


def test_device_matches_golden_fixtures():
    """Device Bessel values and the device-assembled unit-circle operators
    against the mpmath fixtures of tests/golden (values and Fourier
    multipliers)."""
    import json
    import os
    q = qb()
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    c = lambda v: complex(v[0], v[1])
    rows = json.load(open(os.path.join(gold, "bessel_mpmath.json")))
    z = np.array([c(r["z"]) for r in rows])
    got = q.bessel_eval(z)
    for k, g in zip(("h0", "h1", "j0", "j1"), got):
        ref = np.array([c(r[k]) for r in rows])
        assert np.max(np.abs(g - ref) / np.maximum(np.abs(ref), 1e-3)) < 3e-14, k
    mult = json.load(open(os.path.join(gold, "circle_multipliers.json")))
    b = q.make_circle(1.0)
    N = 128
    m_all = np.fft.fftfreq(N, 1.0 / N).astype(int)
    cp = dict(n_panels_h=6000, M_h=7, n_panels_j=3000, M_j=7)
    for key, ent in mult.items():
        zz = c(ent["z"])
        T = q.assemble(q.CFIE_kress(10.0, b), q.evaluate_points(q.CFIE_kress(10.0, b), b, 1.0,
                                                                 N=[N]), [zz], **cp)[0]
        F = q.assemble(q.DLP_kress(10.0, b), q.evaluate_points(q.DLP_kress(10.0, b), b, 1.0,
                                                                N=[N]), [zz], **cp)[0]
        lt, lf = np.fft.fft(np.asarray(T[:, 0])), np.fft.fft(np.asarray(F[:, 0]))
        for e in ent["multipliers"]:
            i = int(np.flatnonzero(m_all == e["m"])[0])
            assert abs(lt[i] - c(e["cfie"])) < 1e-10, (key, e["m"])
            assert abs(lf[i] - c(e["dlp"])) < 1e-10, (key, e["m"])


@pytest.mark.parametrize("kind,Ns,k0,refsign", [("circle_hole", [200, 60], 25.0, False),
                                                ("circle_hole", [200, 60], 25.0, True),
                                                ("star", [384], 40.0, False)])
def test_cfie_derivative_assembly_matches_oracle(kind, Ns, k0, refsign):
    """construct_boundary_matrices_with_derivatives!: T, dT/dk and d2T/dk2 at
    complex wavenumbers against the oracle (same tables), with the
    between-component sign of the derivative of the matrix and of the
    reference."""
    q = qb()
    b, pts = _pts(kind, Ns)
    zj = np.array([k0 + 0.3 + 0.2j, k0 - 0.4 - 0.1j])
    cp = dict(n_panels_h=4000, M_h=7, n_panels_j=1500, M_j=7)
    A, A1, A2 = q.assemble_with_derivatives(q.CFIE_kress(10.0, b), pts, zj,
                                            reference_offblock_sign=refsign, **cp)
    R = O.cfie_assemble_deriv(O.make_geometry(kind, Ns), zj, O.ChebParams(4000, 7, 1500, 7),
                              reference_offblock_sign=refsign)
    for got, ref in zip((A, A1, A2), R):
        e = _relerr(got, ref)
        print(kind, refsign, e.max(axis=0))
        assert e[:, 0].max() < 1e-12 and e[:, 1].max() < 1e-10
    # the undifferentiated output is the plain assembly
    P = q.assemble(q.CFIE_kress(10.0, b), pts, zj, **cp)
    assert np.max(np.abs(np.asarray(A) - np.asarray(P))) <= 1e-14 * np.max(np.abs(P))


def test_dlp_derivative_assembly_matches_oracle():
    q = qb()
    b = q.make_full_ellipse(1.0, 0.6)
    pts = q.evaluate_points(q.DLP_kress(10.0, b), b, 1.0, N=[320])
    zj = np.array([35.0 + 0.25j, 34.6 - 0.3j])
    Tb = [np.zeros((320, 320), complex, order="F") for _ in range(2)]
    dT = [np.zeros((320, 320), complex, order="F") for _ in range(2)]
    ddT = [np.zeros((320, 320), complex, order="F") for _ in range(2)]
    q.construct_boundary_matrices_with_derivatives(Tb, dT, ddT, q.DLP_kress(10.0, b), pts, zj,
                                                   n_panels_h=3000, M_h=7, n_panels_j=1000,
                                                   M_j=7)
    R = O.dlp_assemble_deriv(O.make_geometry("ellipse", [320])[0], zj,
                             O.ChebParams(3000, 7, 1000, 7))
    for got, ref in zip((np.array(Tb), np.array(dT), np.array(ddT)), R):
        e = _relerr(got, ref)
        print(e.max(axis=0))
        assert e[:, 0].max() < 1e-12 and e[:, 1].max() < 1e-10


def test_lu_graph_replay_matches_direct_launches():
    """With hipGraph replay on, the first window of a shape captures the
    batched LU and later windows replay it; every replay (with eager work in
    between) gives bitwise the moments of direct launches, FP64 and mixed."""
    q = qb()
    b, pts = _pts("star", [320])
    solver = q.CFIE_kress(10.0, b)
    cp = dict(n_panels_h=3000, M_h=5, n_panels_j=1000, M_j=5)
    V = q.probe_block(np.random.default_rng(5), 320, 16)
    ctx = q.Context(0)
    geom = q.Geometry(pts, ctx)
    precs = ("fp64", "mixed")
    ref = {}
    for p in precs:
        w = q.Window(solver, pts, 30.0 + 0j, 0.4, 6, precision=p, geom=geom, **cp)
        ref[p] = w.moments(V, 1e-13)
        w.close()
    ctx.set_graphs(True)
    for rep in range(3):
        for p in precs:
            w = q.Window(solver, pts, 30.0 + 0j, 0.4, 6, precision=p, geom=geom, **cp)
            A0, A1 = w.moments(V, 1e-13)
            w.close()
            assert np.array_equal(A0, ref[p][0]) and np.array_equal(A1, ref[p][1]), (rep, p)
    ctx.set_graphs(False)
    geom.close()
    ctx.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_moments_qr_on_device(precision):
    """qb_window_moments_qr: A0 = Q R with Q orthonormal and C = Q^H A1, against
    the moments of qb_window_moments with the same probes; the device-QR route
    of construct_B_matrix gives the eigenvalues and eigenvector span of the
    host-SVD route."""
    q = qb()
    b, pts = _pts("star", [320])
    solver = q.CFIE_kress(10.0, b)
    cp = dict(n_panels_h=3000, M_h=5, n_panels_j=1000, M_j=5)
    V = q.probe_block(np.random.default_rng(5), 320, 16)
    win = q.Window(solver, pts, 30.0 + 0j, 0.4, 8, precision=precision, **cp)
    A0, A1 = win.moments(V, 1e-13)
    Rm, Cm = win.moments_qr(V, 1e-13)
    Q = win.basis(np.eye(16, dtype=complex))
    win.close()
    s0 = np.max(np.abs(A0))
    assert np.max(np.abs(Q.conj().T @ Q - np.eye(16))) < 1e-13
    assert np.max(np.abs(Q @ Rm - A0)) < 1e-12 * s0
    assert np.max(np.abs(np.tril(Rm, -1))) == 0.0
    assert np.max(np.abs(Q.conj().T @ A1 - Cm)) < 1e-12 * np.max(np.abs(A1))
    # the two routes of construct_B_matrix on the unit-circle window
    bc = q.make_circle(1.0)
    sc = q.CFIE_kress(10.0, bc)
    pc = q.evaluate_points(sc, bc, 20.0, N=[256])
    out = []
    for dq in (True, False):
        B, Uk, info = q.construct_B_matrix(sc, pc, 256, 20.0 + 0j, 0.5, nq=32, r=24,
                                           svd_tol=1e-10, rng=np.random.default_rng(1),
                                           precision=precision, device_qr=dq, return_info=True)
        lam = np.linalg.eigvals(B)
        out.append((np.sort_complex(lam[np.abs(lam - 20.0) <= 0.5]), Uk, info["sigma"]))
    (l1, U1, S1), (l2, U2, S2) = out
    # inside the disk the eigenvalues agree (outside, leaked poles are set by
    # singular values near the threshold and are not compared)
    assert len(l1) == len(l2) == 10 and np.max(np.abs(l1 - l2) / np.abs(l2)) < 1e-12
    assert np.max(np.abs(S1 - S2)) <= 1e-13 * S2[0]
    # the span of the well-separated singular directions agrees
    lead = int(np.count_nonzero(S2 >= 1e-6 * S2[0]))
    P1 = U1[:, :lead] @ U1[:, :lead].conj().T
    P2 = U2[:, :lead] @ U2[:, :lead].conj().T
    assert np.max(np.abs(U1.conj().T @ U1 - np.eye(U1.shape[1]))) < 1e-12
    assert np.max(np.abs(P1 - P2)) < 1e-8


@pytest.mark.parametrize("r", [16, 40, 96])
def test_moments_qr_blocked_matches_lapack(r):
    """The device QR of the moment block at widths that take the unblocked
    sweep (r <= 32) and the compact-WY sweep (panels of 32 reflectors, the
    last one partial): R equals LAPACK's zgeqrf R of the same A0 (same
    Householder sign convention, beta = -sign(Re alpha) |x|) to rounding,
    Q is orthonormal, Q R = A0 and C = Q^H A1."""
    q = qb()
    b, pts = _pts("star", [320])
    solver = q.CFIE_kress(10.0, b)
    cp = dict(n_panels_h=3000, M_h=5, n_panels_j=1000, M_j=5)
    V = q.probe_block(np.random.default_rng(7), 320, r)
    win = q.Window(solver, pts, 30.0 + 0j, 0.4, 8, precision="fp64", **cp)
    A0, A1 = win.moments(V, 1e-13)
    Rm, Cm = win.moments_qr(V, 1e-13)
    Q = win.basis(np.eye(r, dtype=complex))
    win.close()
    s0 = np.max(np.abs(A0))
    R_np = np.linalg.qr(A0, mode="r")
    assert np.max(np.abs(np.tril(Rm, -1))) == 0.0
    assert np.all(np.abs(np.diag(Rm).imag) == 0.0)
    assert np.max(np.abs(Rm - R_np)) < 1e-12 * s0
    assert np.max(np.abs(Q.conj().T @ Q - np.eye(r))) < 1e-13
    assert np.max(np.abs(Q @ Rm - A0)) < 1e-12 * s0
    assert np.max(np.abs(Q.conj().T @ A1 - Cm)) < 1e-12 * np.max(np.abs(A1))


@pytest.mark.parametrize("kind,Ns,k0,op,M,nq", [
    ("star", [512], 60.0, "cfie", 5, 8),
    ("star", [512], 60.0, "cfie", 7, 8),
    ("circle_hole", [200, 60], 25.0, "cfie", 5, 8),
    ("ellipse", [384], 40.0, "dlp", 7, 6),
    ("star", [512], 10.0, "cfie", 5, 8),
])
def test_conj_pair_assembly_matches_oracle(kind, Ns, k0, op, M, nq):
    """k_assemble_env, the kernel that assembles the T(z_j) of every Beyn
    window with real k0 (QB_ASM_CONJ_PAIRS routes qb_assemble through it),
    against the oracle, which reads H and J from the reference's uniform
    tables (cheb-CFIE:543-665, cheb-DLP:702-769), normwise to 1e-12 and
    entrywise to 1e-11, and entry by entry against the exact-Hankel route
    (AMOS at every distance) to 1e-12:
      * each pair (z, conj z) of contour nodes is assembled together from the
        two nodes' envelope tables of H_n e^{-izr} (quarter-octave Chebyshev
        panels in LDS) and takes J_n(z r) = (H_n(z r) + conj H_n(conj(z) r)) / 2;
        the real nodes theta = 0, pi pair with themselves (J = Re H);
      * the mirror entries (j, i) of both nodes go through LDS;
      * the two-component circle with a hole exercises the off-component
        block (no J there) and the DLP ellipse the DLP entry;
      * at k0 = 10 the nearest neighbours of the star lie below |k r| = 0.2,
        where both H and the derived J come from the direct evaluation.
    The uniform-table variant of the pair kernel (k_assemble_conj,
    QB_ASM_TABLES=1) passes the same checks."""
    q = qb()
    if op == "dlp":
        b = q.make_full_ellipse(1.0, 0.6)
        solver = q.DLP_kress(10.0, b)
        pts = q.evaluate_points(solver, b, 1.0, N=Ns)
    else:
        b, pts = _pts(kind, Ns)
        solver = q.CFIE_kress(10.0, b)
    zj, _ = O.contour(k0, 0.5, nq)
    assert np.sum(np.abs(zj.imag) < 1e-14) == 2  # two self-paired real nodes
    nh = 4000 if op == "cfie" else 3000
    cp = dict(n_panels_h=nh, M_h=M, n_panels_j=1500, M_j=M)
    T = q.assemble(solver, pts, zj, conj_pairs=True, **cp)
    ocp = O.ChebParams(nh, M, 1500, M)
    comps = O.make_geometry(kind, Ns)
    if op == "dlp":
        Tref = O.dlp_assemble(comps[0], zj, ocp)
    else:
        Tref = O.cfie_assemble(comps, zj, ocp)
    if k0 == 10.0:
        c = comps[0]
        r = np.hypot(c.x - np.roll(c.x, 1), c.y - np.roll(c.y, 1))
        assert np.min(r) * np.min(np.abs(zj)) < 0.2  # the direct branch is reached
    # the envelope tables are closer to the exact Hankels than the uniform
    # tables of degree M: against the oracle's table route the difference is
    # that route's own error (4e-11 at M = 5, k0 = 60), plus 1e-12
    Tex = O.dlp_assemble_exact(comps[0], zj) if op == "dlp" else O.cfie_assemble(comps, zj,
                                                                                  exact=True)
    ex = _relerr(T, Tex)
    print("vs exact Hankels", ex.max(axis=0))
    assert ex[:, 0].max() < 1e-12 and ex[:, 1].max() < 1e-12
    own = _relerr(Tref, Tex).max(axis=0)
    e = _relerr(T, Tref)
    print(kind, op, M, k0, e.max(axis=0), "oracle tables vs exact", own)
    assert e[:, 0].max() < own[0] + 1e-12 and e[:, 1].max() < own[1] + 1e-11
    # the J-table kernel (k_assemble) gives the same matrices to the tables' tolerance
    P = q.assemble(solver, pts, zj, **cp)
    assert _relerr(T, P)[:, 0].max() < own[0] + 1e-12
    # the uniform-table variant of the pair kernel
    import os
    os.environ["QB_ASM_TABLES"] = "1"
    try:
        Tt = q.assemble(solver, pts, zj, conj_pairs=True, **cp)
    finally:
        del os.environ["QB_ASM_TABLES"]
    et = _relerr(Tt, Tref)
    assert et[:, 0].max() < 1e-12 and et[:, 1].max() < 1e-11


@pytest.mark.parametrize("op", ["cfie", "dlp"])
def test_envelope_assembly_variants_are_bitwise_equal(op):
    """The envelope-table kernel's two variants (mirror tile staged 16 columns
    at a time at two workgroups per CU, or 4 columns at three) compute the
    same entries with the same arithmetic: QB_ASM_ENV_OCC=2 and =3 give
    bitwise the same matrices, on the star (CFIE) and the ellipse (DLP), with
    a partial last tile (N not a multiple of 64)."""
    import os
    q = qb()
    if op == "dlp":
        b = q.make_full_ellipse(1.0, 0.6)
        solver = q.DLP_kress(10.0, b)
        pts = q.evaluate_points(solver, b, 1.0, N=[330])
    else:
        b, pts = _pts("star", [330])
        solver = q.CFIE_kress(10.0, b)
    zj, _ = O.contour(40.0, 0.5, 6)
    cp = dict(n_panels_h=3000, M_h=7, n_panels_j=1000, M_j=7)
    out = {}
    try:
        for occ in ("2", "3"):
            os.environ["QB_ASM_ENV_OCC"] = occ
            out[occ] = q.assemble(solver, pts, zj, conj_pairs=True, **cp)
    finally:
        del os.environ["QB_ASM_ENV_OCC"]
    assert np.array_equal(out["2"], out["3"])
    assert np.array_equal(q.assemble(solver, pts, zj, conj_pairs=True, **cp), out["2"])


def test_conj_pair_assembly_rejects_unpaired_nodes():
    """QB_ASM_CONJ_PAIRS with a node whose conjugate is missing is an argument
    error (the library does not fall back silently)."""
    from qbhip._lib import QBArgError
    q = qb()
    b, pts = _pts("star", [128])
    with pytest.raises(QBArgError):
        q.assemble(q.CFIE_kress(10.0, b), pts, [20.0 + 0.3j, 20.0 - 0.2j], conj_pairs=True,
                   n_panels_h=2000, M_h=5, n_panels_j=800, M_j=5)
