"""GPU tests at the full workloads of the star window (N = 8192, 48 nodes),
the uneven windows on the disk with a hole and the serving stream.

Each test prints what it runs before the GPU work starts, so that a failing
run names its case."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def qb():
    import qbhip
    return qbhip


def _match(a, b, tol):
    """Greedy one-to-one match of two root sets; returns (matched, worst
    relative difference)."""
    a, b = np.asarray(a), np.asarray(b)
    used = np.zeros(len(b), bool)
    n, worst = 0, 0.0
    for x in a:
        if len(b) == 0:
            break
        d = np.abs(b - x) / np.abs(x)
        d[used] = np.inf
        i = int(np.argmin(d))
        if d[i] <= tol:
            used[i] = True
            n += 1
            worst = max(worst, float(d[i]))
    return n, worst


def test_star_N8192_window_full_size():
    """The star r = 1 + 0.2 cos 5t, CFIE-Kress, N = 8192, 48 contour nodes
    800 + 0.5 e^{i theta}, tables tuned on the device to 1e-12, probes from
    the Weyl count of the disk (768: 64 probes saturate the rank of the ~400
    levels inside), mixed precision against the complex128 path.
      (a) Sampled columns of the mixed window's own matrices (the
          conjugate-pair kernel with envelope tables: a conjugate pair of
          nodes and both real nodes) against the oracle with AMOS tables of
          the picked sizes (the reference's route) and against the
          exact-Hankel route: <= 1e-12 normwise.
      (b) The refinement ends with every node's full-column backward error
          <= refine_tol = 1e-12 (refine_resid_max).
      (c) Inside the disk the mixed eigenvalues match the complex128 path's
          one to one within 1e-10 relative.
      (d) The reference's residual filter (imag_k_check_EXPERIMENTAL, every
          candidate examined, beyn.jl:623-721) keeps exactly the real roots
          |Im k| <= 1e-8 |k| that the bench counts."""
    q = qb()
    from qbhip import serve as S
    N, nq, k0, R, svd_tol = 8192, 48, 800.0, 0.5, 1e-10
    print(f"star N={N} nq={nq} k0={k0}", flush=True)
    b = q.make_star_polar_boundary(1.0, 0.2, 5)
    solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(solver, b, k0, N=[N])
    ctx = q.Context(0)
    geom = q.Geometry(pts, ctx)
    zj, _ = O.contour(k0, R, nq)
    nh, mh, nj, mj, errs = q.chebyshev_params(solver, pts, zj, tol=1e-12, geom=geom)
    assert max(errs) < 1e-12
    cheb = dict(n_panels_h=nh, M_h=mh, n_panels_j=nj, M_j=mj)
    r = q.probe_count(b, k0, R, nq, svd_tol)
    assert r == 768
    win = q.Window(solver, pts, k0, R, nq, precision="mixed", geom=geom, **cheb)
    # (a) the window's matrices against the oracle
    comp = O.make_geometry("star", [N])[0]
    rmin, rmax = O._rrange_single_fast(comp)
    assert abs(win.stats["rmin"] - rmin) <= 1e-14 * rmin
    assert abs(win.stats["rmax"] - rmax) <= 1e-14 * rmax
    ocp = O.ChebParams(nh, mh, nj, mj)
    rng = np.random.default_rng(3)
    worst = 0.0
    assert abs(zj[0].imag) < 1e-14 and abs(zj[24].imag) < 1e-14
    assert abs(zj[5] - np.conj(zj[43])) < 1e-13
    for node in (0, 5, 43, 24, 17):
        j0 = int(rng.integers(0, N - 32))
        G = win.matrix_cols(node, j0, 32)
        F = O.cfie_cols_c(comp, complex(zj[node]), ocp, j0, j0 + 32, rmin, rmax)
        assert np.all(np.isfinite(F))  # no |k r| < 0.2 pair at this size
        Fx = O.cfie_assemble_cols(comp, complex(zj[node]), ocp, j0, j0 + 32, rmin, rmax,
                                  exact=True)
        e = np.max(np.abs(G - F)) / np.max(np.abs(F))
        ex = np.max(np.abs(G - Fx)) / np.max(np.abs(Fx))
        print("node", node, "cols", j0, "normwise", e, "vs exact", ex, flush=True)
        worst = max(worst, e, ex)
    assert worst < 1e-12
    # (b) + (c) the mixed window against the complex128 window
    B, Uk, info = q.construct_B_matrix(solver, pts, N, complex(k0), R, nq=nq, r=r,
                                       svd_tol=svd_tol, rng=np.random.default_rng(1),
                                       precision="mixed", refine_tol=1e-12, window=win,
                                       return_info=True)
    win.close()
    st = info["moments"]
    print("mixed", st, "rk", info["rk"], flush=True)
    assert info["rk"] < r and info["r"] == r
    assert 0.0 < st["refine_resid_max"] <= 1e-12
    lam, Y = np.linalg.eig(B)
    w64 = q.Window(solver, pts, k0, R, nq, precision="fp64", geom=geom, **cheb)
    B64, _ = q.construct_B_matrix(solver, pts, N, complex(k0), R, nq=nq, r=r, svd_tol=svd_tol,
                                  rng=np.random.default_rng(1), precision="fp64", window=w64)
    w64.close()
    lam64 = np.linalg.eigvals(B64)
    ins = np.flatnonzero(np.abs(lam - k0) <= R)
    real = ins[np.abs(lam[ins].imag) <= S.REAL_TOL * np.abs(lam[ins])]
    real64 = lam64[(np.abs(lam64 - k0) <= R) & (np.abs(lam64.imag) <= S.REAL_TOL * np.abs(lam64))]
    n, wv = _match(lam[real], real64, 1e-10)
    print("real roots", len(real), len(real64), "matched", n, "worst", wv, flush=True)
    assert len(real) > 350 and n == len(real) == len(real64)
    # (d) the residual filter over every candidate inside the disk
    idx, resid = q.imag_k_check_EXPERIMENTAL(solver, [lam], [Uk], [Y], [complex(k0)], [R], [pts],
                                             res_tol=1e-6, pad=10 ** 6, group_size=64,
                                             geoms=[geom], **cheb)
    print("filter keeps", len(idx[0]), "of", len(ins), "max resid", np.max(resid[0]), flush=True)
    assert set(idx[0].tolist()) == set(real.tolist())
    geom.close()
    ctx.close()


UNEVEN = (116, 109, 9, 14)  # N = 1536, 4096, 6400, 10496 of uneven_windows(128, seed=4)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_uneven_windows_alone_equal_in_mix(precision):
    """Windows of the uneven workload on the disk with a hole of radius 0.3 at
    (0.35, 0) (two boundary components, the off-component blocks), N from
    1536 to 10496, 32 nodes each: run alone one after another, and all
    submitted at once to the serving loop (two slots, admission by HBM
    footprint), every window returns the same roots.  The real roots of the
    largest window (N = 10496) are those the reference's residual filter
    keeps (imag_k_check_EXPERIMENTAL, every candidate examined) on the same
    disk."""
    q = qb()
    from qbhip import serve as S
    allreq = S.uneven_windows(128, seed=4, precision=precision)
    reqs = [allreq[i] for i in UNEVEN]
    print(precision, [(rq.rid, round(rq.k0, 2), sum(rq.N)) for rq in reqs], flush=True)
    assert [sum(rq.N) for rq in reqs] == [1536, 4096, 6400, 10496]
    srv = S.Server(slots=1)
    alone = {rq.rid: srv.submit(rq).result() for rq in reqs}
    srv.close()
    srv = S.Server(slots=2)
    mixed = {f.result().rid: f.result() for f in [srv.submit(rq) for rq in reqs]}
    srv.close()
    for rid in alone:
        a, m = alone[rid], mixed[rid]
        print(rid, a.N, len(a.real), len(a.inside), a.stats.get("precision_used"), flush=True)
        assert not a.stats["rank_saturated"] and not m.stats["rank_saturated"]
        assert len(a.real) > 0 and len(a.real) == len(m.real), rid
        assert _match(a.real, m.real, 1e-10)[0] == len(a.real), rid
    # the residual filter on the largest disk (complex128 path, same probes)
    big = allreq[UNEVEN[-1]]
    b = S.SHAPES["circle_hole"]()
    solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(solver, b, big.k0, N=big.N)
    zj = big.k0 + big.R * np.exp(2j * np.pi * np.arange(big.nq) / big.nq)
    nh, mh, nj, mj, _ = q.chebyshev_params(solver, pts, zj, tol=1e-12)
    cp = dict(n_panels_h=nh, M_h=mh, n_panels_j=nj, M_j=mj)
    lam, Uk, Y, _, _, _ = q.solve_vect(solver, None, pts, big.k0 + 0j, big.R, nq=big.nq,
                                       r=alone[big.rid].r, svd_tol=big.svd_tol,
                                       rng=np.random.default_rng(big.seed), **cp)
    idx, _ = q.imag_k_check_EXPERIMENTAL(solver, [lam], [Uk], [Y], [big.k0 + 0j], [big.R],
                                         [pts], res_tol=1e-6, pad=10 ** 6, group_size=64, **cp)
    kept = lam[idx[0]]
    n, wv = _match(alone[big.rid].real, kept, 1e-10)
    print("largest window: kept", len(kept), "real", len(alone[big.rid].real), "matched", n, wv,
          flush=True)
    assert n == len(kept) == len(alone[big.rid].real)


def test_serving_stream_alone_equals_in_mix():
    """A seeded Poisson stream of the serving workload: ellipse (1, 0.6),
    stadium (radius 1, half-length 1) and the star, k0 spanning [50, 1000]
    (N from 10 points per wavelength: 422 ... 11 688), R in [0.25, 0.5],
    tolerances 1e-8 and 1e-12, mixed precision.  Served with two slots as the
    requests arrive, each window returns the roots it returns when it runs
    alone; the summary reports p50/p99 latency."""
    from qbhip import serve as S
    rng = np.random.default_rng(11)
    spec = [("ellipse", 52.0), ("stadium", 110.0), ("star", 240.0), ("ellipse", 470.0),
            ("stadium", 640.0), ("star", 960.0)]
    t, reqs = 0.0, []
    for i, (shape, k0) in enumerate(spec):
        t += rng.exponential(0.25)
        reqs.append(S.WindowRequest(i, shape, k0, float(rng.uniform(0.25, 0.5)),
                                    tol=(1e-8, 1e-12)[i % 2], nq=32, seed=300 + i, arrival=t,
                                    precision="mixed"))
    print([(rq.shape, rq.k0, round(rq.R, 3), rq.tol) for rq in reqs], flush=True)
    srv = S.Server(slots=1)
    alone = {rq.rid: srv.submit(rq).result() for rq in reqs}
    srv.close()
    srv = S.Server(slots=2)
    res = srv.run_stream(reqs)
    srv.close()
    summ = S.summarize(res)
    print(summ, flush=True)
    assert summ["windows"] == len(reqs) and summ["rejected"] == 0
    assert summ["p99_latency_s"] >= summ["p50_latency_s"] > 0
    Ns = [r.N for r in res]
    assert min(Ns) < 500 and max(Ns) > 11000
    for m in res:
        a = alone[m.rid]
        print(m.rid, m.N, len(m.real), len(m.inside), flush=True)
        assert not m.stats["rank_saturated"] and m.stats["cheb_tol_met"]
        assert len(a.real) > 0 and len(a.real) == len(m.real) and len(a.inside) == len(m.inside)
        assert _match(a.real, m.real, 1e-10)[0] == len(a.real), m.rid


def test_serving_grows_an_undersized_probe_block():
    """A request that starts with too few probes (r = 4 for the ten J_m zeros
    of the unit disk at k0 = 20, R = 0.5) saturates the rank of A0; the server
    solves the resident factors again with twice the probes until the rank no
    longer saturates (beyn.jl:291-309) and returns the zeros of J_m."""
    from qbhip import serve as S
    req = S.WindowRequest(0, "circle", 20.0, 0.5, tol=1e-12, nq=32, N=[256], r=4, seed=5)
    srv = S.Server(slots=1)
    res = srv.submit(req).result()
    srv.close()
    print("r", res.r, "rk", res.rk, res.stats["rank_saturated"], flush=True)
    assert res.r >= 16 and res.rk < res.r and not res.stats["rank_saturated"]
    ex = O.jm_zeros_in_disk(20.0, 0.5)
    got = np.sort(res.inside.real)
    assert len(got) == len(ex) and np.max(np.abs(got - ex) / ex) < 1e-10
