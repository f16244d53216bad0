"""Serving loop: continuous batching with HBM-footprint admission.  Each
window's result must be the same when it runs alone and when it runs in a
mix with others (same probes, same kernels)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _reqs():
    from qbhip.serve import WindowRequest
    return [
        WindowRequest(0, "circle", 20.0, 0.5, tol=1e-12, nq=24, ppw=13, seed=11, precision="mixed"),
        WindowRequest(1, "ellipse", 24.0, 0.4, tol=1e-8, nq=24, ppw=12, seed=12, precision="mixed"),
        WindowRequest(2, "star", 18.0, 0.45, tol=1e-12, nq=24, ppw=12, seed=13, precision="fp64"),
        WindowRequest(3, "stadium", 15.0, 0.3, tol=1e-8, nq=24, ppw=12, seed=14,
                      precision="mixed"),
        WindowRequest(4, "circle_hole", 16.0, 0.35, tol=1e-12, nq=24, ppw=12, seed=15,
                      precision="mixed"),
        WindowRequest(5, "circle", 26.0, 0.25, tol=1e-12, nq=24, ppw=12, seed=16),
    ]


def test_alone_equals_mixed_and_circle_is_exact():
    from qbhip.serve import Server, summarize
    reqs = _reqs()
    srv = Server(slots=1)
    alone = {q.rid: srv.submit(q).result() for q in reqs}
    srv.close()
    srv = Server(slots=3)
    futs = [srv.submit(q) for q in reqs]
    mixed = {f.result().rid: f.result() for f in futs}
    srv.close()
    for rid in alone:
        a, m = alone[rid].inside, mixed[rid].inside
        assert len(a) == len(m) and len(a) > 0, rid
        assert np.max(np.abs(a - m) / np.abs(a)) <= 1e-10, rid
    # the circle windows are checked against the zeros of J_m
    for rid, k0, R in ((0, 20.0, 0.5), (5, 26.0, 0.25)):
        ex = O.jm_zeros_in_disk(k0, R)
        got = np.sort(mixed[rid].inside.real)
        assert len(got) == len(ex)
        assert np.max(np.abs(got - ex) / ex) < 1e-10
    s = summarize(list(mixed.values()))
    assert s["eigenvalues"] > 0 and s["p99_latency_s"] >= s["p50_latency_s"]


def test_admission_by_footprint():
    from qbhip.beyn import footprint
    from qbhip.serve import Server, WindowRequest
    reqs = _reqs()[:3]
    srv0 = Server(slots=1)
    fps = [srv0._prepare((q, None))[4] for q in reqs]
    srv0.close()
    # budget for the largest one only: windows must run one at a time
    srv = Server(slots=3, hbm_budget=max(fps))
    res = [f.result() for f in [srv.submit(q) for q in reqs]]
    srv.close()
    spans = sorted((r.t_admit, r.t_admit + r.device_s) for r in res)
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1 - 1e-3, "admitted while the budget was full"
    # a request whose footprint exceeds the whole budget is rejected
    srv = Server(slots=1, hbm_budget=min(fps) // 2)
    with pytest.raises(MemoryError):
        srv.submit(reqs[0]).result()
    srv.close()


def test_mixed_request_falls_back_to_fp64_when_only_that_fits():
    """An int8-residual window whose footprint exceeds the whole budget runs
    with the complex128-GEMM residual (no int8 residues of T: 24 instead of
    56 bytes per matrix entry) when that fits, else in FP64 (16 bytes) when
    that fits, with the same eigenvalues; the per-request outcome of a stream
    includes the rejection of a window that fits in none."""
    import dataclasses
    from qbhip.beyn import footprint
    from qbhip.serve import Server, WindowRequest, summarize
    req = dataclasses.replace(_reqs()[0], precision="mixed_int8res")
    srv0 = Server(slots=1)
    b, Ns, N, r, fp_int8, prec = srv0._prepare((req, None))
    ref = srv0.submit(req).result()
    srv0.close()
    assert prec == "mixed_int8res"
    fp64 = footprint(N, req.nq, "fp64", r)
    fpres = footprint(N, req.nq, "mixed", r)
    assert fp64 < fpres < fp_int8
    srv = Server(slots=1, hbm_budget=(fpres + fp_int8) // 2)
    got = srv.submit(req).result()
    srv.close()
    assert got.stats["precision_used"] == "mixed"
    assert len(got.inside) == len(ref.inside)
    assert np.max(np.abs(got.inside - ref.inside) / np.abs(ref.inside)) <= 1e-10
    srv = Server(slots=1, hbm_budget=(fp64 + fpres) // 2)
    got = srv.submit(req).result()
    assert got.stats["precision_used"] == "fp64"
    assert len(got.inside) == len(ref.inside)
    assert np.max(np.abs(got.inside - ref.inside) / np.abs(ref.inside)) <= 1e-10
    big = WindowRequest(9, "circle", 60.0, 0.5, nq=24, ppw=40)
    out = srv.run_stream([req, big])
    srv.close()
    assert isinstance(out[1], MemoryError)
    s = summarize(out)
    assert s["windows"] == 1 and s["rejected"] == 1


@pytest.mark.parametrize("batched", [True, False])
def test_symmetric_request_equals_the_unreduced_one(batched):
    """A request served as the reduced windows of its shape's symmetry group
    (both reflections of the ellipse and stadium, the C_5 rotations of the
    star), all irreps in one window or one job per irrep, returns the
    eigenvalues the unreduced window finds in the disk; on the disk they are
    the zeros of J_m."""
    from qbhip.serve import Server, WindowRequest
    reqs = [WindowRequest(0, "ellipse", 24.0, 0.4, tol=1e-12, nq=32, ppw=12, seed=21),
            WindowRequest(1, "star", 30.0, 0.3, tol=1e-12, nq=32, ppw=12, seed=22),
            WindowRequest(2, "stadium", 15.0, 0.3, tol=1e-12, nq=32, ppw=12, seed=23),
            WindowRequest(3, "circle", 20.0, 0.5, tol=1e-12, nq=32, ppw=13, seed=24)]
    srv = Server(slots=2, irreps_batched=batched)
    full, red = {}, {}
    for q in reqs:
        # the same node counts (a multiple of 4 / 10) for both
        sym = WindowRequest(**{**q.__dict__, "symmetry": True})
        f = srv.submit(sym)
        rs = f.result()
        Ns = srv._prepare((sym, __import__("qbhip").serve.SYM_GROUPS[q.shape]()[0]))[1]
        full[q.rid] = srv.submit(WindowRequest(**{**q.__dict__, "N": list(Ns)})).result()
        red[q.rid] = rs
    srv.close()
    for rid in full:
        a, m = full[rid].inside, red[rid].inside
        print(rid, len(a), len(m), red[rid].irreps)
        assert len(a) == len(m) and len(a) > 0, rid
        a, m = np.sort_complex(a), np.sort_complex(m)
        assert np.max(np.abs(a - m) / np.abs(a)) <= 1e-10, rid
        assert red[rid].irreps == (5 if reqs[rid].shape == "star" else 4)
    ex = O.jm_zeros_in_disk(20.0, 0.5)
    got = np.sort(red[3].inside.real)
    assert len(got) == len(ex) and np.max(np.abs(got - ex) / ex) < 1e-10


def test_requests_sharing_a_geometry_are_batched():
    """Requests with the same shape, node counts and settings that wait for a
    slot run as one batched window (qb_window_create_multi); each returns the
    eigenvalues of its own disk, as when it runs alone, and on the disk the
    zeros of J_m."""
    from qbhip.serve import Server, WindowRequest
    reqs = [WindowRequest(i, "circle", k0, 0.5, tol=1e-12, nq=24, N=[256], seed=31 + i)
            for i, k0 in enumerate((19.0, 20.0, 21.3))]
    reqs += [WindowRequest(3 + i, "circle_hole", k0, 0.35, tol=1e-12, nq=24, N=[200, 60],
                           seed=41 + i) for i, k0 in enumerate((16.0, 17.1))]
    srv = Server(slots=1, group=False)
    alone = {q.rid: srv.submit(q).result() for q in reqs}
    srv.close()
    srv = Server(slots=1, group=True)
    # occupy the only slot so that the requests below wait together
    blocker = srv.submit(WindowRequest(-1, "ellipse", 30.0, 0.4, nq=32, ppw=12))
    futs = [srv.submit(q) for q in reqs]
    grouped = {f.result().rid: f.result() for f in futs}
    blocker.result()
    srv.close()
    for rid, res in grouped.items():
        assert res.stats["group"] == (3 if rid < 3 else 2), rid
        assert alone[rid].stats["group"] == 1
        a, m = alone[rid].inside, res.inside
        assert len(a) == len(m) and len(a) > 0, rid
        assert np.max(np.abs(a - m) / np.abs(a)) <= 1e-10, rid
    for rid, k0 in enumerate((19.0, 20.0, 21.3)):
        ex = O.jm_zeros_in_disk(k0, 0.5)
        got = np.sort(grouped[rid].inside.real)
        assert len(got) == len(ex) and np.max(np.abs(got - ex) / ex) < 1e-10, rid
