"""Error behaviour and empty inputs through the C ABI: argument errors come
back as QB_E_ARG (raised as an AssertionError, like the reference's size
checks, cheb-CFIE:1433 / cheb-DLP:1117) with the context's message, and
empty batches are valid no-ops."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def q():
    import qbhip
    return qbhip


def _circle(N=64):
    b = q().make_circle(1.0)
    s = q().CFIE_kress(10.0, b)
    return b, s, q().evaluate_points(s, b, 10.0, N=[N])


CP = dict(n_panels_h=500, M_h=5, n_panels_j=200, M_j=5)


def test_argument_errors_raise_assertion_errors():
    Q = q()
    b, s, pts = _circle()
    with pytest.raises(AssertionError, match="Chebyshev parameters"):
        Q.assemble(s, pts, [10.0 + 0.1j], n_panels_h=500, M_h=1, n_panels_j=200, M_j=5)
    with pytest.raises(AssertionError, match="nq"):
        Q.Window(s, pts, 10.0 + 0j, 0.3, 0, **CP)
    with pytest.raises(AssertionError, match="radius"):
        Q.Window(s, pts, 10.0 + 0j, -0.3, 8, **CP)
    w = Q.Window(s, pts, 10.0 + 0j, 0.3, 8, **CP)
    with pytest.raises(AssertionError, match="r must be"):
        w.moments(np.zeros((64, 0), complex))
    w.close()
    # DLP-Kress on two components
    hb = Q.CircularHoleBilliard(1.0, holes=[(0.3, 0.35, 0.0)])
    hp = Q.evaluate_points(Q.CFIE_kress(10.0, hb), hb, 10.0, N=[64, 20])
    with pytest.raises(AssertionError, match="one boundary component"):
        Q.assemble(Q.DLP_kress(10.0, hb), hp, [10.0 + 0j], **CP)
    # Tbufs of the wrong size (host check, as in the reference)
    with pytest.raises(AssertionError):
        Q.construct_boundary_matrices([np.zeros((10, 10), complex)], s, pts, [10.0 + 0j], **CP)
    with pytest.raises(AssertionError):
        Q.residual_norms(s, pts, [10.0 + 0j], np.zeros((63, 1), complex), **CP)
    # use_chebyshev=false raises (the reference: error for DLP, zeros + @error for CFIE)
    with pytest.raises(RuntimeError):
        Q.construct_boundary_matrices([np.zeros((64, 64), complex)], s, pts, [10.0 + 0j],
                                      use_chebyshev=False)


def test_abi_rejects_null_pointers_with_a_message():
    Q = q()
    L = Q.lib()
    ctx = Q.default_context()
    b, s, pts = _circle()
    g = Q.Geometry(pts, ctx)
    cp = Q._lib.ChebParams(500, 5, 200, 5)
    rc = L.qb_assemble(ctx.ptr, g.ptr, 1, None, 1, C.byref(cp), None, 0, None)
    assert rc == Q._lib.QB_E_ARG
    assert b"null" in L.qb_last_error(ctx.ptr)
    rc = L.qb_assemble(ctx.ptr, g.ptr, 7, None, 0, C.byref(cp), None, 0, None)
    assert rc == Q._lib.QB_E_ARG
    g.close()


def test_empty_batches_are_no_ops():
    Q = q()
    b, s, pts = _circle()
    T = Q.assemble(s, pts, np.zeros(0, complex), **CP)
    assert T.shape == (0, 64, 64)
    out = Q.residual_norms(s, pts, np.zeros(0, complex), np.zeros((64, 0), complex), **CP)
    assert all(len(o) == 0 for o in out)
    idx, Phi, tens, tensN, logs = Q.residual_and_norm_select(
        s, np.zeros(0, complex), np.zeros((64, 0), complex), np.zeros((0, 0), complex),
        10.0 + 0j, 0.3, pts, res_tol=1e-8, **CP)
    assert len(idx) == 0 and Phi.shape == (64, 0)
    ks, tens, us, pk, tn = Q.compute_spectrum_beyn(s, b, 10.0, 10.0, m=5)
    assert len(ks) == 0 and len(us) == 0


def test_pinned_probe_block_gives_the_same_moments():
    """qb_host_alloc memory (qbhip.pinned_array) as the probe block: the
    moments equal those from ordinary host memory bit for bit."""
    Q = q()
    b, s, pts = _circle(96)
    rng = np.random.default_rng(4)
    V = Q.probe_block(rng, 96, 8)
    P = Q.pinned_array((96, 8))
    P[...] = V
    assert P.flags.f_contiguous
    w = Q.Window(s, pts, 10.0 + 0j, 0.3, 8, **CP)
    A0, A1 = w.moments(V)
    B0, B1 = w.moments(P)
    w.close()
    assert np.array_equal(A0, B0) and np.array_equal(A1, B1)
    Vp = Q.probe_block(np.random.default_rng(4), 96, 8, pinned=True)
    assert np.array_equal(Vp, V)
