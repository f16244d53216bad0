"""qb_assemble_noalloc, the assembly entry point that allocates no device
memory (for callers that cannot allocate), against qb_assemble on the same
inputs: two calls in a row, the second with fewer wavenumbers and a smaller
geometry, each bitwise equal to qb_assemble's matrices; a call whose tables
exceed the fixed device array is rejected with QB_E_ARG and a clear message."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libqbhip links against, the copy already mapped into
    this process (device buffers for the test)."""
    from qbhip import _lib
    _lib.lib()
    path = "libamdhip64.so.7"
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    L = C.CDLL(path)
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return L


H2D, D2H = 1, 2


def _dev(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
    return p


def _noalloc(q, solver, pts, zj, cp, hip):
    from qbhip import _lib
    from qbhip.solvers import Geometry, _cheb
    ctx = q.default_context()
    g = Geometry(pts, ctx, solver)
    zj = np.ascontiguousarray(np.asarray(zj, np.complex128))
    nz, N = len(zj), g.dim
    zd = _dev(hip, zj.nbytes)
    od = _dev(hip, nz * N * N * 16)
    try:
        assert hip.hipMemcpy(zd, zj.ctypes.data, zj.nbytes, H2D) == 0
        rr = np.zeros(2)
        rc = _lib.lib().qb_assemble_noalloc(ctx.ptr, g.ptr, solver.op, zd, nz,
                                            C.byref(_cheb(**cp)), od, 0, rr.ctypes.data)
        _lib.check(ctx.ptr, rc)
        raw = np.empty((nz, N, N), np.complex128)
        assert hip.hipMemcpy(raw.ctypes.data, od, raw.nbytes, D2H) == 0
    finally:
        hip.hipFree(zd)
        hip.hipFree(od)
        g.close()
    return raw.transpose(0, 2, 1), rr


def test_noalloc_assembly_equals_qb_assemble_twice():
    import qbhip as q
    from oracle import oracle as O
    hip = _hip()
    b = q.make_star_polar_boundary(1.0, 0.2, 5)
    solver = q.CFIE_kress(10.0, b)
    cases = [([640], O.contour(45.0, 0.5, 8)[0], dict(n_panels_h=6000, M_h=7, n_panels_j=2000,
                                                       M_j=7)),
             ([384], O.contour(30.0, 0.4, 3)[0], dict(n_panels_h=3000, M_h=5, n_panels_j=1000,
                                                       M_j=5))]
    for Ns, zj, cp in cases:
        print("noalloc star N", Ns, "nz", len(zj), cp, flush=True)
        pts = q.evaluate_points(solver, b, 1.0, N=Ns)
        got, rr = _noalloc(q, solver, pts, zj, cp, hip)
        ref = q.assemble(solver, pts, zj, **cp)
        assert np.array_equal(got, ref)
        assert rr[0] > 0 and rr[1] > rr[0]


def test_noalloc_rejects_tables_beyond_the_array():
    import qbhip as q
    from qbhip._lib import QBArgError
    from oracle import oracle as O
    hip = _hip()
    b = q.make_star_polar_boundary(1.0, 0.2, 5)
    solver = q.CFIE_kress(10.0, b)
    pts = q.evaluate_points(solver, b, 1.0, N=[128])
    zj = O.contour(20.0, 0.5, 64)[0]
    # 64 * (200000 * 8 + 1000 * 8) * 32 bytes > 128 MiB
    cp = dict(n_panels_h=200000, M_h=7, n_panels_j=1000, M_j=7)
    with pytest.raises(QBArgError, match="fixed device array"):
        _noalloc(q, solver, pts, zj, cp, hip)
