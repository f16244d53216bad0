"""Launches for hardware-counter runs (rocprofv3 --pmc): a plain copy kernel
moving a known number of bytes, then the complex64 and complex128 GEMMs at
the LU trailing-update shape (M = N = 7936, K = 256, 4 nodes) and the
complex128 GEMM at the refinement-residual shape (8192 x 768, K = 8192).
With the argument `window`: one mixed-precision Beyn window at the bench
workload instead (star r = 1 + 0.2 cos 5t, CFIE-Kress, N = 8192, k0 = 800,
R = 0.5, 48 nodes, 768 probes, tables tuned to 1e-12, GEMM tiles autotuned),
preceded by a 1 MiB k_copy16 that marks where the window's dispatches start.
Used by the profiling commands recorded under profiles/."""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(HERE, "..", ".."), os.path.join(HERE, "..", "..", "quantumbilliards.jl_amd")]

import qbhip  # noqa: E402

L = C.CDLL(os.path.join(HERE, "libqbt_linalg.so"))
L.qbt_gemm_bench.restype = C.c_double
L.qbt_copy_bench.restype = C.c_double
ctx = qbhip.default_context()
p = C.c_void_p(ctx.ptr.value)
if len(sys.argv) > 1 and sys.argv[1] == "window":
    import numpy as np
    N, nq, k0, R, r = 8192, 48, 800.0, 0.5, 768
    b = qbhip.make_star_polar_boundary(1.0, 0.2, 5)
    solver = qbhip.CFIE_kress(10.0, b)
    pts = qbhip.evaluate_points(solver, b, k0, N=[N])
    geom = qbhip.Geometry(pts, ctx)
    zj = k0 + R * np.exp(2j * np.pi * np.arange(nq) / nq)
    nh, mh, nj, mj, _ = qbhip.chebyshev_params(solver, pts, zj, tol=1e-12, geom=geom)
    ctx.autotune(N, nq, r, "mixed")
    print("copy 1 GiB: %.0f GB/s" % L.qbt_copy_bench(p, C.c_int64(1 << 30), 1), flush=True)
    L.qbt_copy_bench(p, C.c_int64(1 << 20), 1)
    win = qbhip.Window(solver, pts, k0, R, nq, precision="mixed", geom=geom, n_panels_h=nh,
                       M_h=mh, n_panels_j=nj, M_j=mj)
    A0, A1 = win.moments(qbhip.probe_block(np.random.default_rng(1), N, r), 1e-12)
    print("window stats", win.stats, flush=True)
    win.close()
    sys.exit(0)
print("copy 1 GiB: %.0f GB/s" % L.qbt_copy_bench(p, C.c_int64(1 << 30), 1), flush=True)
print("c64  LU trailing 7936x7936x256 x4: %.1f TF" % L.qbt_gemm_bench(p, 7936, 7936, 256, 4, 1, 0), flush=True)
print("c128 LU trailing 7936x7936x256 x4: %.1f TF" % L.qbt_gemm_bench(p, 7936, 7936, 256, 4, 1, 1), flush=True)
print("c128 residual 8192x768x8192 x4: %.1f TF" % L.qbt_gemm_bench(p, 8192, 768, 8192, 4, 1, 1), flush=True)
