"""Launches for hardware-counter runs (rocprofv3 --pmc): a plain copy kernel
moving a known number of bytes, then the complex64 and complex128 GEMMs at
the LU trailing-update shape (M = N = 7936, K = 256, 4 nodes) and the
complex128 GEMM at the refinement-residual shape (8192 x 768, K = 8192).
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
print("copy 1 GiB: %.0f GB/s" % L.qbt_copy_bench(p, C.c_int64(1 << 30), 1), flush=True)
print("c64  LU trailing 7936x7936x256 x4: %.1f TF" % L.qbt_gemm_bench(p, 7936, 7936, 256, 4, 1, 0), flush=True)
print("c128 LU trailing 7936x7936x256 x4: %.1f TF" % L.qbt_gemm_bench(p, 7936, 7936, 256, 4, 1, 1), flush=True)
print("c128 residual 8192x768x8192 x4: %.1f TF" % L.qbt_gemm_bench(p, 8192, 768, 8192, 4, 1, 1), flush=True)
