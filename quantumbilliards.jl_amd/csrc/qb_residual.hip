// Residual check of Beyn eigenpairs: y = T(lambda) phi for a batch of
// assembled matrices (one candidate eigenvalue and density per matrix),
// with the absolute column sums (opnorm(T, 1)) and row sums (opnorm(T, Inf))
// of each matrix.  Replaces the mul!(y, Tbufs[j], phi) / opnorm loop of
// residual_and_norm_select (beyn.jl:557-593) and of
// imag_k_check_EXPERIMENTAL (beyn.jl:666-687).
//
// Both kernels stream the N x N complex128 matrices once, column-major:
//   k_gemv_part: a workgroup owns 256 rows and one of `nsplit` column
//     chunks; lanes read consecutive rows of a column (coalesced), the chunk
//     of phi is staged in LDS; partial products (and partial row sums) go to
//     a [nsplit][nb][N] buffer that the host adds up, so the result does not
//     depend on launch order.
//   k_colabs: one wave per column, lanes stride the column, shuffle
//     reduction.
#include "qb_internal.h"
#include "qb_linalg.h"

__global__ void __launch_bounds__(256) k_gemv_part(const cd* __restrict__ T,
                                                   const cd* __restrict__ phi,
                                                   cd* __restrict__ part,
                                                   double* __restrict__ rowabs, int64_t N,
                                                   int64_t chunk, int nb) {
  __shared__ cd ph[256];
  const int b = blockIdx.z, s = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const cd* Tb = T + (int64_t)b * N * N;
  const cd* pb = phi + (int64_t)b * N;
  const int64_t j0 = (int64_t)s * chunk;
  const int64_t j1 = (j0 + chunk < N) ? j0 + chunk : N;
  double ar = 0, ai = 0, sa = 0;
  for (int64_t jb = j0; jb < j1; jb += 256) {
    __syncthreads();
    if (jb + threadIdx.x < j1) ph[threadIdx.x] = pb[jb + threadIdx.x];
    __syncthreads();
    const int m = (j1 - jb < 256) ? (int)(j1 - jb) : 256;
    if (i < N) {
      const cd* col = Tb + i + jb * N;
      int q = 0;
      for (; q + 4 <= m; q += 4) {
        cd t[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) t[u] = col[(int64_t)(q + u) * N];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          cd p = ph[q + u];
          ar += t[u].re * p.re - t[u].im * p.im;
          ai += t[u].re * p.im + t[u].im * p.re;
          sa += hypot(t[u].re, t[u].im);
        }
      }
      for (; q < m; ++q) {
        cd t = col[(int64_t)q * N];
        cd p = ph[q];
        ar += t.re * p.re - t.im * p.im;
        ai += t.re * p.im + t.im * p.re;
        sa += hypot(t.re, t.im);
      }
    }
  }
  if (i < N) {
    const int64_t o = ((int64_t)s * nb + b) * N + i;
    part[o] = cd{ar, ai};
    rowabs[o] = sa;
  }
}

__global__ void __launch_bounds__(256) k_colabs(const cd* __restrict__ T,
                                                double* __restrict__ colsum, int64_t N) {
  const int b = blockIdx.y;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t j = (int64_t)blockIdx.x * 4 + w;
  if (j >= N) return;  // whole wave leaves; the kernel has no barrier
  const cd* c = T + (int64_t)b * N * N + j * N;
  double s = 0;
  for (int64_t i = lane; i < N; i += 64) s += hypot(c[i].re, c[i].im);
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) colsum[(int64_t)b * N + j] = s;
}

int qb_gemv_norms(qb_ctx* ctx, const cd* T, const cd* phi, int64_t N, int nb, int nsplit,
                  cd* part, double* rowabs, double* colsum) {
  if (N <= 0 || nb <= 0) return QB_OK;
  if (nsplit < 1) nsplit = 1;
  const int64_t chunk = (N + nsplit - 1) / nsplit;
  dim3 g((unsigned)((N + 255) / 256), (unsigned)nsplit, (unsigned)nb);
  int slot = qb_prof_begin(ctx, QB_K_OTHER, 8.0 * (double)N * N * nb, 16.0 * (double)N * N * nb);
  hipLaunchKernelGGL(k_gemv_part, g, dim3(256), 0, ctx->stream, T, phi, part, rowabs, N, chunk,
                     nb);
  QB_HIP(ctx, hipGetLastError());
  dim3 g2((unsigned)((N + 3) / 4), (unsigned)nb);
  hipLaunchKernelGGL(k_colabs, g2, dim3(256), 0, ctx->stream, T, colsum, N);
  qb_prof_end(ctx, slot);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
