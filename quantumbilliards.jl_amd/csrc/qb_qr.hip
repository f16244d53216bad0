// Householder QR of the Beyn moment block on the device.
//
// The reference takes the thin SVD of the N x r block A0 on the host
// (svd(A0), beyn.jl:286) and forms B = U_k^H A1 W_k Sigma_k^{-1}
// (beyn.jl:310-321).  With A0 = Q R (Householder, backward stable also for
// the rank-deficient A0 of a saturated window), svd(R) = U_R Sigma W^H gives
// the same singular values and W, U_k = Q U_R[:, :k], and
// U_k^H A1 = U_R[:, :k]^H (Q^H A1)[:r, :].  So only the r x r blocks R and
// C = (Q^H A1)[:r, :] leave the device, and U_k is formed on the device when
// it is asked for.
//
// Storage: one N x (2r) column-major block [A0 | A1] (ld = N).  Column j of
// the A0 half is reduced by H_j = I - tau_j v_j v_j^H (v_j[j] = 1, v_j below
// the diagonal in place, LAPACK zgeqr2 / zlarfg conventions), applied as
// H_j^H to the columns right of j (the rest of A0 and all of A1).
#include "qb_internal.h"
#include "qb_linalg.h"

// zlarfg for column j: rows j..N-1.  One workgroup per block; block b =
// blockIdx.y holds its matrix at A + b * bstride and its tau at tau + b * tstride.
__global__ void __launch_bounds__(1024) k_house(cd* __restrict__ A, int64_t N, int64_t ld, int j,
                                                cd* __restrict__ tau, int64_t bstride,
                                                int64_t tstride) {
  __shared__ double red[16];
  __shared__ cd s_scale;
  A += (int64_t)blockIdx.y * bstride;
  tau += (int64_t)blockIdx.y * tstride;
  cd* x = A + (int64_t)j * ld;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double s = 0;
  for (int64_t i = j + 1 + tid; i < N; i += blockDim.x) s += x[i].re * x[i].re + x[i].im * x[i].im;
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) red[wv] = s;
  __syncthreads();
  if (tid == 0) {
    double sig2 = 0;
    for (int q = 0; q < (int)(blockDim.x + 63) / 64; ++q) sig2 += red[q];
    const cd alpha = x[j];
    if (sig2 == 0.0 && alpha.im == 0.0) {
      tau[j] = mk(0.0, 0.0);
      s_scale = mk(0.0, 0.0);  // no reflection: the column stays as it is
    } else {
      const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + sig2);
      const double beta = alpha.re >= 0.0 ? -nrm : nrm;
      tau[j] = mk((beta - alpha.re) / beta, -alpha.im / beta);
      s_scale = crecip(mk(alpha.re - beta, alpha.im));
      x[j] = mk(beta, 0.0);
    }
  }
  __syncthreads();
  const cd sc = s_scale;
  if (sc.re == 0.0 && sc.im == 0.0) return;
  for (int64_t i = j + 1 + tid; i < N; i += blockDim.x) x[i] = x[i] * sc;
}

// Columns c0 + blockIdx.x (< c1) of A: a <- a - t v (v^H a) over rows j..N-1,
// with v the stored reflector of column j (v[j] = 1) of V and t = conj(tau_j)
// (ADJ: applying H_j^H, the QR sweep) or tau_j (applying H_j, forming Q).
// Block b = blockIdx.y: V, A at + b * bstride, tau at + b * tstride.
template <bool ADJ>
__global__ void __launch_bounds__(256) k_house_apply(const cd* __restrict__ V, int64_t ldv,
                                                     cd* __restrict__ A, int64_t N, int64_t lda,
                                                     int j, int c0, const cd* __restrict__ tau,
                                                     int64_t bstride, int64_t tstride) {
  __shared__ double red_re[4], red_im[4];
  V += (int64_t)blockIdx.y * bstride;
  A += (int64_t)blockIdx.y * bstride;
  tau += (int64_t)blockIdx.y * tstride;
  const cd t = ADJ ? cconj(tau[j]) : tau[j];
  if (t.re == 0.0 && t.im == 0.0) return;  // uniform over the workgroup
  const cd* v = V + (int64_t)j * ldv;
  cd* a = A + (int64_t)(c0 + blockIdx.x) * lda;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // w = v^H a
  double wr = 0, wi = 0;
  for (int64_t i = j + tid; i < N; i += 256) {
    const cd vi = (i == j) ? mk(1.0, 0.0) : v[i];
    const cd ai = a[i];
    wr += vi.re * ai.re + vi.im * ai.im;
    wi += vi.re * ai.im - vi.im * ai.re;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    wr += __shfl_xor(wr, off);
    wi += __shfl_xor(wi, off);
  }
  if (lane == 0) {
    red_re[wv] = wr;
    red_im[wv] = wi;
  }
  __syncthreads();
  const cd w = mk(red_re[0] + red_re[1] + red_re[2] + red_re[3],
                  red_im[0] + red_im[1] + red_im[2] + red_im[3]);
  const cd tw = t * w;
  for (int64_t i = j + tid; i < N; i += 256) {
    const cd vi = (i == j) ? mk(1.0, 0.0) : v[i];
    a[i] = a[i] - vi * tw;
  }
}

int qb_house_qr(qb_ctx* ctx, cd* A, int64_t N, int r, int ncols, cd* tau, int nb,
                int64_t bstride, int64_t tstride) {
  if (r > N || ncols < r) return qb_fail(ctx, QB_E_ARG, "QR needs r <= N and ncols >= r");
  if (nb < 1 || nb > 65535) return qb_fail(ctx, QB_E_ARG, "QR batch must be in [1, 65535]");
  for (int j = 0; j < r; ++j) {
    hipLaunchKernelGGL(k_house, dim3(1, nb), dim3(1024), 0, ctx->stream, A, N, N, j, tau,
                       bstride, tstride);
    if (ncols - j - 1 > 0)
      hipLaunchKernelGGL(k_house_apply<true>, dim3(ncols - j - 1, nb), dim3(256), 0,
                         ctx->stream, A, N, A, N, N, j, j + 1, tau, bstride, tstride);
  }
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

int qb_house_apply_q(qb_ctx* ctx, const cd* V, int64_t N, int r, const cd* tau, cd* Y,
                     int ncols) {
  for (int j = r - 1; j >= 0; --j)
    hipLaunchKernelGGL(k_house_apply<false>, dim3(ncols), dim3(256), 0, ctx->stream, V, N, Y, N,
                       N, j, 0, tau, (int64_t)0, (int64_t)0);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
