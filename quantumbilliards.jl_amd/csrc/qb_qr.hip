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

// ---------------------------------------------------------------------------
// Blocked (compact WY) application of a panel of QB_QR_NB reflectors to the
// trailing columns (LAPACK zlarft + zlarfb, forward, columnwise): with
// V = [v_j0 ... v_j0+jb-1] (unit diagonal, zeros above) and the upper
// triangular T of H_j0 ... H_j0+jb-1 = I - V T V^H, the sweep applies
// (I - V T V^H)^H = I - V T^H V^H: W = V^H C, W <- T^H W, C <- C - V W, as
// three GEMMs instead of one pass over C per reflector.
// ---------------------------------------------------------------------------
#define QB_QR_NB 32

// V (m x jb, ld m) and Vh = -V^H (jb x m, ld jb) of the panel at column j0:
// rows j0..N-1 of columns j0..j0+jb-1.  Grid (ceil(m/256), jb, nb).
__global__ void k_qr_form_v(const cd* __restrict__ A, int64_t N, int j0, int jb, cd* __restrict__ V,
                            cd* __restrict__ Vh, int64_t bstride, int64_t vstride) {
  const int64_t m = N - j0;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= m) return;
  A += (int64_t)blockIdx.z * bstride;
  V += (int64_t)blockIdx.z * vstride;
  Vh += (int64_t)blockIdx.z * vstride;
  cd v = mk(0.0, 0.0);
  if (i == c)
    v = mk(1.0, 0.0);
  else if (i > c)
    v = A[(j0 + i) + (int64_t)(j0 + c) * N];
  V[i + (int64_t)c * m] = v;
  Vh[c + (int64_t)i * jb] = mk(-v.re, v.im);
}

// T of the panel from tau and G = V^H V (jb x jb, ld jb): T(i,i) = tau_i,
// T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i); stored (jb x jb, ld jb) as
// -T^H (ADJ, the QR sweep) or -T (forming Q).  One workgroup of jb threads
// per block (jb <= QB_QR_NB).
template <bool ADJ>
__global__ void k_qr_form_t(const cd* __restrict__ G, const cd* __restrict__ tau, int j0, int jb,
                            cd* __restrict__ Th, int64_t gstride, int64_t tstride) {
  __shared__ cd T[QB_QR_NB][QB_QR_NB + 1];
  __shared__ cd z[QB_QR_NB];
  G += (int64_t)blockIdx.x * gstride;
  tau += (int64_t)blockIdx.x * tstride;
  Th += (int64_t)blockIdx.x * gstride;
  const int t = threadIdx.x;
  for (int c = 0; c < jb; ++c) T[t][c] = mk(0.0, 0.0);
  __syncthreads();
  for (int i = 0; i < jb; ++i) {
    const cd ti = tau[j0 + i];
    if (t == 0) T[i][i] = ti;
    if (t < i) z[t] = G[t + (int64_t)i * jb];
    __syncthreads();
    if (t < i) {
      cd s = mk(0.0, 0.0);
      for (int c = t; c < i; ++c) s = s + T[t][c] * z[c];
      T[t][i] = mk(-(ti.re * s.re - ti.im * s.im), -(ti.re * s.im + ti.im * s.re));
    }
    __syncthreads();
  }
  for (int c = 0; c < jb; ++c) {
    if (ADJ) {
      const cd x = T[c][t];  // Th(t, c) = -conj(T(c, t))
      Th[t + (int64_t)c * jb] = mk(-x.re, x.im);
    } else {
      const cd x = T[t][c];
      Th[t + (int64_t)c * jb] = mk(-x.re, -x.im);
    }
  }
}

__global__ void k_qr_zero(cd* __restrict__ X, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) X[i] = mk(0.0, 0.0);
}

// out[i] = sum over the nch partial products part[c * n + i].
__global__ void k_qr_sum_parts(const cd* __restrict__ part, int nch, int64_t n, cd* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cd s = part[i];
  for (int c = 1; c < nch; ++c) s = s + part[(int64_t)c * n + i];
  out[i] = s;
}

// out (M x N, ld M) = -Ah B with Ah M x K (ld M) and B K x N (ld ldb), M <= 32,
// K long: the products of K-chunks of QB_QR_KC run as one batch of 32 x 32
// tiles into part (zeroed first) and are summed in chunk order.  A single
// tile row over the whole K would leave most of the chip idle.
#define QB_QR_KC 256
static int qb_qr_skinny(qb_ctx* ctx, int M, int N, int K, const cd* Ah, const cd* B, int64_t ldb,
                        cd* part, cd* out) {
  const int nfull = K / QB_QR_KC, rem = K % QB_QR_KC, nch = nfull + (rem > 0);
  const int64_t n = (int64_t)M * N;
  hipLaunchKernelGGL(k_qr_zero, dim3((unsigned)((nch * n + 255) / 256)), dim3(256), 0, ctx->stream,
                     part, nch * n);
  int rc = gemm_sub_v<cd>(ctx, QB_GEMM_LEAN_C128, M, N, QB_QR_KC, Ah, M, (int64_t)QB_QR_KC * M, B,
                          ldb, QB_QR_KC, part, M, n, nfull);
  if (!rc && rem)
    rc = gemm_sub_v<cd>(ctx, QB_GEMM_LEAN_C128, M, N, rem, Ah + (int64_t)nfull * QB_QR_KC * M, M, 0,
                        B + (int64_t)nfull * QB_QR_KC, ldb, 0, part + nfull * n, M, 0, 1);
  if (rc) return rc;
  hipLaunchKernelGGL(k_qr_sum_parts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     part, nch, n, out);
  return QB_OK;
}

int qb_house_qr(qb_ctx* ctx, cd* A, int64_t N, int r, int ncols, cd* tau, int nb,
                int64_t bstride, int64_t tstride) {
  if (r > N || ncols < r) return qb_fail(ctx, QB_E_ARG, "QR needs r <= N and ncols >= r");
  if (nb < 1 || nb > 65535) return qb_fail(ctx, QB_E_ARG, "QR batch must be in [1, 65535]");
  const int* bix = ctx->bix;
  ctx->bix = nullptr;  // the GEMMs below address their blocks themselves
  const int jbm = QB_QR_NB;
  const bool blocked = r > jbm && ncols > r;
  // work buffers: V and Vh (N x jb), G and Th (jb x jb), W and W2 (jb x
  // ncols) per block, and the K-chunk partial products of one block
  const int64_t vstride = N * jbm, gstride = (int64_t)jbm * jbm, wstride = (int64_t)jbm * ncols;
  const int64_t pn = ((N + QB_QR_KC - 1) / QB_QR_KC) * wstride;
  const size_t welems = (size_t)nb * (2 * vstride + 2 * gstride + 2 * wstride) + (size_t)pn;
  const size_t wbytes = blocked ? welems * sizeof(cd) : 0;
  cd* wk = nullptr;
  int rc = blocked ? qb_dev_alloc(ctx, (void**)&wk, wbytes) : QB_OK;
  if (rc) {
    ctx->bix = bix;
    return rc;
  }
  cd *V = wk, *Vh = V + nb * vstride, *G = Vh + nb * vstride, *Th = G + nb * gstride,
     *W = Th + nb * gstride, *W2 = W + nb * wstride, *part = W2 + nb * wstride;
  for (int j0 = 0; j0 < r && !rc; j0 += (blocked ? jbm : r)) {
    const int jb = blocked ? (r - j0 < jbm ? r - j0 : jbm) : r;
    // the panel: reflectors j0..j0+jb-1 applied within the panel's columns
    // (all trailing columns when unblocked)
    const int cend = blocked ? j0 + jb : ncols;
    for (int j = j0; j < j0 + jb; ++j) {
      hipLaunchKernelGGL(k_house, dim3(1, nb), dim3(1024), 0, ctx->stream, A, N, N, j, tau,
                         bstride, tstride);
      if (cend - j - 1 > 0)
        hipLaunchKernelGGL(k_house_apply<true>, dim3(cend - j - 1, nb), dim3(256), 0,
                           ctx->stream, A, N, A, N, N, j, j + 1, tau, bstride, tstride);
    }
    const int nc = ncols - (j0 + jb);
    if (!blocked || nc <= 0) continue;
    const int64_t m = N - j0;
    hipLaunchKernelGGL(k_qr_form_v, dim3((unsigned)((m + 255) / 256), jb, nb), dim3(256), 0,
                       ctx->stream, A, N, j0, jb, V, Vh, bstride, vstride);
    // per block: G = V^H V and W = V^H C (C = columns j0+jb.. of rows
    // j0..N-1), as -Vh V and -Vh C
    for (int b = 0; b < nb && !rc; ++b) {
      rc = qb_qr_skinny(ctx, jb, jb, (int)m, Vh + b * vstride, V + b * vstride, m, part,
                        G + b * gstride);
      if (!rc)
        rc = qb_qr_skinny(ctx, jb, nc, (int)m, Vh + b * vstride,
                          A + b * bstride + j0 + (int64_t)(j0 + jb) * N, N, part, W + b * wstride);
    }
    if (rc) break;
    hipLaunchKernelGGL(k_qr_form_t<true>, dim3(nb), dim3(jb), 0, ctx->stream, G, tau, j0, jb, Th,
                       gstride, tstride);
    // W2 = T^H W = 0 - Th W
    hipLaunchKernelGGL(k_qr_zero, dim3((unsigned)((nb * wstride + 255) / 256)), dim3(256), 0,
                       ctx->stream, W2, (int64_t)nb * wstride);
    rc = gemm_sub_v<cd>(ctx, QB_GEMM_LEAN_C128, jb, nc, jb, Th, jb, gstride, W, jb, wstride, W2, jb,
                        wstride, nb);
    // C -= V W2
    cd* Cm = A + j0 + (int64_t)(j0 + jb) * N;
    if (!rc) rc = gemm_sub<cd>(ctx, (int)m, nc, jb, V, m, vstride, W2, jb, wstride, Cm, N, bstride, nb);
  }
  if (blocked) qb_dev_free(ctx, wk, wbytes);
  ctx->bix = bix;
  if (rc) return rc;
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

// Y <- Q Y = H_0 H_1 ... H_{r-1} Y (Y N x ncols, ld N): one reflector at a
// time for r <= QB_QR_NB, else panel by panel from the last, each as
// Y(j0:, :) -= V (T (V^H Y(j0:, :))).
int qb_house_apply_q(qb_ctx* ctx, const cd* V, int64_t N, int r, const cd* tau, cd* Y,
                     int ncols) {
  if (r <= QB_QR_NB) {
    for (int j = r - 1; j >= 0; --j)
      hipLaunchKernelGGL(k_house_apply<false>, dim3(ncols), dim3(256), 0, ctx->stream, V, N, Y, N,
                         N, j, 0, tau, (int64_t)0, (int64_t)0);
    QB_HIP(ctx, hipGetLastError());
    return QB_OK;
  }
  const int* bix = ctx->bix;
  ctx->bix = nullptr;
  const int jbm = QB_QR_NB;
  const int64_t vn = N * jbm, gn = (int64_t)jbm * jbm, wn = (int64_t)jbm * ncols;
  const int64_t pn = ((N + QB_QR_KC - 1) / QB_QR_KC) * (wn > gn ? wn : gn);
  const size_t wbytes = (size_t)(2 * vn + 2 * gn + 2 * wn + pn) * sizeof(cd);
  cd* wk = nullptr;
  int rc = qb_dev_alloc(ctx, (void**)&wk, wbytes);
  if (rc) {
    ctx->bix = bix;
    return rc;
  }
  cd *Vb = wk, *Vh = Vb + vn, *G = Vh + vn, *Tn = G + gn, *W = Tn + gn, *W2 = W + wn,
     *part = W2 + wn;
  for (int j0 = ((r - 1) / jbm) * jbm; j0 >= 0 && !rc; j0 -= jbm) {
    const int jb = r - j0 < jbm ? r - j0 : jbm;
    const int64_t m = N - j0;
    hipLaunchKernelGGL(k_qr_form_v, dim3((unsigned)((m + 255) / 256), jb, 1), dim3(256), 0,
                       ctx->stream, V, N, j0, jb, Vb, Vh, (int64_t)0, (int64_t)0);
    rc = qb_qr_skinny(ctx, jb, jb, (int)m, Vh, Vb, m, part, G);
    if (!rc) rc = qb_qr_skinny(ctx, jb, ncols, (int)m, Vh, Y + j0, N, part, W);
    if (rc) break;
    hipLaunchKernelGGL(k_qr_form_t<false>, dim3(1), dim3(jb), 0, ctx->stream, G, tau, j0, jb, Tn,
                       (int64_t)0, (int64_t)0);
    // W2 = T W = 0 - (-T) W, then Y(j0:, :) -= V W2
    hipLaunchKernelGGL(k_qr_zero, dim3((unsigned)((wn + 255) / 256)), dim3(256), 0, ctx->stream,
                       W2, wn);
    rc = gemm_sub_v<cd>(ctx, QB_GEMM_LEAN_C128, jb, ncols, jb, Tn, jb, 0, W, jb, 0, W2, jb, 0, 1);
    if (!rc) rc = gemm_sub<cd>(ctx, (int)m, ncols, jb, Vb, m, 0, W2, jb, 0, Y + j0, N, 0, 1);
  }
  qb_dev_free(ctx, wk, wbytes);
  ctx->bix = bix;
  if (rc) return rc;
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
