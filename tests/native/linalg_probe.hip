// Test harness: host-array wrappers around libqbhip's internal batched GEMM,
// LU and triangular solves, so GPU unit tests can compare each against NumPy.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../quantumbilliards.jl_amd/csrc/qb_linalg.h"

template <typename T>
static int run_gemm(qb_ctx* ctx, int M, int N, int K, const T* A, const T* B, T* C, int batch) {
  T *dA, *dB, *dC;
  size_t a = (size_t)M * K * batch, b = (size_t)K * N * batch, c = (size_t)M * N * batch;
  if (hipMalloc(&dA, a * sizeof(T)) || hipMalloc(&dB, b * sizeof(T)) || hipMalloc(&dC, c * sizeof(T)))
    return -2;
  hipMemcpy(dA, A, a * sizeof(T), hipMemcpyHostToDevice);
  hipMemcpy(dB, B, b * sizeof(T), hipMemcpyHostToDevice);
  hipMemcpy(dC, C, c * sizeof(T), hipMemcpyHostToDevice);
  int rc = gemm_sub<T>(ctx, M, N, K, dA, M, (int64_t)M * K, dB, K, (int64_t)K * N, dC, M,
                       (int64_t)M * N, batch);
  hipStreamSynchronize(ctx->stream);
  hipMemcpy(C, dC, c * sizeof(T), hipMemcpyDeviceToHost);
  hipFree(dA); hipFree(dB); hipFree(dC);
  return rc;
}

template <typename T>
static int run_getrf(qb_ctx* ctx, int n, T* A, int* ipiv, int batch) {
  T* dA;
  int *dP, *dI;
  size_t a = (size_t)n * n * batch;
  if (hipMalloc(&dA, a * sizeof(T)) || hipMalloc(&dP, (size_t)n * batch * sizeof(int)) ||
      hipMalloc(&dI, batch * sizeof(int)))
    return -2;
  hipMemcpy(dA, A, a * sizeof(T), hipMemcpyHostToDevice);
  hipMemset(dI, 0, batch * sizeof(int));
  int rc = qb_getrf_batched<T>(ctx, dA, n, n, (int64_t)n * n, dP, n, dI, batch);
  hipStreamSynchronize(ctx->stream);
  hipMemcpy(A, dA, a * sizeof(T), hipMemcpyDeviceToHost);
  hipMemcpy(ipiv, dP, (size_t)n * batch * sizeof(int), hipMemcpyDeviceToHost);
  hipFree(dA); hipFree(dP); hipFree(dI);
  return rc;
}

template <typename T>
static int run_trsm(qb_ctx* ctx, int n, const T* LU, T* B, int nrhs, int batch) {
  T *dA, *dB;
  size_t a = (size_t)n * n * batch, b = (size_t)n * nrhs * batch;
  if (hipMalloc(&dA, a * sizeof(T)) || hipMalloc(&dB, b * sizeof(T))) return -2;
  hipMemcpy(dA, LU, a * sizeof(T), hipMemcpyHostToDevice);
  hipMemcpy(dB, B, b * sizeof(T), hipMemcpyHostToDevice);
  int rc = qb_trsm_lu_batched<T>(ctx, dA, n, n, (int64_t)n * n, dB, nrhs, n, (int64_t)n * nrhs,
                                 batch);
  hipStreamSynchronize(ctx->stream);
  hipMemcpy(B, dB, b * sizeof(T), hipMemcpyDeviceToHost);
  hipFree(dA); hipFree(dB);
  return rc;
}

extern "C" {
int qbt_gemm_sub_c128(qb_ctx* c, int M, int N, int K, const cd* A, const cd* B, cd* C, int b) {
  return run_gemm<cd>(c, M, N, K, A, B, C, b);
}
int qbt_gemm_sub_c64(qb_ctx* c, int M, int N, int K, const cf* A, const cf* B, cf* C, int b) {
  return run_gemm<cf>(c, M, N, K, A, B, C, b);
}
int qbt_getrf_c128(qb_ctx* c, int n, cd* A, int* ipiv, int b) { return run_getrf<cd>(c, n, A, ipiv, b); }
int qbt_getrf_c64(qb_ctx* c, int n, cf* A, int* ipiv, int b) { return run_getrf<cf>(c, n, A, ipiv, b); }
int qbt_trsm_c128(qb_ctx* c, int n, const cd* LU, cd* B, int nrhs, int b) {
  return run_trsm<cd>(c, n, LU, B, nrhs, b);
}
int qbt_trsm_c64(qb_ctx* c, int n, const cf* LU, cf* B, int nrhs, int b) {
  return run_trsm<cf>(c, n, LU, B, nrhs, b);
}
void qbt_set_trsm_leaf_lds(int mode) { qb_set_trsm_leaf_lds(mode); }
}

// Timing of gemm_sub at one shape: buffers filled on the device, `iters`
// launches timed with events on the context stream; returns TFLOP/s.
__global__ void k_fill(double* p, int64_t n, double s) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = s * (double)((i * 2654435761ull) % 1000) / 1000.0;
}
template <typename T>
static double bench_gemm(qb_ctx* ctx, int M, int N, int K, int batch, int iters) {
  T *dA, *dB, *dC;
  size_t a = (size_t)M * K * batch, b = (size_t)K * N * batch, c = (size_t)M * N * batch;
  if (hipMalloc(&dA, a * sizeof(T)) || hipMalloc(&dB, b * sizeof(T)) || hipMalloc(&dC, c * sizeof(T)))
    return -1;
  int64_t na = a * sizeof(T) / 8, nb = b * sizeof(T) / 8, nc = c * sizeof(T) / 8;
  if (sizeof(T) == 16) {
    hipLaunchKernelGGL(k_fill, dim3((na + 255) / 256), dim3(256), 0, ctx->stream, (double*)dA, na, 1e-3);
    hipLaunchKernelGGL(k_fill, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, (double*)dB, nb, 1e-3);
    hipLaunchKernelGGL(k_fill, dim3((nc + 255) / 256), dim3(256), 0, ctx->stream, (double*)dC, nc, 1e-3);
  } else {
    hipMemsetAsync(dA, 0, a * sizeof(T), ctx->stream);
    hipMemsetAsync(dB, 0, b * sizeof(T), ctx->stream);
    hipMemsetAsync(dC, 0, c * sizeof(T), ctx->stream);
  }
  gemm_sub<T>(ctx, M, N, K, dA, M, (int64_t)M * K, dB, K, (int64_t)K * N, dC, M, (int64_t)M * N, batch);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, ctx->stream);
  for (int i = 0; i < iters; ++i)
    gemm_sub<T>(ctx, M, N, K, dA, M, (int64_t)M * K, dB, K, (int64_t)K * N, dC, M, (int64_t)M * N, batch);
  hipEventRecord(e1, ctx->stream);
  hipEventSynchronize(e1);
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  hipFree(dA); hipFree(dB); hipFree(dC);
  return 8.0 * M * N * (double)K * batch * iters / (ms * 1e-3) / 1e12;
}
extern "C" double qbt_gemm_bench(qb_ctx* c, int M, int N, int K, int batch, int iters, int dbl) {
  return dbl ? bench_gemm<cd>(c, M, N, K, batch, iters) : bench_gemm<cf>(c, M, N, K, batch, iters);
}

// Plain copy of `bytes` bytes (16 B per thread per step): the reference
// kernel for scaling FETCH_SIZE / WRITE_SIZE counters.  Returns GB/s.
__global__ void k_copy16(const double2* __restrict__ a, double2* __restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) b[i] = a[i];
}
extern "C" double qbt_copy_bench(qb_ctx* ctx, int64_t bytes, int iters) {
  double2 *a, *b;
  int64_t n = bytes / 16;
  if (hipMalloc(&a, n * 16) || hipMalloc(&b, n * 16)) return -1;
  hipMemsetAsync(a, 0, n * 16, ctx->stream);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, ctx->stream);
  for (int i = 0; i < iters; ++i)
    hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, ctx->stream, a, b, n);
  hipEventRecord(e1, ctx->stream);
  hipEventSynchronize(e1);
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  hipFree(a);
  hipFree(b);
  return 2.0 * n * 16 * iters / (ms * 1e-3) / 1e9;
}

// Force every shape class of one operand type to tile variant v (-1 restores
// the default choice), so tests can exercise each variant.
extern "C" void qbt_force_variant(qb_ctx* ctx, int ty, int v) {
  for (int k = 0; k < QB_GEMM_CLASSES; ++k) ctx->gemm_pick[ty][k] = ctx->gemm_pick_conc[ty][k] = v;
  ++ctx->gemm_epoch;
}

// The int8 (Ozaki) residual R[b] = V - T[b] X[b] for nb nodes (host arrays:
// T [nb][N][N], X [nb][N][r] column-major, V [N][r]); R out like X.
extern "C" int qbt_oz_residual(qb_ctx* ctx, int N, int r, int nb, const cd* T, const cd* X,
                               const cd* V, cd* R) {
  const size_t NN = (size_t)N * N, nv = (size_t)N * r;
  const size_t abytes = qb_oz_window_bytes(N, nb) - (size_t)nb * N * sizeof(int);
  const size_t sb = qb_oz_step_bytes(N, r, nb);
  const size_t Np = (size_t)qb_oz_pad(N), rp = (size_t)qb_oz_pad(r);
  const size_t bc = 16 * 2 * rp * Np * nb;
  cd *dT, *dX, *dV, *dR;
  signed char *A, *Bw, *Cw;
  int *E, *F;
  if (hipMalloc(&dT, NN * nb * sizeof(cd)) || hipMalloc(&dX, nv * nb * sizeof(cd)) ||
      hipMalloc(&dV, nv * sizeof(cd)) || hipMalloc(&dR, nv * nb * sizeof(cd)) ||
      hipMalloc(&A, abytes) || hipMalloc(&E, (size_t)nb * N * sizeof(int)) ||
      hipMalloc(&Bw, bc) || hipMalloc(&Cw, bc) || hipMalloc(&F, rp * nb * sizeof(int)))
    return -2;
  (void)sb;
  hipMemcpy(dT, T, NN * nb * sizeof(cd), hipMemcpyHostToDevice);
  hipMemcpy(dX, X, nv * nb * sizeof(cd), hipMemcpyHostToDevice);
  hipMemcpy(dV, V, nv * sizeof(cd), hipMemcpyHostToDevice);
  int rc = qb_oz_prepare(ctx, dT, N, nb, A, E);
  if (!rc) rc = qb_oz_residual(ctx, A, E, N, dX, (int64_t)nv, r, dV, dR, (int64_t)nv, nb, Bw, F, Cw);
  hipStreamSynchronize(ctx->stream);
  hipMemcpy(R, dR, nv * nb * sizeof(cd), hipMemcpyDeviceToHost);
  hipFree(dT); hipFree(dX); hipFree(dV); hipFree(dR); hipFree(A); hipFree(E); hipFree(Bw);
  hipFree(Cw); hipFree(F);
  return rc;
}

// The int8 product itself on one modulus-free path: time of qb_oz_residual
// (ms per call, after one warmup) for nb nodes of N x N with r columns.
extern "C" double qbt_oz_bench(qb_ctx* ctx, int N, int r, int nb, int iters) {
  const size_t NN = (size_t)N * N, nv = (size_t)N * r;
  const size_t abytes = qb_oz_window_bytes(N, nb) - (size_t)nb * N * sizeof(int);
  const size_t Np = (size_t)qb_oz_pad(N), rp = (size_t)qb_oz_pad(r);
  const size_t bc = 16 * 2 * rp * Np * nb;
  cd *dT, *dX, *dV, *dR;
  signed char *A, *Bw, *Cw;
  int *E, *F;
  if (hipMalloc(&dT, NN * nb * sizeof(cd)) || hipMalloc(&dX, nv * nb * sizeof(cd)) ||
      hipMalloc(&dV, nv * sizeof(cd)) || hipMalloc(&dR, nv * nb * sizeof(cd)) ||
      hipMalloc(&A, abytes) || hipMalloc(&E, (size_t)nb * N * sizeof(int)) ||
      hipMalloc(&Bw, bc) || hipMalloc(&Cw, bc) || hipMalloc(&F, rp * nb * sizeof(int)))
    return -1.0;
  hipMemset(dT, 0x3f, NN * nb * sizeof(cd));
  hipMemset(dX, 0x3f, nv * nb * sizeof(cd));
  hipMemset(dV, 0, nv * sizeof(cd));
  qb_oz_prepare(ctx, dT, N, nb, A, E);
  qb_oz_residual(ctx, A, E, N, dX, (int64_t)nv, r, dV, dR, (int64_t)nv, nb, Bw, F, Cw);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, ctx->stream);
  for (int it = 0; it < iters; ++it)
    qb_oz_residual(ctx, A, E, N, dX, (int64_t)nv, r, dV, dR, (int64_t)nv, nb, Bw, F, Cw);
  hipEventRecord(e1, ctx->stream);
  hipEventSynchronize(e1);
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  hipFree(dT); hipFree(dX); hipFree(dV); hipFree(dR); hipFree(A); hipFree(E); hipFree(Bw);
  hipFree(Cw); hipFree(F);
  return ms / iters;
}
