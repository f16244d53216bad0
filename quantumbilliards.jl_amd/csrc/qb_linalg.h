// Batched complex LU / triangular solves / helpers (qb_linalg.hip).
#pragma once
#include "qb_internal.h"

template <typename T>
int gemm_sub(qb_ctx* ctx, int M, int N, int K, const T* A, int64_t lda, int64_t sA, const T* B,
             int64_t ldb, int64_t sB, T* C, int64_t ldc, int64_t sC, int batch);
// gemm_sub with the given variant of the operand type's tile table (for a
// shape outside the tuned classes); an unknown variant takes the default tile.
template <typename T>
int gemm_sub_v(qb_ctx* ctx, int v, int M, int N, int K, const T* A, int64_t lda, int64_t sA,
               const T* B, int64_t ldb, int64_t sB, T* C, int64_t ldc, int64_t sC, int batch);
// complex128 variant 13: the 32 x 32 x 8 Gauss-3M tile with 19 KB of LDS
#define QB_GEMM_LEAN_C128 13
// first_done: the first panel was already factored by qb_getrf_first_panel
// (the call then starts by waiting for it).
template <typename T>
int qb_getrf_batched(qb_ctx* ctx, T* A, int n, int64_t lda, int64_t sA, int* ipiv, int64_t sP,
                     int* info, int batch, bool first_done = false);
// The first panel of qb_getrf_batched (columns [0, min(n, qb_lu_nb(n)))) on
// the context's second stream, after the work already enqueued on the main
// stream; pass first_done to the qb_getrf_batched / qb_getrf_graph call that
// factors the rest.  Work on other columns may run on the main stream
// meanwhile.
template <typename T>
int qb_getrf_first_panel(qb_ctx* ctx, T* A, int n, int64_t lda, int64_t sA, int* ipiv,
                         int64_t sP, int* info, int batch);
int qb_lu_nb(int n);
// qb_getrf_batched through a cached hipGraph when ctx->graphs is set and
// launch accounting is off: the first call with given buffers captures the
// launch sequence (both streams) and instantiates it, later calls replay it.
// With first_done the call is not captured.
template <typename T>
int qb_getrf_graph(qb_ctx* ctx, T* A, int n, int64_t lda, int64_t sA, int* ipiv, int64_t sP,
                   int* info, int batch, bool first_done = false);
void qb_graphs_clear(qb_ctx* ctx);
// Triangular-solve leaf kernel for every later launch of the process: 1 the
// one-wave LDS leaf (k_trsm_leaf), 0 the register leaf (k_trsm_leaf_reg), -1
// back to the default (QB_TRSM_LEAF=lds, else the register leaf for launches
// of at least 8 waves per CU and the LDS leaf below).
void qb_set_trsm_leaf_lds(int mode);
// Triangular solves only (the row permutation is applied separately, fused
// into the gather that fills B).
template <typename T>
int qb_trsm_lu_batched(qb_ctx* ctx, const T* LU, int n, int64_t lda, int64_t sA, T* B, int nrhs,
                       int64_t ldb, int64_t sB, int batch);

int qb_perm_from_ipiv(qb_ctx* ctx, const int* ipiv, int* perm, int n, int batch);
// dst[b][i + c*n] = src[(b*sb)][perm[b][i] + c*n] for a permuted broadcast /
// copy of right-hand sides (src stride 0 broadcasts one block to all nodes).
int qb_gather_rows_cd(qb_ctx* ctx, const cd* src, int64_t s_src, const int* perm, cd* dst, int n,
                      int nrhs, int batch);
int qb_gather_rows_cd2cf(qb_ctx* ctx, const cd* src, int64_t s_src, const int* perm, cf* dst,
                         int n, int nrhs, int batch);
int qb_add_cf_to_cd(qb_ctx* ctx, const cf* d, cd* x, int64_t n);
// x[b*ns + i] += d[b*ns + i], i < n, for the nb batch entries (ctx->bix aware)
int qb_add_cf_to_cd_nodes(qb_ctx* ctx, const cf* d, cd* x, int64_t n, int64_t ns, int nb);
int qb_cd_to_cf(qb_ctx* ctx, const cd* x, cf* y, int64_t n);
// R[b*ns + e] = V[e] for e < nv, b < nb
int qb_bcast(qb_ctx* ctx, const cd* V, cd* R, int64_t nv, int64_t ns, int nb);
int qb_moments(qb_ctx* ctx, const cd* X, int64_t nv, int nq, const cd* w, const cd* wz, cd* A0,
               cd* A1);
// out[b] = max_{e < n_per} |X[b*stride + e]| for the physical node b of each
// of the nb batch entries (ctx->bix aware); out[0..max(n_out, nb)) cleared first
int qb_node_maxabs(qb_ctx* ctx, const cd* X, int64_t n_per, int64_t stride, int nb,
                   double* out_dev, int n_out);
int qb_node_maxabs32(qb_ctx* ctx, const cf* X, int64_t n_per, int nb, double* out_dev);
int qb_node_rownorm(qb_ctx* ctx, const cd* A, int64_t n, int nb, double* out_dev);

// GEMM autotuning (qb_linalg.hip): shape class, tuning of one shape (returns
// the chosen variant or a negative error), and variant names (ty 0: complex64).
int qb_gemm_class(int M, int N, int K);
// Panel width of the blocked LU of n x n matrices.
int qb_lu_nb(int n);
template <typename T>
int qb_gemm_tune(qb_ctx* ctx, int M, int N, int K, int batch, double* best_tflops);
const char* qb_gemm_variant_name(int ty, int v);

// Residual check (qb_residual.hip): for nb matrices T (N x N, stride N*N) and
// vectors phi (stride N), partial products part[s][b][i] and partial row
// sums rowabs[s][b][i] of |T| over nsplit column chunks, and column sums
// colsum[b][j] of |T|.
int qb_gemv_norms(qb_ctx* ctx, const cd* T, const cd* phi, int64_t N, int nb, int nsplit,
                  cd* part, double* rowabs, double* colsum);

// Householder QR (qb_qr.hip): the first r columns of the N x ncols block A
// (ld = N) are factored in place (R on and above the diagonal, reflectors
// below, tau[r]) and H_j^H is applied to all columns right of each j.  nb
// such blocks are factored together, block b at A + b * bstride with its tau
// at tau + b * tstride (2r launches for all of them).
int qb_house_qr(qb_ctx* ctx, cd* A, int64_t N, int r, int ncols, cd* tau, int nb,
                int64_t bstride, int64_t tstride);
// Y <- Q Y for the N x ncols block Y (ld = N), Q = H_0 ... H_{r-1} from V.
int qb_house_apply_q(qb_ctx* ctx, const cd* V, int64_t N, int r, const cd* tau, cd* Y,
                     int ncols);

// FP64-accurate refinement residual on int8 MFMA (qb_ozaki.hip).
// Padded size (multiple of 128) of the residue planes.
int64_t qb_oz_pad(int64_t n);
// Device bytes of the residues of T for nq nodes, and of the per-step
// buffers of nslots nodes with r probes.
size_t qb_oz_window_bytes(int64_t N, int nq);
size_t qb_oz_step_bytes(int64_t N, int r, int nslots);
// Row exponents and residues of T (nq matrices N x N, stride N*N):
// A [nq][16][2][Np][Np] int8, erow [nq][N].
int qb_oz_prepare(qb_ctx* ctx, const cd* T, int64_t N, int nq, signed char* A, int* erow);
// R[b] = V - T[b] X[b] for the nb nodes of ctx->bix (all nodes 0..nb-1 when
// null); X and R strided sX / sR per node, ld N; Bw, fcol, Cw per-step
// buffers of nb slots (qb_oz_step_bytes).
int qb_oz_residual(qb_ctx* ctx, const signed char* A, const int* erow, int64_t N, const cd* X,
                   int64_t sX, int r, const cd* V, cd* R, int64_t sR, int nb, signed char* Bw,
                   int* fcol, signed char* Cw);
