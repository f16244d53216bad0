// FP64 residual R = V - T X of the mixed-precision refinement, computed
// exactly enough on the int8 matrix cores (the Ozaki scheme with the Chinese
// remainder theorem, "Ozaki scheme II": Ozaki, Uchino, Imamura 2025).
//
// The refinement of the Beyn window (beyn.jl:278-285 solved in complex64,
// then corrected) needs R_j = V - T(z_j) X_j to about FP64 accuracy for
// every contour node j.  On MI355X the int8 MFMA rate (5.0 POP/s dense) is
// 64x the FP64 rate (78.6 TF/s), so an exact integer product modulo small
// moduli beats the FP64 GEMM even though it takes 16 x 4 int8 products per
// complex product:
//
//  1. Row i of T_j is scaled by 2^e_i so that its largest real or imaginary
//     part lies in [2^52, 2^53), and rounded to integers T' (exact in a
//     double; T' is what an FP64 entry of that row resolves).  Column c of
//     X_j is scaled by 2^f_c the same way, X'.  The error of T' X' against
//     T X (scaled) is below 2^-53 N max|T_i.| max|X_.c| per entry.
//  2. The residues of T' and X' modulo 16 pairwise coprime moduli m_l <= 256
//     are int8 (symmetric range).  T' is reduced once per window (T is
//     fixed), X' once per refinement step.
//  3. For every modulus, the complex product over the int8 residues is 4
//     int8 matrix products with exact int32 sums (|sum| <= 2 N 128^2 < 2^31
//     for N < 2^16): Re = Tr Xr + Ti (-Xi), Im = Tr Xi + Ti Xr.
//     Their residues modulo m_l are stored (int8).
//  4. Garner's algorithm turns the 16 residues into mixed-radix digits, and
//     Horner's rule in FP64 into the integer (T' X')_ic to 16 ulp; the
//     product M of the moduli (2^125.4) exceeds 2 |Re|, 2 |Im| <= 4 N 2^106,
//     so the representation is unique for N up to 2^17.  Then
//     R_ic = V_ic - 2^-(e_i + f_c) (T' X')_ic.
//
// Layouts (Np = N rounded up to 128, rp = r rounded up to 128):
//   A  [node][l][plane re/im][i][k]  int8, row-major (k fastest), Np x Np
//   B  [slot][l][plane re/im][c][k]  int8, k fastest, rp x Np
//   Cr [slot][l][plane re/im][c][i]  int8 residues of T' X', rp x Np
//   erow [node][N], fcol [slot][rp]  the exponents e_i, f_c
// Padding rows / columns hold zero residues.  `slot` is the position of a
// node in the launch's node list (ctx->bix), so the per-step buffers hold
// only the nodes still being refined.
#include <math.h>
#include <string.h>

#include "qb_internal.h"
#include "qb_linalg.h"

#define OZ_L 16

typedef int v4i __attribute__((ext_vector_type(4)));

struct OzConst {
  int m[OZ_L];
  double rcp[OZ_L];
  float rcpf[OZ_L];
  unsigned char inv[OZ_L][OZ_L];  // inv[l][p] = m_p^{-1} mod m_l, p < l
};

static OzConst oz_const() {
  // pairwise coprime: 2^8, 3*5*17, 11*23, 251, 13*19, 241, 239, 233, 229, 227,
  // 223, 7*31, 211, 199, 197, 193; product 2^125.37
  static const int mods[OZ_L] = {256, 255, 253, 251, 247, 241, 239, 233,
                                 229, 227, 223, 217, 211, 199, 197, 193};
  OzConst c;
  memset(&c, 0, sizeof(c));
  for (int l = 0; l < OZ_L; ++l) {
    c.m[l] = mods[l];
    c.rcp[l] = 1.0 / mods[l];
    c.rcpf[l] = 1.0f / (float)mods[l];
    for (int p = 0; p < l; ++p) {
      const int a = mods[p] % mods[l];
      int inv = 0;
      for (int x = 1; x < mods[l]; ++x)
        if ((a * x) % mods[l] == 1) {
          inv = x;
          break;
        }
      c.inv[l][p] = (unsigned char)inv;
    }
  }
  return c;
}

// Symmetric residue in [-(m/2), (m-1)/2] of an integer-valued double
// |x| < 2^53: x - q m is exact in the fused multiply-add and off by at most
// one modulus from the symmetric range.
__device__ __forceinline__ int oz_sres_d(double x, int m, double rcp) {
  const double q = rint(x * rcp);
  int r = (int)fma(-q, (double)m, x);
  if (r > (m - 1) / 2) r -= m;
  if (r < -(m / 2)) r += m;
  return r;
}

// The same for |x| < 2^17 through single precision (x * rcp errs by < 2^-7).
__device__ __forceinline__ int oz_sres_i(int x, int m, float rcp) {
  const int q = (int)rintf((float)x * rcp);
  int r = x - q * m;
  if (r > (m - 1) / 2) r -= m;
  if (r < -(m / 2)) r += m;
  return r;
}

// Row exponents: e_i = 52 - ilogb(max_k max(|Re T_ik|, |Im T_ik|)) (0 for a
// zero row).  One thread per row, coalesced along the column.
__global__ __launch_bounds__(256) void k_oz_rowexp(const cd* __restrict__ T, int64_t N,
                                                   int64_t sT, int* __restrict__ erow) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t node = blockIdx.y;
  if (i >= N) return;
  const cd* t = T + node * sT + i;
  double m0 = 0.0, m1 = 0.0;
  int64_t j = 0;
  for (; j + 1 < N; j += 2) {
    const cd a = t[j * N], b = t[(j + 1) * N];
    m0 = fmax(m0, fmax(fabs(a.re), fabs(a.im)));
    m1 = fmax(m1, fmax(fabs(b.re), fabs(b.im)));
  }
  if (j < N) {
    const cd a = t[j * N];
    m0 = fmax(m0, fmax(fabs(a.re), fabs(a.im)));
  }
  const double mx = fmax(m0, m1);
  erow[node * N + i] = mx > 0.0 ? 52 - ilogb(mx) : 0;
}

// T' residues, transposed to k-fastest rows through LDS: one block per 64 x 64
// tile (rows i0.., columns k0..) of one node; zeros beyond N.
__global__ __launch_bounds__(256) void k_oz_slice_T(const cd* __restrict__ T, int64_t N,
                                                    int64_t Np, int64_t sT,
                                                    const int* __restrict__ erow,
                                                    signed char* __restrict__ A, OzConst oc) {
  __shared__ __align__(16) signed char tile[2][64][80];  // [plane][i][k]
  const int64_t node = blockIdx.z;
  const int64_t i0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  const int tid = threadIdx.x, il = tid & 63, kq = tid >> 6;
  const int64_t i = i0 + il;
  const int e = i < N ? erow[node * N + i] : 0;
  double vr[16], vi[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int64_t k = k0 + kq + 4 * s;
    cd v = mk(0.0, 0.0);
    if (i < N && k < N) v = T[node * sT + i + k * N];
    vr[s] = rint(ldexp(v.re, e));
    vi[s] = rint(ldexp(v.im, e));
  }
  const size_t plane = (size_t)Np * Np;
  for (int l = 0; l < OZ_L; ++l) {
    const int m = oc.m[l];
    const double rc = oc.rcp[l];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      tile[0][il][kq + 4 * s] = (signed char)oz_sres_d(vr[s], m, rc);
      tile[1][il][kq + 4 * s] = (signed char)oz_sres_d(vi[s], m, rc);
    }
    __syncthreads();
    // 2 planes x 64 rows x 4 chunks of 16 bytes
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      const int q = tid + 256 * w;
      const int p = q >> 8, rl = (q >> 2) & 63, part = q & 3;
      const int4 v = *(const int4*)&tile[p][rl][16 * part];
      signed char* dst = A + ((size_t)(node * OZ_L + l) * 2 + p) * plane +
                         (size_t)(i0 + rl) * Np + k0 + 16 * part;
      *(int4*)dst = v;
    }
    __syncthreads();
  }
}

// X' residues of the listed nodes: one block per (column c < rp, slot b);
// the column maximum first, then the scaled integers modulo every m_l, four
// rows per thread and store.
__global__ __launch_bounds__(256) void k_oz_slice_X(const cd* __restrict__ X, int64_t N,
                                                    int64_t Np, int r, int rp, int64_t sX,
                                                    const int* __restrict__ bix,
                                                    signed char* __restrict__ B,
                                                    int* __restrict__ fcol, OzConst oc) {
  const int c = blockIdx.x;
  const int b = blockIdx.y;
  const int64_t node = bix ? bix[b] : b;
  const int tid = threadIdx.x;
  const cd* x = X + node * sX + (int64_t)c * N;
  double mx = 0.0;
  if (c < r)
    for (int64_t k = tid; k < N; k += 256) mx = fmax(mx, fmax(fabs(x[k].re), fabs(x[k].im)));
  __shared__ double red[4];
  for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const int f = mx > 0.0 ? 52 - ilogb(mx) : 0;
  if (tid == 0) fcol[(int64_t)b * rp + c] = f;
  const size_t plane = (size_t)rp * Np;
  for (int64_t k = 4 * (int64_t)tid; k < Np; k += 1024) {
    double xr[4], xi[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      cd v = mk(0.0, 0.0);
      if (c < r && k + s < N) v = x[k + s];
      xr[s] = rint(ldexp(v.re, f));
      xi[s] = rint(ldexp(v.im, f));
    }
    for (int l = 0; l < OZ_L; ++l) {
      const int m = oc.m[l];
      const double rc = oc.rcp[l];
      unsigned pr = 0, pi = 0;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        pr |= (unsigned)(oz_sres_d(xr[s], m, rc) & 0xff) << (8 * s);
        pi |= (unsigned)(oz_sres_d(xi[s], m, rc) & 0xff) << (8 * s);
      }
      signed char* base = B + ((size_t)((int64_t)b * OZ_L + l) * 2) * plane + (size_t)c * Np + k;
      *(unsigned*)base = pr;
      *(unsigned*)(base + plane) = pi;
    }
  }
}

// Bytewise two's-complement negation of four packed int8.
__device__ __forceinline__ int oz_neg4(int x) {
  const unsigned t = ~(unsigned)x;
  return (int)(((t & 0x7f7f7f7fu) + 0x01010101u) ^ (t & 0x80808080u));
}

// Residues of T' X' modulo m_l for every listed node and modulus:
// workgroup tile 128 (rows i) x 128 (columns c), k-tile 64, 4 waves of 64 x 64
// (4 x 4 blocks of v_mfma_i32_16x16x64_i8).  Operand lanes: row / column
// lane % 16, bytes k = 16 (lane / 16) + 0..15 of the k-tile for both A and B
// (the same k assignment on both sides, so the products pair up); result
// register v of a lane holds D[4 (lane / 16) + v][lane % 16].  grid.x walks
// column tiles fastest, then row tiles, so the workgroups that share an A
// tile run together; grid.y = modulus, grid.z = slot.
#define OZ_BM 128
#define OZ_BN 128
#define OZ_BK 64
#define OZ_LDS_ROW 80  // bytes per LDS row: 64 + 16 of padding (conflict-free b128 reads)
__global__ void __launch_bounds__(256, 1)
k_oz_gemm(const signed char* __restrict__ A, const signed char* __restrict__ B,
          signed char* __restrict__ Cr, int64_t Np, int rp, int ntn, const int* __restrict__ bix,
          OzConst oc) {
  __shared__ __align__(16) signed char As[2][2][OZ_BM * OZ_LDS_ROW];
  __shared__ __align__(16) signed char Bs[2][2][OZ_BN * OZ_LDS_ROW];
  const int l = blockIdx.y;
  const int64_t b = blockIdx.z;
  const int64_t node = bix ? bix[b] : b;
  const int nt = blockIdx.x % ntn, mt = blockIdx.x / ntn;
  const int64_t m0 = (int64_t)mt * OZ_BM, n0 = (int64_t)nt * OZ_BN;
  const size_t aplane = (size_t)Np * Np, bplane = (size_t)rp * Np;
  const signed char* Ap = A + ((size_t)(node * OZ_L + l) * 2) * aplane;
  const signed char* Bp = B + ((size_t)(b * OZ_L + l) * 2) * bplane;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave & 1) * 64, wn0 = (wave >> 1) * 64;
  const int li = lane & 15, lk = lane >> 4;

  v4i acc_re[4][4], acc_im[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc_re[i][j] = v4i{0, 0, 0, 0};
      acc_im[i][j] = v4i{0, 0, 0, 0};
    }
  // 1024 chunks of 16 bytes per operand and k-tile: chunk q -> plane q >> 9,
  // row (q >> 2) & 127, bytes 16 (q & 3)
  int4 ra[4], rb[4];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int q = tid + 256 * s;
      const int p = q >> 9, row = (q >> 2) & 127, part = q & 3;
      ra[s] = *(const int4*)(Ap + p * aplane + (size_t)(m0 + row) * Np + k0 + 16 * part);
      rb[s] = *(const int4*)(Bp + p * bplane + (size_t)(n0 + row) * Np + k0 + 16 * part);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int q = tid + 256 * s;
      const int p = q >> 9, row = (q >> 2) & 127, part = q & 3;
      *(int4*)&As[buf][p][row * OZ_LDS_ROW + 16 * part] = ra[s];
      *(int4*)&Bs[buf][p][row * OZ_LDS_ROW + 16 * part] = rb[s];
    }
  };
  auto compute = [&](int buf) {
    v4i ar[4], ai[4], br[4], bi[4], nbi[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int off = (wm0 + 16 * i + li) * OZ_LDS_ROW + 16 * lk;
      ar[i] = *(const v4i*)&As[buf][0][off];
      ai[i] = *(const v4i*)&As[buf][1][off];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int off = (wn0 + 16 * j + li) * OZ_LDS_ROW + 16 * lk;
      br[j] = *(const v4i*)&Bs[buf][0][off];
      bi[j] = *(const v4i*)&Bs[buf][1][off];
      nbi[j] = v4i{oz_neg4(bi[j][0]), oz_neg4(bi[j][1]), oz_neg4(bi[j][2]), oz_neg4(bi[j][3])};
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc_re[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ar[i], br[j], acc_re[i][j], 0, 0, 0);
        acc_im[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ar[i], bi[j], acc_im[i][j], 0, 0, 0);
        acc_re[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ai[i], nbi[j], acc_re[i][j], 0, 0, 0);
        acc_im[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ai[i], br[j], acc_im[i][j], 0, 0, 0);
      }
  };
  const int nk = (int)(Np / OZ_BK);
  gload(0);
  sstore(0);
  __syncthreads();
  for (int t = 0; t < nk; ++t) {
    const int buf = t & 1;
    if (t + 1 < nk) gload((int64_t)(t + 1) * OZ_BK);
    compute(buf);
    if (t + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }
  // residues of the exact int32 sums, four consecutive rows packed per store
  const int m = oc.m[l];
  const double rc = oc.rcp[l];
  signed char* Cre = Cr + ((size_t)(b * OZ_L + l) * 2) * bplane;
  signed char* Cim = Cre + bplane;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t col = n0 + wn0 + 16 * j + li;
      const int64_t row = m0 + wm0 + 16 * i + 4 * lk;
      unsigned pr = 0, pi = 0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        pr |= (unsigned)(oz_sres_d((double)acc_re[i][j][v], m, rc) & 0xff) << (8 * v);
        pi |= (unsigned)(oz_sres_d((double)acc_im[i][j][v], m, rc) & 0xff) << (8 * v);
      }
      *(unsigned*)(Cre + (size_t)col * Np + row) = pr;
      *(unsigned*)(Cim + (size_t)col * Np + row) = pi;
    }
}

// Garner's algorithm on 16 symmetric residues -> mixed-radix digits, then
// Horner's rule in FP64: the integer x = d_0 + m_0 (d_1 + m_1 (d_2 + ...)).
__device__ __forceinline__ double oz_crt(const int* x, const OzConst& oc) {
  int d[OZ_L];
  d[0] = x[0];
#pragma unroll
  for (int l = 1; l < OZ_L; ++l) {
    const int m = oc.m[l];
    const float rc = oc.rcpf[l];
    int t = x[l];
#pragma unroll
    for (int p = 0; p < l; ++p) t = oz_sres_i((t - d[p]) * (int)oc.inv[l][p], m, rc);
    d[l] = t;
  }
  double y = (double)d[OZ_L - 1];
#pragma unroll
  for (int l = OZ_L - 2; l >= 0; --l) y = fma(y, (double)oc.m[l], (double)d[l]);
  return y;
}

// R[node] = V - 2^-(e_i + f_c) (T' X') for every listed node: one thread per
// entry (i, c), i < N, c < r.
__global__ __launch_bounds__(256) void k_oz_recon(const signed char* __restrict__ Cr, int64_t N,
                                                  int64_t Np, int r, int rp,
                                                  const int* __restrict__ erow,
                                                  const int* __restrict__ fcol,
                                                  const cd* __restrict__ V, cd* __restrict__ R,
                                                  int64_t sR, const int* __restrict__ bix,
                                                  OzConst oc) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t b = blockIdx.y;
  if (e >= N * (int64_t)r) return;
  const int64_t node = bix ? bix[b] : b;
  const int64_t c = e / N, i = e - c * N;
  const size_t bplane = (size_t)rp * Np;
  const signed char* cre = Cr + ((size_t)(b * OZ_L) * 2) * bplane + (size_t)c * Np + i;
  int xr[OZ_L], xi[OZ_L];
#pragma unroll
  for (int l = 0; l < OZ_L; ++l) {
    xr[l] = cre[(size_t)(2 * l) * bplane];
    xi[l] = cre[(size_t)(2 * l + 1) * bplane];
  }
  const int sh = -(erow[node * N + i] + fcol[b * rp + c]);
  const double pre = ldexp(oz_crt(xr, oc), sh);
  const double pim = ldexp(oz_crt(xi, oc), sh);
  const cd v = V[e];
  R[node * sR + e] = mk(v.re - pre, v.im - pim);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
int64_t qb_oz_pad(int64_t n) { return (n + 127) / 128 * 128; }

size_t qb_oz_window_bytes(int64_t N, int nq) {
  const size_t Np = (size_t)qb_oz_pad(N);
  return (size_t)nq * OZ_L * 2 * Np * Np + (size_t)nq * N * sizeof(int);
}

size_t qb_oz_step_bytes(int64_t N, int r, int nslots) {
  const size_t Np = (size_t)qb_oz_pad(N), rp = (size_t)qb_oz_pad(r);
  return (size_t)nslots * (2 * OZ_L * 2 * rp * Np + rp * sizeof(int));
}

int qb_oz_prepare(qb_ctx* ctx, const cd* T, int64_t N, int nq, signed char* A, int* erow) {
  if (N < 1 || N > 65535) return qb_fail(ctx, QB_E_ARG, "Ozaki residual needs 1 <= N < 65536");
  const int64_t Np = qb_oz_pad(N);
  const OzConst oc = oz_const();
  const int64_t NN = N * N;
  int slot = qb_prof_begin(ctx, QB_K_OTHER, 0.0, (double)nq * (NN * 16.0 + 2.0 * OZ_L * Np * Np));
  hipLaunchKernelGGL(k_oz_rowexp, dim3((unsigned)((N + 255) / 256), (unsigned)nq), dim3(256), 0,
                     ctx->stream, T, N, NN, erow);
  QB_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_oz_slice_T, dim3((unsigned)(Np / 64), (unsigned)(Np / 64), (unsigned)nq),
                     dim3(256), 0, ctx->stream, T, N, Np, NN, erow, A, oc);
  qb_prof_end(ctx, slot);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

int qb_oz_residual(qb_ctx* ctx, const signed char* A, const int* erow, int64_t N, const cd* X,
                   int64_t sX, int r, const cd* V, cd* R, int64_t sR, int nb, signed char* Bw,
                   int* fcol, signed char* Cw) {
  if (nb < 1) return QB_OK;
  if (nb > 65535) return qb_fail(ctx, QB_E_ARG, "Ozaki residual: too many nodes in one launch");
  const int64_t Np = qb_oz_pad(N);
  const int rp = (int)qb_oz_pad(r);
  const OzConst oc = oz_const();
  const int* bix = ctx->bix;
  int s0 = qb_prof_begin(ctx, QB_K_OTHER, 0.0, (double)nb * N * r * 16.0);
  hipLaunchKernelGGL(k_oz_slice_X, dim3((unsigned)rp, (unsigned)nb), dim3(256), 0, ctx->stream,
                     X, N, Np, r, rp, sX, bix, Bw, fcol, oc);
  qb_prof_end(ctx, s0);
  QB_HIP(ctx, hipGetLastError());
  const int ntn = rp / OZ_BN, ntm = (int)(Np / OZ_BM);
  // algorithmic work of the integer products: 4 real products per modulus
  const double ops = 8.0 * OZ_L * (double)Np * rp * Np * nb;
  // algorithmic: the unpadded N x r x N products; issued: as launched (padded)
  int s1 = qb_prof_begin(ctx, QB_K_OZAKI, 8.0 * OZ_L * (double)N * r * N * nb,
                         (double)nb * OZ_L * 2 * ((double)N * N + 2.0 * r * N));
  if (ctx->prof) ctx->prof_issued[QB_K_OZAKI] += ops;
  hipLaunchKernelGGL(k_oz_gemm, dim3((unsigned)(ntn * ntm), OZ_L, (unsigned)nb), dim3(256), 0,
                     ctx->stream, A, Bw, Cw, Np, rp, ntn, bix, oc);
  qb_prof_end(ctx, s1);
  QB_HIP(ctx, hipGetLastError());
  const int64_t ne = N * (int64_t)r;
  int s2 = qb_prof_begin(ctx, QB_K_OTHER, 0.0, (double)nb * ne * 32.0);
  hipLaunchKernelGGL(k_oz_recon, dim3((unsigned)((ne + 255) / 256), (unsigned)nb), dim3(256), 0,
                     ctx->stream, Cw, N, Np, r, rp, erow, fcol, V, R, sR, bix, oc);
  qb_prof_end(ctx, s2);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
