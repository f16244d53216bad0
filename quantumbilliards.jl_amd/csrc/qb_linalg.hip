// Batched dense complex linear algebra for the Beyn contour loop:
// blocked LU with partial pivoting (replaces lu!(T; check=false), i.e.
// LAPACK zgetrf, beyn.jl:270-277), the two triangular solves with the row
// permutation (ldiv!, zgetrs, beyn.jl:281), the FP64 residual of the
// mixed-precision refinement, and the moment sums (axpy!, beyn.jl:282-283).
//
// Every routine works on a batch of matrices with a fixed stride (one matrix
// per contour node), so one launch covers all nodes of a window.  Matrices
// are column-major, complex interleaved.  T = cf (complex64) or cd
// (complex128).
//
// Structure (right-looking, panel width NB = qb_lu_nb(n), 256 or 512):
//   panel (recursive, leaf width 32: single-workgroup column elimination)
//   -> row swaps on the other columns -> unit-lower TRSM of the block row
//   -> trailing update C -= A*B on MFMA.
// The trailing update and every GEMM inside the recursions run in
// k_gemm_sub, which splits complex operands into real/imaginary planes in
// LDS and issues 4 real 16x16x4 MFMAs per complex product block
// (v_mfma_f32_16x16x4_f32 for complex64, v_mfma_f64_16x16x4_f64 for
// complex128) with accumulation in the operand precision.
#include <stdlib.h>

#include <type_traits>
#include <vector>

#include "qb_internal.h"
#include "qb_linalg.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Scal;
template <>
struct Scal<cf> {
  typedef float S;
  typedef floatx4 V4;
};
template <>
struct Scal<cd> {
  typedef double S;
  typedef doublex4 V4;
};

__device__ __forceinline__ floatx4 mfma16(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ doublex4 mfma16(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// GEMM: C[b] -= A[b] * B[b]  (A: M x K, B: K x N, C: M x N)
// ---------------------------------------------------------------------------
// Workgroup tile BM x BN, k-tile BK, 4 waves arranged WGM x WGN; each wave
// computes a (BM/WGM) x (BN/WGN) tile of 16 x 16 MFMA blocks, one
// accumulator block for the real and one for the imaginary part.  MFMA
// 16x16x4 operand layout (one value per lane): A[i][k] with i = lane % 16,
// k = lane / 16; B[k][j] with j = lane % 16, k = lane / 16.  Result
// register v of a lane holds D[4*(lane/16) + v][lane % 16] for the f32
// instruction and D[4*v + lane/16][lane % 16] for the f64 one (checked on
// gfx950).
//
// Pipeline: the next k-tile is loaded from global memory into registers
// while the current one is multiplied out of LDS; the two LDS buffers
// alternate, so each k-tile costs one barrier.  A is stored k-major
// ([k][m], row pad of 16 words puts the four k rows a wave reads in
// different bank quarters), B n-major ([n][k], row stride below) so that
// both the coalesced global loads and the fragment reads are conflict-free.  A is
// negated when it is written to LDS, so the accumulators hold -A*B and the
// epilogue adds them to C.
template <typename T, int BM, int BN, int BK, int WGM, int WGN, bool XCD, bool G3, int PF,
          int OCC>
__global__ void __launch_bounds__(256, OCC) k_gemm_sub(int M, int N, int K, const T* __restrict__ A,
                                                  int64_t lda, int64_t sA,
                                                  const T* __restrict__ B, int64_t ldb,
                                                  int64_t sB, T* __restrict__ C, int64_t ldc,
                                                  int64_t sC, const int* __restrict__ bix) {
  typedef typename Scal<T>::S S;
  typedef typename Scal<T>::V4 V4;
  static_assert(WGM * WGN == 4, "4 waves");
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int APAD = (sizeof(S) == 4) ? 16 : 8;
  constexpr int AST = BM + APAD;       // A row stride (elements)
  // B row stride: the f32 fragment read (li * BST + lk, li < 16, lk < 4)
  // maps the 64 lanes of a wave onto 64 distinct 4-byte banks when BST is 4
  // times an odd number (BK + 4 for BK = 16, 32); BK + 1 wraps lane 15 onto
  // lane 0.
  constexpr int BST = (sizeof(S) == 4) ? BK + 4 : BK + 1;
  constexpr int NA = BM * BK / 256;    // complex elements of A per thread per tile
  constexpr int NB = BN * BK / 256;
  static_assert(NA * 256 == BM * BK && NB * 256 == BN * BK, "tile/thread mismatch");
  __shared__ S As[2][2][BK * AST];     // [buf][re/im]
  __shared__ S Bs[2][2][BN * BST];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave % WGM) * WM, wn0 = (wave / WGM) * WN;
  int mt = blockIdx.x, ntile = blockIdx.y;
  if constexpr (XCD) {
    // Workgroups are dispatched round-robin over the 8 XCDs (each with its
    // own L2).  Give XCD x a contiguous range of tile ids, and order tile
    // ids in groups of 8 m-tiles sweeping n, so the workgroups resident on
    // one XCD share A and B panels in its L2.  Only when the tile count is a
    // multiple of 8 (every batch entry then starts on XCD 0).
    const unsigned gx = gridDim.x, gy = gridDim.y, ntiles = gx * gy;
    if ((ntiles & 7u) == 0) {
      const unsigned L = blockIdx.x + gx * blockIdx.y;
      const unsigned t = (L & 7u) * (ntiles >> 3) + (L >> 3);
      const unsigned grp = t / (8u * gy), first = grp * 8u;
      const unsigned gm = gx - first < 8u ? gx - first : 8u;
      const unsigned loc = t - first * gy;
      mt = (int)(first + loc % gm);
      ntile = (int)(loc / gm);
    }
  }
  const int m0 = mt * BM, n0 = ntile * BN;
  const int64_t b = bix ? bix[blockIdx.z] : blockIdx.z;
  A += b * sA;
  B += b * sB;
  C += b * sC;
  const int li = lane & 15, lk = lane >> 4;

  // acc_re / acc_im hold -Re(AB) / -Im(AB); with G3 (Gauss's three-product
  // form) they hold P1 = (-Ar) Br and P2 = (-Ai) Bi, and acc_s holds
  // P3 = -(Ar + Ai)(Br + Bi): -Re = P1 - P2, -Im = P3 - P1 - P2.
  V4 acc_re[FM][FN], acc_im[FM][FN], acc_s[G3 ? FM : 1][G3 ? FN : 1];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      acc_re[i][j] = V4{0, 0, 0, 0};
      acc_im[i][j] = V4{0, 0, 0, 0};
      if constexpr (G3) acc_s[i][j] = V4{0, 0, 0, 0};
    }

  // PF register slots stage k-tiles on their way from global memory to LDS:
  // with PF = 2 the loads of k-tile t+2 are issued while k-tile t is
  // multiplied, so they have two k-tiles of MFMA work to arrive in.
  T ra[PF][NA], rb[PF][NB];
  // tiles fully inside M x N (all but the edge tiles) load k-tiles that are
  // fully inside K without per-element bounds checks
  const bool full_mn = (m0 + BM <= M) && (n0 + BN <= N);
  auto gload = [&](auto slot, int k0) {
    constexpr int sl = decltype(slot)::value;
    if (full_mn && k0 + BK <= K) {
#pragma unroll
      for (int p = 0; p < NA; ++p) {
        int e = tid + 256 * p;
        int m = e % BM, kk = e / BM;
        ra[sl][p] = A[(m0 + m) + (int64_t)(k0 + kk) * lda];
      }
#pragma unroll
      for (int p = 0; p < NB; ++p) {
        int e = tid + 256 * p;
        int kk = e % BK, n = e / BK;
        rb[sl][p] = B[(k0 + kk) + (int64_t)(n0 + n) * ldb];
      }
      return;
    }
#pragma unroll
    for (int p = 0; p < NA; ++p) {
      int e = tid + 256 * p;
      int m = e % BM, kk = e / BM;
      int gm = m0 + m, gk = k0 + kk;
      ra[sl][p] = (gm < M && gk < K) ? A[gm + (int64_t)gk * lda] : T{0, 0};
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      int e = tid + 256 * p;
      int kk = e % BK, n = e / BK;
      int gn = n0 + n, gk = k0 + kk;
      rb[sl][p] = (gn < N && gk < K) ? B[gk + (int64_t)gn * ldb] : T{0, 0};
    }
  };
  auto sstore = [&](auto slot, int buf) {
    constexpr int sl = decltype(slot)::value;
#pragma unroll
    for (int p = 0; p < NA; ++p) {
      int e = tid + 256 * p;
      int m = e % BM, kk = e / BM;
      As[buf][0][kk * AST + m] = -ra[sl][p].re;
      As[buf][1][kk * AST + m] = -ra[sl][p].im;
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      int e = tid + 256 * p;
      int kk = e % BK, n = e / BK;
      Bs[buf][0][n * BST + kk] = rb[sl][p].re;
      Bs[buf][1][n * BST + kk] = rb[sl][p].im;
    }
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, PF - 1>;

  auto compute = [&](int buf) {
    const S* Ar = As[buf][0];
    const S* Ai = As[buf][1];
    const S* Br = Bs[buf][0];
    const S* Bi = Bs[buf][1];
#pragma unroll
    for (int k4 = 0; k4 < BK; k4 += 4) {
      S ar[FM], ai[FM], nai[FM], br[FN], bi[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        int off = (k4 + lk) * AST + wm0 + i * 16 + li;
        ar[i] = Ar[off];
        ai[i] = Ai[off];
        nai[i] = -ai[i];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        int off = (wn0 + j * 16 + li) * BST + k4 + lk;
        br[j] = Br[off];
        bi[j] = Bi[off];
      }
      if constexpr (G3) {
        S as[FM], bs[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) as[i] = ar[i] + ai[i];
#pragma unroll
        for (int j = 0; j < FN; ++j) bs[j] = br[j] + bi[j];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            acc_re[i][j] = mfma16(ar[i], br[j], acc_re[i][j]);
            acc_im[i][j] = mfma16(ai[i], bi[j], acc_im[i][j]);
            acc_s[i][j] = mfma16(as[i], bs[j], acc_s[i][j]);
          }
      } else {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            acc_re[i][j] = mfma16(ar[i], br[j], acc_re[i][j]);
            acc_re[i][j] = mfma16(nai[i], bi[j], acc_re[i][j]);
            acc_im[i][j] = mfma16(ar[i], bi[j], acc_im[i][j]);
            acc_im[i][j] = mfma16(ai[i], br[j], acc_im[i][j]);
          }
      }
    }
  };

  const int nt = (K + BK - 1) / BK;
  if constexpr (PF == 1) {
    gload(S0{}, 0);
    sstore(S0{}, 0);
    __syncthreads();
    for (int t = 0; t < nt; ++t) {
      const int buf = t & 1;
      if (t + 1 < nt) gload(S0{}, (t + 1) * BK);
      compute(buf);
      if (t + 1 < nt) sstore(S0{}, buf ^ 1);
      __syncthreads();
    }
  } else {
    // k-tile t sits in LDS buffer t & 1 and came through slot t & 1; when it
    // is multiplied, slot t & 1 is free for k-tile t+2 and slot (t+1) & 1
    // holds k-tile t+1, stored to the other buffer at the end of the step.
    gload(S0{}, 0);
    if (nt > 1) gload(S1{}, BK);
    sstore(S0{}, 0);
    __syncthreads();
    auto step = [&](auto P, int t) {
      constexpr int p = decltype(P)::value;
      if (t + 2 < nt) gload(std::integral_constant<int, p>{}, (t + 2) * BK);
      compute(p);
      if (t + 1 < nt) sstore(std::integral_constant<int, 1 - p>{}, 1 - p);
      __syncthreads();
    };
    for (int t = 0; t < nt; t += 2) {
      step(S0{}, t);
      if (t + 1 < nt) step(S1{}, t + 1);
    }
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      int gn = n0 + wn0 + j * 16 + li;
      if (gn >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int gm = m0 + wm0 + i * 16 + (sizeof(S) == 8 ? 4 * v + lk : lk * 4 + v);
        if (gm < M) {
          T* p = C + gm + (int64_t)gn * ldc;
          T c = *p;
          if constexpr (G3) {
            const S p1 = acc_re[i][j][v], p2 = acc_im[i][j][v];
            c.re += p1 - p2;
            c.im += acc_s[i][j][v] - p1 - p2;
          } else {
            c.re += acc_re[i][j][v];
            c.im += acc_im[i][j][v];
          }
          *p = c;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// C[b] -= A[b] * B[b] (complex128 on FP64 MFMA, complex64 on FP32 MFMA) with
// the operands streamed from global memory straight into the MFMA operand
// registers: no LDS and no barrier.  Each wave owns a WM x WN tile of C
// (WGM x WGN waves per workgroup share their A rows and B columns through
// L1) and keeps the fragments of the next D k-steps (4-deep each) in flight
// in a register ring.  Lane l holds A[m0 + 16i + l % 16][k + l / 16] and
// B[k + l / 16][n0 + 16j + l % 16]: 16 lanes read one contiguous run of an A
// column, 4 lanes one of a B column.  Gauss's three products as in
// k_gemm_sub (P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi)), accumulated
// as +AB and subtracted in the epilogue, the same k order and MFMA as
// k_gemm_sub, so the two give the same bits.  XCD: the tile order of
// k_gemm_sub, with GRP m-tiles per group (2-4 keep the in-flight A and B
// panels of an XCD within its L2).
// ---------------------------------------------------------------------------
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

typedef double qb_d2 __attribute__((ext_vector_type(2)));
typedef float qb_f2 __attribute__((ext_vector_type(2)));
template <typename T>
struct RegVec;
template <>
struct RegVec<cd> {
  typedef qb_d2 V2;
};
template <>
struct RegVec<cf> {
  typedef qb_f2 V2;
};

template <typename T, int WM, int WN, int WGM, int WGN, int D, int OCC, bool XCD, int GRP = 8>
__global__ void __launch_bounds__(64 * WGM * WGN, OCC)
    k_gemm_reg(int M, int N, int K, const T* __restrict__ A, int64_t lda, int64_t sA,
               const T* __restrict__ B, int64_t ldb, int64_t sB, T* __restrict__ C,
               int64_t ldc, int64_t sC, const int* __restrict__ bix) {
  typedef typename Scal<T>::S S;
  typedef typename Scal<T>::V4 V4;
  typedef typename RegVec<T>::V2 V2;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int BM = WM * WGM, BN = WN * WGN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  int mt = blockIdx.x, ntile = blockIdx.y;
  if constexpr (XCD) {
    // the XCD-aware tile order of k_gemm_sub, in groups of GRP m-tiles
    const unsigned gx = gridDim.x, gy = gridDim.y, ntiles = gx * gy;
    if ((ntiles & 7u) == 0) {
      const unsigned L = blockIdx.x + gx * blockIdx.y;
      const unsigned t = (L & 7u) * (ntiles >> 3) + (L >> 3);
      const unsigned grp = t / ((unsigned)GRP * gy), first = grp * (unsigned)GRP;
      const unsigned gm = gx - first < (unsigned)GRP ? gx - first : (unsigned)GRP;
      const unsigned loc = t - first * gy;
      mt = (int)(first + loc % gm);
      ntile = (int)(loc / gm);
    }
  }
  const int m0 = mt * BM + (wave % WGM) * WM, n0 = ntile * BN + (wave / WGM) * WN;
  if (m0 >= M || n0 >= N) return;  // the wave has no rows or columns (no barrier follows)
  const int64_t b = bix ? bix[blockIdx.z] : blockIdx.z;
  A += b * sA;
  B += b * sB;
  C += b * sC;
  // rows, columns and k outside the operands are read at a clamped address
  // (so every load is in bounds and unconditional) and zeroed when used
  const T* Ap[FM];
  const T* Bp[FN];
  bool aok[FM], bok[FN];
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int row = m0 + 16 * i + li;
    aok[i] = row < M;
    Ap[i] = A + (aok[i] ? row : M - 1);
  }
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    const int col = n0 + 16 * j + li;
    bok[j] = col < N;
    Bp[j] = B + (int64_t)(bok[j] ? col : N - 1) * ldb;
  }
  const int ns = (K + 3) >> 2;  // k-steps of 4
  V2 ra[D][FM], rb[D][FN];  // (re, im) of the lane's A and B elements, as loaded
  auto load = [&](auto slot, int s) {
    constexpr int d = decltype(slot)::value;
    const int k = 4 * s + lk < K ? 4 * s + lk : K - 1;
#pragma unroll
    for (int i = 0; i < FM; ++i) ra[d][i] = *(const V2*)(Ap[i] + (int64_t)k * lda);
#pragma unroll
    for (int j = 0; j < FN; ++j) rb[d][j] = *(const V2*)(Bp[j] + k);
  };
  V4 p1[FM][FN], p2[FM][FN], p3[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      p1[i][j] = V4{0, 0, 0, 0};
      p2[i][j] = V4{0, 0, 0, 0};
      p3[i][j] = V4{0, 0, 0, 0};
    }
  // k-step s: multiply ring slot s % D, masking rows, columns and k outside
  // the operands
  // (MASK: rows, columns or k of the wave's tile fall outside the operands;
  // interior tiles with K a multiple of 4 multiply the loads as they are)
  auto mult = [&](auto slot, int s, auto mask) {
    constexpr int d = decltype(slot)::value;
    constexpr bool MASK = decltype(mask)::value;
    const bool kok = 4 * s + lk < K;
    S xr[FM], xi[FM], yr[FN], yi[FN], as[FM], bs[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const bool ok = !MASK || (aok[i] && kok);
      xr[i] = ok ? ra[d][i].x : S(0);
      xi[i] = ok ? ra[d][i].y : S(0);
      as[i] = xr[i] + xi[i];
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const bool ok = !MASK || (bok[j] && kok);
      yr[j] = ok ? rb[d][j].x : S(0);
      yi[j] = ok ? rb[d][j].y : S(0);
      bs[j] = yr[j] + yi[j];
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        p1[i][j] = mfma16(xr[i], yr[j], p1[i][j]);
        p2[i][j] = mfma16(xi[i], yi[j], p2[i][j]);
        p3[i][j] = mfma16(as[i], bs[j], p3[i][j]);
      }
  };
  // the ring is refilled without conditions (clamped addresses are always
  // valid), so the loop body is straight-line code and the compiler waits
  // only for the slot it is about to multiply
  static_for<0, D>([&](auto slot) { load(slot, decltype(slot)::value); });
  int s0 = 0;
  auto rounds = [&](auto mask) {
    for (; s0 + D <= ns; s0 += D) {
      static_for<0, D>([&](auto slot) {
        mult(slot, s0 + decltype(slot)::value, mask);
        load(slot, s0 + decltype(slot)::value + D);
      });
    }
  };
  const bool interior = m0 + WM <= M && n0 + WN <= N && (K & 3) == 0;  // uniform per wave
  if (interior)
    rounds(std::false_type{});
  else
    rounds(std::true_type{});
  static_for<0, D>([&](auto slot) {
    if (s0 + decltype(slot)::value < ns) mult(slot, s0 + decltype(slot)::value, std::true_type{});
  });
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int gn = n0 + 16 * j + li;
      if (gn >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // result register v of a lane: row 4v + lane/16 (f64), 4(lane/16) + v (f32)
        const int gm = m0 + 16 * i + (sizeof(S) == 8 ? 4 * v + lk : 4 * lk + v);
        if (gm < M) {
          T* p = C + gm + (int64_t)gn * ldc;
          T c = *p;
          const S q1 = p1[i][j][v], q2 = p2[i][j][v];
          c.re -= q1 - q2;
          c.im -= p3[i][j][v] - q1 - q2;
          *p = c;
        }
      }
    }
}

template <int WM, int WN, int WGM, int WGN, int D, int OCC, bool XCD, int GRP = 8, typename T = cd>
static void launch_gemm_reg(qb_ctx* ctx, int M, int N, int K, const T* A, int64_t lda,
                            int64_t sA, const T* B, int64_t ldb, int64_t sB, T* C,
                            int64_t ldc, int64_t sC, int batch) {
  constexpr int BM = WM * WGM, BN = WN * WGN;
  dim3 g((M + BM - 1) / BM, (N + BN - 1) / BN, batch);
  hipLaunchKernelGGL((k_gemm_reg<T, WM, WN, WGM, WGN, D, OCC, XCD, GRP>), g, dim3(64 * WGM * WGN), 0,
                     ctx->stream, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, ctx->bix);
}

// OCC: waves per SIMD the register allocation must allow (__launch_bounds__).
template <typename T, int BM, int BN, int BK, int WGM, int WGN, bool XCD = false,
          bool G3 = false, int PF = 1, int OCC = (PF == 2 ? 2 : 1)>
static void launch_gemm(qb_ctx* ctx, int M, int N, int K, const T* A, int64_t lda, int64_t sA,
                        const T* B, int64_t ldb, int64_t sB, T* C, int64_t ldc, int64_t sC,
                        int batch) {
  dim3 g((M + BM - 1) / BM, (N + BN - 1) / BN, batch);
  hipLaunchKernelGGL((k_gemm_sub<T, BM, BN, BK, WGM, WGN, XCD, G3, PF, OCC>), g, dim3(256), 0,
                     ctx->stream, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, ctx->bix);
}

// ---------------------------------------------------------------------------
// complex64 C[b] -= A[b] * B[b] on the bf16 matrix cores
// ---------------------------------------------------------------------------
// Each FP32 operand x is split into three bf16 parts, x = h + m + l with
// h = bf16(x), m = bf16(x - h), l = bf16(x - h - m) (round to nearest; both
// differences are exact in FP32), which keeps 24 significant bits.  A real
// product a b is then the six products h h + h m + m h + h l + m m + l h with
// FP32 accumulation (the dropped m l, l m, l l are below 2^-24 |a b|): the
// accuracy of an FP32 product at 6 bf16 MFMAs (2.5 PFLOP/s dense) instead of
// one FP32 one (157.3 TFLOP/s), i.e. up to 2.7x the FP32 rate.  A complex
// block product is 4 real ones (Re = Ar Br + Ai (-Bi), Im = Ar Bi + Ai Br):
// 24 v_mfma_f32_16x16x32_bf16 per 16 x 16 block and 32-deep k-tile.
//
// Tiles: workgroup BM x BN, k-tile 32, 4 waves WGM x WGN; operands staged in
// LDS as FP32 planes [row][k] (k contiguous, row stride 36 floats: the
// 32-byte fragment reads of 16 lanes cover the 64 banks once) and split into
// bf16 in registers when the fragments are read.  Operand lanes: row / column
// lane % 16, k = 8 (lane / 16) + 0..7 on both sides; result register v of a
// lane holds D[4 (lane / 16) + v][lane % 16].
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split3_bf16(const float* x, bf16x8& h, bf16x8& m, bf16x8& l) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const __bf16 a = (__bf16)x[e];
    const float r1 = x[e] - (float)a;
    const __bf16 b = (__bf16)r1;
    const float r2 = r1 - (float)b;
    h[e] = a;
    m[e] = b;
    l[e] = (__bf16)r2;
  }
}

__device__ __forceinline__ bf16x8 neg_bf16(bf16x8 v) {
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = -v[e];
  return v;
}

__device__ __forceinline__ floatx4 mfma_bf6(const bf16x8* a, const bf16x8* b, floatx4 c) {
  // a[0..2], b[0..2] = h, m, l: the six products down to 2^-24 (small ones first)
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], c, 0, 0, 0);
  return c;
}

template <int BM, int BN, int WGM, int WGN, bool XCD>
__global__ void __launch_bounds__(256, 1) k_gemm_bf6(int M, int N, int K, const cf* __restrict__ A,
                                                     int64_t lda, int64_t sA,
                                                     const cf* __restrict__ B, int64_t ldb,
                                                     int64_t sB, cf* __restrict__ C, int64_t ldc,
                                                     int64_t sC, const int* __restrict__ bix) {
  static_assert(WGM * WGN == 4, "4 waves");
  constexpr int BK = 32, ST = BK + 4;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int FM = WM / 16, FN = WN / 16;
  constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
  __shared__ __align__(16) float As[2][2][BM * ST];  // [buf][re/im][m][k]
  __shared__ __align__(16) float Bs[2][2][BN * ST];  // [buf][re/im][n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave % WGM) * WM, wn0 = (wave / WGM) * WN;
  int mt = blockIdx.x, ntile = blockIdx.y;
  if constexpr (XCD) {
    // the XCD-aware tile order of k_gemm_sub
    const unsigned gx = gridDim.x, gy = gridDim.y, ntiles = gx * gy;
    if ((ntiles & 7u) == 0) {
      const unsigned L = blockIdx.x + gx * blockIdx.y;
      const unsigned t = (L & 7u) * (ntiles >> 3) + (L >> 3);
      const unsigned grp = t / (8u * gy), first = grp * 8u;
      const unsigned gm = gx - first < 8u ? gx - first : 8u;
      const unsigned loc = t - first * gy;
      mt = (int)(first + loc % gm);
      ntile = (int)(loc / gm);
    }
  }
  const int m0 = mt * BM, n0 = ntile * BN;
  const int64_t b = bix ? bix[blockIdx.z] : blockIdx.z;
  A += b * sA;
  B += b * sB;
  C += b * sC;
  const int li = lane & 15, lg = lane >> 4;
  floatx4 acc_re[FM][FN], acc_im[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      acc_re[i][j] = floatx4{0, 0, 0, 0};
      acc_im[i][j] = floatx4{0, 0, 0, 0};
    }
  cf ra[NA], rb[NB];
  const bool full_mn = (m0 + BM <= M) && (n0 + BN <= N);
  auto gload = [&](int k0) {
    if (full_mn && k0 + BK <= K) {
#pragma unroll
      for (int p = 0; p < NA; ++p) {
        const int e = tid + 256 * p, m = e % BM, kk = e / BM;
        ra[p] = A[(m0 + m) + (int64_t)(k0 + kk) * lda];
      }
#pragma unroll
      for (int p = 0; p < NB; ++p) {
        const int e = tid + 256 * p, kk = e % BK, n = e / BK;
        rb[p] = B[(k0 + kk) + (int64_t)(n0 + n) * ldb];
      }
      return;
    }
#pragma unroll
    for (int p = 0; p < NA; ++p) {
      const int e = tid + 256 * p, m = e % BM, kk = e / BM;
      const int gm = m0 + m, gk = k0 + kk;
      ra[p] = (gm < M && gk < K) ? A[gm + (int64_t)gk * lda] : cf{0, 0};
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      const int e = tid + 256 * p, kk = e % BK, n = e / BK;
      const int gn = n0 + n, gk = k0 + kk;
      rb[p] = (gn < N && gk < K) ? B[gk + (int64_t)gn * ldb] : cf{0, 0};
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int p = 0; p < NA; ++p) {
      const int e = tid + 256 * p, m = e % BM, kk = e / BM;
      As[buf][0][m * ST + kk] = ra[p].re;
      As[buf][1][m * ST + kk] = ra[p].im;
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      const int e = tid + 256 * p, kk = e % BK, n = e / BK;
      Bs[buf][0][n * ST + kk] = rb[p].re;
      Bs[buf][1][n * ST + kk] = rb[p].im;
    }
  };
  auto compute = [&](int buf) {
    bf16x8 ar[FM][3], ai[FM][3];
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int off = (wm0 + 16 * i + li) * ST + 8 * lg;
      float xr[8], xi[8];
      *(float4*)&xr[0] = *(const float4*)&As[buf][0][off];
      *(float4*)&xr[4] = *(const float4*)&As[buf][0][off + 4];
      *(float4*)&xi[0] = *(const float4*)&As[buf][1][off];
      *(float4*)&xi[4] = *(const float4*)&As[buf][1][off + 4];
      split3_bf16(xr, ar[i][0], ar[i][1], ar[i][2]);
      split3_bf16(xi, ai[i][0], ai[i][1], ai[i][2]);
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int off = (wn0 + 16 * j + li) * ST + 8 * lg;
      float yr[8], yi[8];
      *(float4*)&yr[0] = *(const float4*)&Bs[buf][0][off];
      *(float4*)&yr[4] = *(const float4*)&Bs[buf][0][off + 4];
      *(float4*)&yi[0] = *(const float4*)&Bs[buf][1][off];
      *(float4*)&yi[4] = *(const float4*)&Bs[buf][1][off + 4];
      bf16x8 br[3], bi[3], nbi[3];
      split3_bf16(yr, br[0], br[1], br[2]);
      split3_bf16(yi, bi[0], bi[1], bi[2]);
#pragma unroll
      for (int s = 0; s < 3; ++s) nbi[s] = neg_bf16(bi[s]);
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        acc_re[i][j] = mfma_bf6(ar[i], br, acc_re[i][j]);
        acc_re[i][j] = mfma_bf6(ai[i], nbi, acc_re[i][j]);
        acc_im[i][j] = mfma_bf6(ar[i], bi, acc_im[i][j]);
        acc_im[i][j] = mfma_bf6(ai[i], br, acc_im[i][j]);
      }
    }
  };
  const int nt = (K + BK - 1) / BK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int t = 0; t < nt; ++t) {
    const int buf = t & 1;
    if (t + 1 < nt) gload((t + 1) * BK);
    compute(buf);
    if (t + 1 < nt) sstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int gn = n0 + wn0 + j * 16 + li;
      if (gn >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = m0 + wm0 + i * 16 + lg * 4 + v;
        if (gm < M) {
          cf* p = C + gm + (int64_t)gn * ldc;
          cf c = *p;
          c.re -= acc_re[i][j][v];
          c.im -= acc_im[i][j][v];
          *p = c;
        }
      }
    }
}

template <int BM, int BN, int WGM, int WGN, bool XCD = false>
static void launch_gemm_bf6(qb_ctx* ctx, int M, int N, int K, const cf* A, int64_t lda,
                            int64_t sA, const cf* B, int64_t ldb, int64_t sB, cf* C, int64_t ldc,
                            int64_t sC, int batch) {
  dim3 g((M + BM - 1) / BM, (N + BN - 1) / BN, batch);
  hipLaunchKernelGGL((k_gemm_bf6<BM, BN, WGM, WGN, XCD>), g, dim3(256), 0, ctx->stream, M, N, K,
                     A, lda, sA, B, ldb, sB, C, ldc, sC, ctx->bix);
}

// Tile variants and the per-context choice.  Variant 0/1 are the defaults
// (wide / narrow N); the autotuner (qb_gemm_tune) may pick another per shape
// class, measured on the device.
typedef void (*gemm_fn_c64)(qb_ctx*, int, int, int, const cf*, int64_t, int64_t, const cf*,
                            int64_t, int64_t, cf*, int64_t, int64_t, int);
typedef void (*gemm_fn_c128)(qb_ctx*, int, int, int, const cd*, int64_t, int64_t, const cd*,
                             int64_t, int64_t, cd*, int64_t, int64_t, int);
static const gemm_fn_c64 kGemmC64[QB_GEMM_VARIANTS] = {
    launch_gemm<cf, 128, 128, 16, 2, 2>,      launch_gemm<cf, 256, 64, 16, 4, 1>,
    launch_gemm<cf, 64, 256, 16, 1, 4>,       launch_gemm<cf, 128, 64, 16, 2, 2>,
    launch_gemm<cf, 128, 128, 32, 2, 2>,      launch_gemm<cf, 128, 64, 16, 2, 2, true>,
    launch_gemm<cf, 128, 128, 16, 2, 2, true>, launch_gemm<cf, 128, 64, 16, 2, 2, false, true>,
    launch_gemm<cf, 128, 128, 16, 2, 2, false, true>, launch_gemm<cf, 128, 64, 16, 2, 2, true, true>,
    launch_gemm<cf, 64, 64, 16, 2, 2, false, true>, launch_gemm<cf, 64, 64, 16, 2, 2, true, true>,
    launch_gemm<cf, 64, 64, 16, 2, 2>,         launch_gemm_bf6<128, 128, 2, 2>,
    launch_gemm_bf6<128, 128, 2, 2, true>,     launch_gemm_bf6<128, 64, 2, 2>,
    launch_gemm_bf6<64, 64, 2, 2>,
    // 17-22: operands streamed into registers, no LDS (k_gemm_reg)
    launch_gemm_reg<32, 32, 4, 1, 8, 2, true, 2, cf>, launch_gemm_reg<32, 64, 2, 2, 6, 2, true, 2, cf>,
    launch_gemm_reg<32, 32, 2, 2, 8, 2, true, 2, cf>, launch_gemm_reg<64, 32, 2, 2, 6, 2, true, 2, cf>,
    launch_gemm_reg<32, 32, 4, 1, 8, 3, true, 2, cf>, launch_gemm_reg<32, 32, 4, 1, 12, 2, true, 2, cf>};
static const gemm_fn_c128 kGemmC128[QB_GEMM_VARIANTS] = {
    launch_gemm<cd, 128, 64, 16, 2, 2>,       launch_gemm<cd, 128, 32, 16, 4, 1>,
    launch_gemm<cd, 64, 64, 16, 2, 2>,        launch_gemm<cd, 64, 128, 16, 2, 2>,
    launch_gemm<cd, 256, 32, 16, 4, 1>,       launch_gemm<cd, 64, 64, 16, 2, 2, true>,
    launch_gemm<cd, 128, 64, 16, 2, 2, true>, launch_gemm<cd, 64, 64, 16, 2, 2, false, true>,
    launch_gemm<cd, 64, 32, 16, 2, 2, false, true>, launch_gemm<cd, 64, 64, 16, 2, 2, true, false, 2>,
    launch_gemm<cd, 64, 32, 16, 2, 2, false, true, 2>, launch_gemm<cd, 64, 64, 16, 2, 2, false, true, 1, 2>,
    launch_gemm<cd, 64, 64, 16, 2, 2, true, true, 2, 2>,
    // 13: the lean tile of the LU panel beside the trailing update (19 KB of
    // LDS; always taken through qb_ctx::panel_lean, and open to the tuner).  No bf16-split
    // counterpart for complex128: the other slots repeat the default and the
    // tuner skips them (kGemmAvail)
    launch_gemm<cd, 32, 32, 8, 2, 2, false, true>, launch_gemm<cd, 64, 64, 8, 2, 2, true, true, 2, 3>,
    launch_gemm<cd, 64, 64, 8, 2, 2, true, true, 1, 4>, launch_gemm<cd, 64, 64, 8, 2, 2, true, true, 1, 3>,
    // 17-20: operands streamed into registers, no LDS (k_gemm_reg).  Wave
    // tiles of 32 x 64 and 64 x 32 at one wave per SIMD were slower: the
    // compiler shuttles their accumulators between AGPRs and VGPRs
    launch_gemm_reg<32, 32, 4, 1, 7, 2, true>, launch_gemm_reg<32, 32, 4, 1, 8, 2, true>,
    launch_gemm_reg<32, 32, 2, 2, 6, 2, true>,
    // 20: 32 x 128 workgroup tile for updates of few rows (the block-row TRSM)
    launch_gemm_reg<32, 32, 1, 4, 7, 2, true>,
    // 21, 22: the 128 x 32 tile in XCD groups of 2 and 4 m-tiles (less of A in L2)
    launch_gemm_reg<32, 32, 4, 1, 7, 2, true, 2>, launch_gemm_reg<32, 32, 4, 1, 7, 2, true, 4>};
static_assert(QB_GEMM_LEAN_C128 == 13, "the lean complex128 tile is variant 13");
static const char* kGemmNames[2][QB_GEMM_VARIANTS] = {
    {"128x128x16", "256x64x16", "64x256x16", "128x64x16", "128x128x32", "128x64x16/xcd",
     "128x128x16/xcd", "128x64x16/3m", "128x128x16/3m", "128x64x16/3m/xcd", "64x64x16/3m", "64x64x16/3m/xcd", "64x64x16",
     "128x128x32/bf16x6", "128x128x32/bf16x6/xcd", "128x64x32/bf16x6", "64x64x32/bf16x6",
     "128x32x4/3m/xcd2/reg/d8", "64x128x4/3m/xcd2/reg/d6", "64x64x4/3m/xcd2/reg/d8",
     "128x64x4/3m/xcd2/reg/d6", "128x32x4/3m/xcd2/reg/d8/occ3", "128x32x4/3m/xcd2/reg/d12"},
    {"128x64x16", "128x32x16", "64x64x16", "64x128x16", "256x32x16", "64x64x16/xcd",
     "128x64x16/xcd", "64x64x16/3m", "64x32x16/3m", "64x64x16/xcd/pf2", "64x32x16/3m/pf2", "64x64x16/3m/o2",
     "64x64x16/3m/xcd/pf2", "32x32x8/3m/lean", "64x64x8/3m/xcd/pf2/occ3", "64x64x8/3m/xcd/occ4",
     "64x64x8/3m/xcd/occ3", "128x32x4/3m/xcd/reg/d7", "128x32x4/3m/xcd/reg/d8",
     "64x64x4/3m/xcd/reg/d6", "32x128x4/3m/xcd/reg/d7", "128x32x4/3m/xcd2/reg/d7",
     "128x32x4/3m/xcd4/reg/d7"}};
// Variants that exist for the type (the tuner times only these).
static const bool kGemmAvail[2][QB_GEMM_VARIANTS] = {
    {true, true, true, true, true, true, true, true, true, true, true, true, true, true, true, true,
     true, true, true, true, true, true, true},
    {true, true, true, true, true, true, true, true, true, true, true, true, true, true, true,
     true, true, true, true, true, true, true, true}};
// Variants on the bf16 matrix cores (three-way bf16 splits of FP32 operands).
static const bool kGemmBf6[2][QB_GEMM_VARIANTS] = {
    {false, false, false, false, false, false, false, false, false, false, false, false, false,
     true, true, true, true, false, false, false, false, false, false},
    {false}};

// Variants built on Gauss's three-product form (3 real MFMA products per
// complex block product instead of 4).
static const bool kGemmG3[2][QB_GEMM_VARIANTS] = {
    {false, false, false, false, false, false, false, true, true, true, true, true, false, false,
     false, false, false, true, true, true, true, true, true},
    {false, false, false, false, false, false, false, true, true, false, true, true, true, true,
     true, true, true, true, true, true, true, true, true}};

// Register-streamed variants (k_gemm_reg: no LDS, no barriers).
static const bool kGemmReg[2][QB_GEMM_VARIANTS] = {
    {false, false, false, false, false, false, false, false, false, false, false, false, false,
     false, false, false, false, true, true, true, true, true, true},
    {false, false, false, false, false, false, false, false, false, false, false, false, false,
     false, false, false, false, true, true, true, true, true, true}};

// Variants that pack 3-4 waves per SIMD (small tiles or k-tiles, low
// register and LDS use): fastest alone, but they leave a concurrently
// running kernel (the lookahead panel) no room on a CU, so gemm_pick_conc
// excludes them.
static const bool kGemmPacks[2][QB_GEMM_VARIANTS] = {
    {false, false, false, false, false, false, false, false, false, false, true, true, true, false,
     false, false, false, false, false, false, false, true, false},
    {false, false, false, false, false, false, false, false, false, false, false, false, false,
     false, true, true, true, false, false, false, false, false, false}};

int qb_gemm_class(int M, int N, int K) {
  int cm = M < 1024 ? 0 : (M < 4096 ? 1 : 2);
  int cn = N <= 32 ? 0 : N <= 64 ? 1 : N <= 256 ? 2 : N <= 1024 ? 3 : 4;
  int ck = K <= 64 ? 0 : K <= 256 ? 1 : 2;
  return (cm * 5 + cn) * 3 + ck;
}

template <typename T>
static int gemm_default(int M, int N, int K) {
  // the tile a shape gets when the autotuner has not timed its class (the
  // serving loop's windows of many sizes): for complex128 the autotuner's
  // usual picks; complex64 keeps four-product tiles (the three-product form
  // loses about a bit more of FP32 accuracy, which the refinement pays for)
  (void)M;
  if constexpr (sizeof(T) == sizeof(cf))
    return N <= 64 ? 1 : 0;
  else
    return (N <= 32 || K <= 32) ? QB_GEMM_LEAN_C128 : K <= 64 ? 16 : 21;  // 128x32 register tile
}

template <typename T>
static void gemm_launch_variant(qb_ctx* ctx, int v, int M, int N, int K, const T* A, int64_t lda,
                                int64_t sA, const T* B, int64_t ldb, int64_t sB, T* C,
                                int64_t ldc, int64_t sC, int batch) {
  if constexpr (sizeof(T) == sizeof(cf))
    kGemmC64[v](ctx, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch);
  else
    kGemmC128[v](ctx, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch);
}

template <typename T>
int gemm_sub(qb_ctx* ctx, int M, int N, int K, const T* A, int64_t lda, int64_t sA, const T* B,
             int64_t ldb, int64_t sB, T* C, int64_t ldc, int64_t sC, int batch) {
  if (M <= 0 || N <= 0 || K <= 0 || batch <= 0) return QB_OK;
  const int ty = sizeof(T) == sizeof(cf) ? 0 : 1;
  int v = (ctx->gemm_conc ? ctx->gemm_pick_conc : ctx->gemm_pick)[ty][qb_gemm_class(M, N, K)];
  if (ty == 1 && ctx->panel_lean) v = QB_GEMM_LEAN_C128;
  return gemm_sub_v<T>(ctx, v, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch);
}

template <typename T>
int gemm_sub_v(qb_ctx* ctx, int v, int M, int N, int K, const T* A, int64_t lda, int64_t sA,
               const T* B, int64_t ldb, int64_t sB, T* C, int64_t ldc, int64_t sC, int batch) {
  if (M <= 0 || N <= 0 || K <= 0 || batch <= 0) return QB_OK;
  const int ty = sizeof(T) == sizeof(cf) ? 0 : 1;
  if (v < 0 || v >= QB_GEMM_VARIANTS || !kGemmAvail[ty][v]) v = gemm_default<T>(M, N, K);
  const int kind = ty == 0 ? (kGemmBf6[ty][v] ? QB_K_GEMM_C64_BF16 : QB_K_GEMM_C64) : QB_K_GEMM_C128;
  int slot = qb_prof_begin(ctx, kind, 8.0 * M * N * K * (double)batch,
                           (double)sizeof(T) * ((double)M * K + (double)K * N + 2.0 * M * N) * batch);
  if (ctx->prof)
    ctx->prof_issued[kind] +=
        (kGemmBf6[ty][v] ? 48.0 : kGemmG3[ty][v] ? 6.0 : 8.0) * M * N * K * (double)batch;
  gemm_launch_variant<T>(ctx, v, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch);
  qb_prof_end(ctx, slot);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

// Autotuner: time every variant on synthetic operands of the given shape
// (batch capped so the operands stay below ~1.5 GiB) and keep the fastest
// for the shape's class.  Returns the chosen variant.
__global__ void k_fill_bits(double* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 1e-3 * (double)((i * 2654435761ull) % 997) / 997.0;
}

template <typename T>
int qb_gemm_tune(qb_ctx* ctx, int M, int N, int K, int batch, double* best_tflops) {
  const int ty = sizeof(T) == sizeof(cf) ? 0 : 1;
  size_t per = ((size_t)M * K + (size_t)K * N + (size_t)M * N) * sizeof(T);
  int b = batch;
  while (b > 1 && per * b > ((size_t)3 << 29)) --b;
  T *A = nullptr, *B = nullptr, *C = nullptr;
  int rc = qb_dev_alloc(ctx, (void**)&A, (size_t)M * K * b * sizeof(T));
  if (!rc) rc = qb_dev_alloc(ctx, (void**)&B, (size_t)K * N * b * sizeof(T));
  if (!rc) rc = qb_dev_alloc(ctx, (void**)&C, (size_t)M * N * b * sizeof(T));
  int best = gemm_default<T>(M, N, K), best_conc = best;
  double best_ms = 1e300, best_conc_ms = 1e300;
  if (!rc) {
    int64_t na = (int64_t)M * K * b * sizeof(T) / 8, nb = (int64_t)K * N * b * sizeof(T) / 8,
            nc = (int64_t)M * N * b * sizeof(T) / 8;
    hipLaunchKernelGGL(k_fill_bits, dim3((na + 255) / 256), dim3(256), 0, ctx->stream, (double*)A, na);
    hipLaunchKernelGGL(k_fill_bits, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, (double*)B, nb);
    hipLaunchKernelGGL(k_fill_bits, dim3((nc + 255) / 256), dim3(256), 0, ctx->stream, (double*)C, nc);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    // QB_TUNE_SKIP (both operand types), QB_TUNE_SKIP_C64, QB_TUNE_SKIP_C128:
    // comma-separated variant indices the tuner leaves out (a diagnostic for
    // comparing tile families in whole windows)
    std::vector<char> skip(QB_GEMM_VARIANTS, 0);
    for (const char* env : {"QB_TUNE_SKIP", ty == 0 ? "QB_TUNE_SKIP_C64" : "QB_TUNE_SKIP_C128"}) {
      const char* sk = getenv(env);
      for (const char* c = sk; c && *c;) {
        char* end = nullptr;
        const long v = strtol(c, &end, 10);
        if (end == c) {
          ++c;
          continue;
        }
        if (v >= 0 && v < QB_GEMM_VARIANTS) skip[v] = 1;
        c = end;
      }
    }
    for (int v = 0; v < QB_GEMM_VARIANTS; ++v) {
      if (skip[v] || !kGemmAvail[ty][v]) continue;
      gemm_launch_variant<T>(ctx, v, M, N, K, A, M, (int64_t)M * K, B, K, (int64_t)K * N, C, M,
                             (int64_t)M * N, b);
      (void)hipEventRecord(e0, ctx->stream);
      for (int it = 0; it < 3; ++it)
        gemm_launch_variant<T>(ctx, v, M, N, K, A, M, (int64_t)M * K, B, K, (int64_t)K * N, C, M,
                               (int64_t)M * N, b);
      (void)hipEventRecord(e1, ctx->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      if (hipGetLastError() == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
          ms > 0.f) {
        if (ms < best_ms) {
          best_ms = ms;
          best = v;
        }
        if (!kGemmPacks[ty][v] && ms < best_conc_ms) {
          best_conc_ms = ms;
          best_conc = v;
        }
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int cls = qb_gemm_class(M, N, K);
    if (ctx->gemm_pick[ty][cls] != best || ctx->gemm_pick_conc[ty][cls] != best_conc)
      ++ctx->gemm_epoch;
    ctx->gemm_pick[ty][cls] = best;
    ctx->gemm_pick_conc[ty][cls] = best_conc;
  }
  (void)hipStreamSynchronize(ctx->stream);
  qb_dev_free(ctx, A, (size_t)M * K * b * sizeof(T));
  qb_dev_free(ctx, B, (size_t)K * N * b * sizeof(T));
  qb_dev_free(ctx, C, (size_t)M * N * b * sizeof(T));
  if (best_tflops) *best_tflops = best_ms < 1e299 ? 8.0 * M * N * (double)K * b * 3 / (best_ms * 1e-3) / 1e12 : 0;
  return rc ? rc : best;
}

const char* qb_gemm_variant_name(int ty, int v) {
  return (ty >= 0 && ty < 2 && v >= 0 && v < QB_GEMM_VARIANTS) ? kGemmNames[ty][v] : "?";
}

// ---------------------------------------------------------------------------
// Leaf panel factorisation: columns [c0, c0+w) of rows [c0, n), w <= 32.
// One workgroup per matrix, left-looking (Crout) order inside the leaf:
// step c first forms the U part of column c (rows c0..c0+c-1) by forward
// substitution with the unit-lower top block kept in LDS, then updates the
// rest of column c with all previous L columns at once (only loads until the
// single store per row, so a thread keeps many loads in flight), picks the
// pivot on max |re|+|im| (LAPACK's izamax measure), swaps the pivot row into
// place across the w columns and scales the column.  Each step reads c+1
// columns and writes two, against ~2(w-c) read-modify-written columns in the
// right-looking order.  A zero pivot leaves the column unscaled and counts
// in info, the behaviour of zgetf2 behind lu!(check=false).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T cmul_(T a, T b) {
  return T{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// A[i] when ok, else 0, loaded component by component (a conditional
// aggregate load can leave a per-thread temporary in memory, which the
// compiler then places in LDS, 16 bytes per thread of the workgroup).
template <typename T>
__device__ __forceinline__ T ld_or0(const T* p, bool ok) {
  T v{0, 0};
  if (ok) {
    v.re = p->re;
    v.im = p->im;
  }
  return v;
}

template <typename T>
__global__ void __launch_bounds__(1024) k_getf2(T* __restrict__ A, int64_t lda, int64_t sA,
                                                int n, int c0, int w, int* __restrict__ ipiv,
                                                int64_t sP, int* __restrict__ info) {
  typedef typename Scal<T>::S S;
  __shared__ T Lt[32][33];  // Lt[j][k] = L(c0 + j, c0 + k), k < j
  __shared__ T u[32];       // U(c0 + k, col), k < c
  __shared__ T prow[32];
  __shared__ S red_v[16];
  __shared__ int red_i[16];
  __shared__ int s_piv;
  const int64_t b = blockIdx.x;
  A += b * sA;
  ipiv += b * sP;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wv = tid >> 6;
  const T* L0 = A + (int64_t)c0 * lda;  // first panel column
  for (int c = 0; c < w; ++c) {
    const int col = c0 + c;
    T* Ac = A + (int64_t)col * lda;
    // (1) U(c0 + k, col) = L11^{-1} A(c0.., col), k < c, on wave 0
    if (wv == 0) {
      T a = ld_or0(Ac + c0 + lane, lane < c);
      for (int k = 0; k < c; ++k) {
        T uk = T{__shfl(a.re, k), __shfl(a.im, k)};
        if (lane > k && lane < c) {
          T l = Lt[lane][k];
          a.re -= l.re * uk.re - l.im * uk.im;
          a.im -= l.re * uk.im + l.im * uk.re;
        }
      }
      if (lane < c) {
        u[lane] = a;
        Ac[c0 + lane] = a;
      }
    }
    __syncthreads();
    // (2) rows col..n-1: A(r, col) -= sum_k L(r, c0 + k) U(c0 + k, col); pivot search
    S best = -1;
    int bi = n;
    for (int r = col + tid; r < n; r += 2 * nt) {
      const int r2 = r + nt;
      const bool two = r2 < n;
      T a0 = Ac[r];
      T a1 = ld_or0(Ac + r2, two);
      const T* Lr0 = L0 + r;
      int k = 0;
      for (; k + 4 <= c; k += 4) {
        T l[4], m[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          l[q] = Lr0[(int64_t)(k + q) * lda];
          m[q] = ld_or0(Lr0 + (int64_t)(k + q) * lda + nt, two);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          T uk = u[k + q];
          a0.re -= l[q].re * uk.re - l[q].im * uk.im;
          a0.im -= l[q].re * uk.im + l[q].im * uk.re;
          a1.re -= m[q].re * uk.re - m[q].im * uk.im;
          a1.im -= m[q].re * uk.im + m[q].im * uk.re;
        }
      }
      for (; k < c; ++k) {
        T uk = u[k];
        T l = Lr0[(int64_t)k * lda];
        T m = ld_or0(Lr0 + (int64_t)k * lda + nt, two);
        a0.re -= l.re * uk.re - l.im * uk.im;
        a0.im -= l.re * uk.im + l.im * uk.re;
        a1.re -= m.re * uk.re - m.im * uk.im;
        a1.im -= m.re * uk.im + m.im * uk.re;
      }
      Ac[r] = a0;
      S v0 = fabs(a0.re) + fabs(a0.im);
      if (v0 > best) {
        best = v0;
        bi = r;
      }
      if (two) {
        Ac[r2] = a1;
        S v1 = fabs(a1.re) + fabs(a1.im);
        if (v1 > best) {
          best = v1;
          bi = r2;
        }
      }
    }
    for (int off = 32; off >= 1; off >>= 1) {
      S ob = __shfl_xor(best, off);
      int oi = __shfl_xor(bi, off);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if (lane == 0) {
      red_v[wv] = best;
      red_i[wv] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      S bv = red_v[0];
      int bix = red_i[0];
      for (int q = 1; q < (nt + 63) / 64; ++q)
        if (red_v[q] > bv || (red_v[q] == bv && red_i[q] < bix)) {
          bv = red_v[q];
          bix = red_i[q];
        }
      if (bix >= n) bix = col;
      s_piv = bix;
      ipiv[col] = bix;
    }
    __syncthreads();
    const int piv = s_piv;
    // (3) swap rows col and piv across the w panel columns; the pivot row's
    // L part becomes row c of the top block
    for (int cc = tid; cc < w; cc += nt) {
      T* Acc = A + (int64_t)(c0 + cc) * lda;
      // component by component: a struct copy here keeps x in memory (16
      // bytes of LDS per thread once the compiler promotes it)
      T x;
      x.re = Acc[piv].re;
      x.im = Acc[piv].im;
      if (piv != col) {
        const auto yr = Acc[col].re, yi = Acc[col].im;
        Acc[piv].re = yr;
        Acc[piv].im = yi;
        Acc[col].re = x.re;
        Acc[col].im = x.im;
      }
      prow[cc] = x;
      if (cc < c) Lt[c][cc] = x;
    }
    __syncthreads();
    // (4) scale the column below the pivot
    const T p = prow[c];
    const bool nz = (p.re != 0 || p.im != 0);
    if (!nz && tid == 0 && info) atomicCAS(info + b, 0, col + 1);
    if (nz) {
      S den = p.re * p.re + p.im * p.im;
      T inv = T{p.re / den, -p.im / den};
      for (int r = col + 1 + tid; r < n; r += nt) Ac[r] = cmul_(Ac[r], inv);
    }
    __syncthreads();
  }
}

// Row interchanges ipiv[k0..k1) applied in order to columns [cb, ce)
// (LAPACK zlaswp).  One thread per column.
template <typename T>
__global__ void k_laswp(T* __restrict__ A, int64_t lda, int64_t sA, int cb, int ce, int k0, int k1,
                        const int* __restrict__ ipiv, int64_t sP) {
  const int64_t b = blockIdx.y;
  int col = cb + blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ce) return;
  A += b * sA + (int64_t)col * lda;
  ipiv += b * sP;
  for (int s = k0; s < k1; ++s) {
    int p = ipiv[s];
    if (p != s) {  // component by component (no per-thread temporary in memory)
      const auto tr = A[s].re, ti = A[s].im;
      const auto pr = A[p].re, pi = A[p].im;
      A[s].re = pr;
      A[s].im = pi;
      A[p].re = tr;
      A[p].im = ti;
    }
  }
}

// The interchanges ipiv[k0..k1) of columns [cb, ce) when every row they touch
// lies in [k0, n) with n - k0 small (the last steps of the LU): each
// workgroup turns the interchanges into a permutation of those R = n - k0
// rows (in LDS, one thread), then moves TC columns at a time through LDS,
// loading and storing along the rows (coalesced) instead of one lane per
// column.  Same result as k_laswp.
template <typename T>
__global__ void __launch_bounds__(256) k_laswp_tail(T* __restrict__ A, int64_t lda, int64_t sA,
                                                    int cb, int ce, int k0, int k1, int n,
                                                    const int* __restrict__ ipiv, int64_t sP,
                                                    int cpw, int tc) {
  extern __shared__ __align__(16) unsigned char sm_tail[];
  const int R = n - k0, ns = k1 - k0;
  T* tile = (T*)sm_tail;                                   // [tc][R]
  int* perm = (int*)(sm_tail + (size_t)tc * R * sizeof(T));  // [R]
  int* ip = perm + R;                                      // [ns]
  const int64_t b = blockIdx.y;
  A += b * sA;
  ipiv += b * sP;
  for (int r = threadIdx.x; r < R; r += 256) perm[r] = r;
  for (int i = threadIdx.x; i < ns; i += 256) ip[i] = ipiv[k0 + i] - k0;
  __syncthreads();
  __shared__ int any_moved;
  if (threadIdx.x == 0) {
    int moved = 0;
    for (int i = 0; i < ns; ++i) {
      const int j = ip[i];
      if (j != i) {
        const int t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
        moved = 1;
      }
    }
    any_moved = moved;
  }
  __syncthreads();
  if (!any_moved) return;  // uniform: no interchange in this step
  // only the rows that move are loaded and stored (the permutation maps the
  // moved rows onto themselves)
  const int c_begin = cb + blockIdx.x * cpw;
  const int c_end = min(ce, c_begin + cpw);
  for (int c0 = c_begin; c0 < c_end; c0 += tc) {
    const int nc = min(tc, c_end - c0);
    for (int e = threadIdx.x; e < nc * R; e += 256) {
      const int c = e / R, r = e - c * R;
      if (perm[r] != r) tile[c * R + r] = A[(int64_t)(c0 + c) * lda + k0 + r];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nc * R; e += 256) {
      const int c = e / R, r = e - c * R;
      const int q = perm[r];
      if (q != r) A[(int64_t)(c0 + c) * lda + k0 + r] = tile[c * R + q];
    }
    __syncthreads();
  }
}

// Triangular-solve leaf, n <= 32 rows: X = L^{-1} B (unit lower) or
// X = U^{-1} B (upper).  One wave per 64 right-hand-side columns: the 32 x 64
// tile of B is staged through LDS with coalesced loads/stores (each column's
// 32 rows are contiguous), every lane then substitutes down its own column
// with the triangle read from LDS.
template <typename T, bool UNIT_LOWER>
__global__ void __launch_bounds__(64) k_trsm_leaf(int n, int nrhs, const T* __restrict__ L,
                                                  int64_t ldl, int64_t sL, T* __restrict__ B,
                                                  int64_t ldb, int64_t sB,
                                                  const int* __restrict__ bix) {
  __shared__ T Ls[32][33];
  __shared__ T Bt[64][33];
  const int64_t b = bix ? bix[blockIdx.y] : blockIdx.y;
  L += b * sL;
  B += b * sB;
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * 64;
  for (int e = tid; e < 32 * 32; e += 64) {
    int i = e % 32, j = e / 32;
    Ls[i][j] = (i < n && j < n) ? L[i + (int64_t)j * ldl] : T{0, 0};
  }
  // coalesced load: lanes 0..31 read rows of one column, 32..63 the next
  for (int cc = tid >> 5; cc < 64; cc += 2) {
    int i = tid & 31, col = c0 + cc;
    Bt[cc][i] = (i < n && col < nrhs) ? B[i + (int64_t)col * ldb] : T{0, 0};
  }
  __syncthreads();
  if (!UNIT_LOWER && tid < 32) {
    T d = tid < n ? Ls[tid][tid] : T{1, 0};
    auto den = d.re * d.re + d.im * d.im;
    Ls[tid][tid] = T{d.re / den, -d.im / den};
  }
  __syncthreads();
  if (c0 + tid < nrhs) {
    // blocks of 8 rows in registers; rows already final are read back from
    // the lane's own LDS column (conflict-free: stride 33)
    T* x = Bt[tid];
    if (UNIT_LOWER) {
#pragma unroll 1
      for (int ib = 0; ib < 32; ib += 8) {
        T v[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) v[a] = x[ib + a];
#pragma unroll 1
        for (int j = 0; j < ib; ++j) {
          T xj = x[j];
#pragma unroll
          for (int a = 0; a < 8; ++a) {
            T l = Ls[ib + a][j];
            v[a].re -= l.re * xj.re - l.im * xj.im;
            v[a].im -= l.re * xj.im + l.im * xj.re;
          }
        }
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
          for (int c = 0; c < a; ++c) {
            T l = Ls[ib + a][ib + c];
            v[a].re -= l.re * v[c].re - l.im * v[c].im;
            v[a].im -= l.re * v[c].im + l.im * v[c].re;
          }
#pragma unroll
        for (int a = 0; a < 8; ++a) x[ib + a] = v[a];
      }
    } else {
#pragma unroll 1
      for (int ib = 24; ib >= 0; ib -= 8) {
        T v[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) v[a] = x[ib + a];
#pragma unroll 1
        for (int j = ib + 8; j < 32; ++j) {
          T xj = x[j];
#pragma unroll
          for (int a = 0; a < 8; ++a) {
            T u = Ls[ib + a][j];
            v[a].re -= u.re * xj.re - u.im * xj.im;
            v[a].im -= u.re * xj.im + u.im * xj.re;
          }
        }
#pragma unroll
        for (int a = 7; a >= 0; --a) {
#pragma unroll
          for (int c = a + 1; c < 8; ++c) {
            T u = Ls[ib + a][ib + c];
            v[a].re -= u.re * v[c].re - u.im * v[c].im;
            v[a].im -= u.re * v[c].im + u.im * v[c].re;
          }
          T d = Ls[ib + a][ib + a];
          v[a] = T{v[a].re * d.re - v[a].im * d.im, v[a].re * d.im + v[a].im * d.re};
        }
#pragma unroll
        for (int a = 0; a < 8; ++a) x[ib + a] = v[a];
      }
    }
  }
  __syncthreads();
  for (int cc = tid >> 5; cc < 64; cc += 2) {
    int i = tid & 31, col = c0 + cc;
    if (i < n && col < nrhs) B[i + (int64_t)col * ldb] = Bt[cc][i];
  }
}

// The same leaf with W waves per workgroup (W x 64 right-hand-side columns)
// and each lane's column held in registers (up to 256 VGPRs: two waves per
// SIMD): the triangle is staged once per workgroup and read by broadcast, and
// each wave moves its columns between global memory and registers through a
// 16-column LDS stage (coalesced: 32 lanes per column, four rounds).  For
// complex128, 17 + 8.4 W KB of LDS instead of 51 KB per wave, so a CU holds 8
// waves (W = 2) or 6 (W = 1) instead of 3.  Every row takes the operations of
// k_trsm_leaf in the same order.
template <typename T, bool UNIT_LOWER, int W>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2)))
k_trsm_leaf_reg(int n, int nrhs, const T* __restrict__ L, int64_t ldl, int64_t sL,
                T* __restrict__ B, int64_t ldb, int64_t sB, const int* __restrict__ bix) {
  __shared__ T Ls[32][33];
  __shared__ T St[W][16][33];
  const int64_t b = bix ? bix[blockIdx.y] : blockIdx.y;
  L += b * sL;
  B += b * sB;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int c0 = blockIdx.x * (64 * W) + w * 64;  // first column of this wave
  for (int e = tid; e < 32 * 32; e += 64 * W) {
    const int i = e % 32, j = e / 32;
    Ls[i][j] = (i < n && j < n) ? L[i + (int64_t)j * ldl] : T{0, 0};
  }
  T x[32];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int cc = 2 * t + (lane >> 5), i = lane & 31, col = c0 + 16 * q + cc;
      St[w][cc][i] = (i < n && col < nrhs) ? B[i + (int64_t)col * ldb] : T{0, 0};
    }
    __syncthreads();
    if ((lane >> 4) == q) {
#pragma unroll
      for (int i = 0; i < 32; ++i) x[i] = St[w][lane & 15][i];
    }
    __syncthreads();
  }
  if (!UNIT_LOWER && tid < 32) {
    T d = tid < n ? Ls[tid][tid] : T{1, 0};
    auto den = d.re * d.re + d.im * d.im;
    Ls[tid][tid] = T{d.re / den, -d.im / den};
  }
  __syncthreads();
  if (UNIT_LOWER) {
#pragma unroll
    for (int ib = 0; ib < 32; ib += 8) {
      T v[8];
#pragma unroll
      for (int a = 0; a < 8; ++a) v[a] = x[ib + a];
#pragma unroll
      for (int j = 0; j < ib; ++j) {
        const T xj = x[j];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          const T l = Ls[ib + a][j];
          v[a].re -= l.re * xj.re - l.im * xj.im;
          v[a].im -= l.re * xj.im + l.im * xj.re;
        }
      }
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < a; ++c) {
          const T l = Ls[ib + a][ib + c];
          v[a].re -= l.re * v[c].re - l.im * v[c].im;
          v[a].im -= l.re * v[c].im + l.im * v[c].re;
        }
#pragma unroll
      for (int a = 0; a < 8; ++a) x[ib + a] = v[a];
    }
  } else {
#pragma unroll
    for (int ib = 24; ib >= 0; ib -= 8) {
      T v[8];
#pragma unroll
      for (int a = 0; a < 8; ++a) v[a] = x[ib + a];
#pragma unroll
      for (int j = ib + 8; j < 32; ++j) {
        const T xj = x[j];
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          const T u = Ls[ib + a][j];
          v[a].re -= u.re * xj.re - u.im * xj.im;
          v[a].im -= u.re * xj.im + u.im * xj.re;
        }
      }
#pragma unroll
      for (int a = 7; a >= 0; --a) {
#pragma unroll
        for (int c = a + 1; c < 8; ++c) {
          const T u = Ls[ib + a][ib + c];
          v[a].re -= u.re * v[c].re - u.im * v[c].im;
          v[a].im -= u.re * v[c].im + u.im * v[c].re;
        }
        const T d = Ls[ib + a][ib + a];
        v[a] = T{v[a].re * d.re - v[a].im * d.im, v[a].re * d.im + v[a].im * d.re};
      }
#pragma unroll
      for (int a = 0; a < 8; ++a) x[ib + a] = v[a];
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if ((lane >> 4) == q) {
#pragma unroll
      for (int i = 0; i < 32; ++i) St[w][lane & 15][i] = x[i];
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int cc = 2 * t + (lane >> 5), i = lane & 31, col = c0 + 16 * q + cc;
      if (i < n && col < nrhs) B[i + (int64_t)col * ldb] = St[w][cc][i];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
// QB_LASWP_TAIL=0 (read once): the last steps' interchanges one lane per column.
static bool laswp_tail_off() {
  static const bool off = [] {
    const char* e = getenv("QB_LASWP_TAIL");
    return e && e[0] == '0';
  }();
  return off;
}

template <typename T>
static int laswp(qb_ctx* ctx, T* A, int64_t lda, int64_t sA, int cb, int ce, int k0, int k1,
                 const int* ipiv, int64_t sP, int batch) {
  if (ce <= cb || k1 <= k0) return QB_OK;
  dim3 g((ce - cb + 127) / 128, batch);
  hipLaunchKernelGGL(k_laswp<T>, g, dim3(128), 0, ctx->stream, A, lda, sA, cb, ce, k0, k1, ipiv,
                     sP);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

// laswp for interchanges whose rows all lie in [k0, n): the LDS block kernel
// when n - k0 <= 2048 and n >= 4096, else k_laswp.  Measured: the last steps
// at n = 8192 (48 matrices) gain 5 ms per window; at n = 1640 (240 matrices,
// mostly diagonal pivots, which k_laswp skips at no cost) every step would
// take the kernel and the C_5 windows lost 7%.
template <typename T>
static int laswp_rows(qb_ctx* ctx, T* A, int64_t lda, int64_t sA, int cb, int ce, int k0, int k1,
                      int n, const int* ipiv, int64_t sP, int batch) {
  if (ce <= cb || k1 <= k0) return QB_OK;
  const int R = n - k0;
  if (R > 2048 || n < 4096 || laswp_tail_off())
    return laswp(ctx, A, lda, sA, cb, ce, k0, k1, ipiv, sP, batch);
  const int tc = (int)((size_t)48 * 1024 / ((size_t)R * sizeof(T)));
  const int cpw = 256;
  const size_t lds = (size_t)tc * R * sizeof(T) + (size_t)(R + (k1 - k0)) * sizeof(int);
  dim3 g((ce - cb + cpw - 1) / cpw, batch);
  hipLaunchKernelGGL(k_laswp_tail<T>, g, dim3(256), lds, ctx->stream, A, lda, sA, cb, ce, k0, k1,
                     n, ipiv, sP, cpw, tc);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

static int split32(int n) {
  int h = (n / 2 + 31) / 32 * 32;
  return h >= n ? n / 2 : h;
}

// Leaf launch (n <= 32): k_trsm_leaf_reg with two waves per workgroup when
// the launch has at least 8 waves per CU to give (the LU's block-row solves),
// else the one-wave LDS leaf k_trsm_leaf, which is the faster of the two at
// low occupancy (the moment solves: 576 waves, 270 against 274 ms per window).
// QB_TRSM_LEAF=lds (read once) or qb_set_trsm_leaf_lds(1) takes the LDS leaf
// everywhere; qb_set_trsm_leaf_lds(0) the register leaf everywhere (one wave
// per workgroup for the small launches).
static int g_trsm_leaf_lds = -1;
void qb_set_trsm_leaf_lds(int mode) { g_trsm_leaf_lds = mode; }
static int trsm_leaf_mode() {
  static const int env = [] {
    const char* e = getenv("QB_TRSM_LEAF");
    return e && e[0] == 'l' ? 1 : -1;
  }();
  return g_trsm_leaf_lds >= 0 ? g_trsm_leaf_lds : env;
}

template <typename T, bool LOWER>
static int trsm_leaf(qb_ctx* ctx, int n, int nrhs, const T* L, int64_t ldl, int64_t sL, T* B,
                     int64_t ldb, int64_t sB, int batch) {
  int slot = qb_prof_begin(ctx, QB_K_TRSM, 4.0 * n * n * (double)nrhs * batch, 0);
  const int64_t waves = (int64_t)((nrhs + 63) / 64) * batch;
  const int mode = trsm_leaf_mode();
  if (mode == 1 || (mode < 0 && waves < 8 * 256)) {
    hipLaunchKernelGGL((k_trsm_leaf<T, LOWER>), dim3((nrhs + 63) / 64, batch), dim3(64), 0,
                       ctx->stream, n, nrhs, L, ldl, sL, B, ldb, sB, ctx->bix);
  } else if (waves >= 8 * 256) {
    hipLaunchKernelGGL((k_trsm_leaf_reg<T, LOWER, 2>), dim3((nrhs + 127) / 128, batch),
                       dim3(128), 0, ctx->stream, n, nrhs, L, ldl, sL, B, ldb, sB, ctx->bix);
  } else {
    hipLaunchKernelGGL((k_trsm_leaf_reg<T, LOWER, 1>), dim3((nrhs + 63) / 64, batch), dim3(64),
                       0, ctx->stream, n, nrhs, L, ldl, sL, B, ldb, sB, ctx->bix);
  }
  qb_prof_end(ctx, slot);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

// X = L^{-1} B with L the unit-lower n x n block at L (ld ldl).
// Unit-lower triangular-solve leaf (n <= 32 rows) without LDS, for the LU
// panel beside the trailing update: each lane holds its right-hand-side
// column in registers and reads the triangle with wave-uniform loads.  Row i
// subtracts L(i, j) x_j in increasing j, the order of k_trsm_leaf.
template <typename T>
__global__ void __launch_bounds__(64) k_trsm_leaf_lower_reg(int n, int nrhs, const T* __restrict__ L,
                                                           int64_t ldl, int64_t sL,
                                                           T* __restrict__ B, int64_t ldb,
                                                           int64_t sB, const int* __restrict__ bix) {
  typedef typename Scal<T>::S S;
  const int64_t b = bix ? bix[blockIdx.y] : blockIdx.y;
  L += b * sL;
  B += b * sB;
  const int col = blockIdx.x * 64 + threadIdx.x;
  if (col >= nrhs) return;
  T* Bc = B + (int64_t)col * ldb;
  S xr[32], xi[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    xr[i] = i < n ? Bc[i].re : S(0);
    xi[i] = i < n ? Bc[i].im : S(0);
  }
#pragma unroll
  for (int j = 0; j < 31; ++j) {
    if (j + 1 < n) {
#pragma unroll
      for (int i = j + 1; i < 32; ++i) {
        if (i < n) {
          const S lr = L[i + (int64_t)j * ldl].re, li = L[i + (int64_t)j * ldl].im;
          xr[i] -= lr * xr[j] - li * xi[j];
          xi[i] -= lr * xi[j] + li * xr[j];
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 32; ++i)
    if (i < n) {
      Bc[i].re = xr[i];
      Bc[i].im = xi[i];
    }
}

template <typename T>
static int trsm_lower(qb_ctx* ctx, int n, int nrhs, const T* L, int64_t ldl, int64_t sL, T* B,
                      int64_t ldb, int64_t sB, int batch) {
  if (n <= 0 || nrhs <= 0) return QB_OK;
  if (n <= 32 && ctx->panel_lean) {
    dim3 g((nrhs + 63) / 64, batch);
    int slot = qb_prof_begin(ctx, QB_K_TRSM, 4.0 * n * n * (double)nrhs * batch, 0);
    hipLaunchKernelGGL((k_trsm_leaf_lower_reg<T>), g, dim3(64), 0, ctx->stream, n, nrhs, L, ldl,
                       sL, B, ldb, sB, ctx->bix);
    qb_prof_end(ctx, slot);
    QB_HIP(ctx, hipGetLastError());
    return QB_OK;
  }
  if (n <= 32) return trsm_leaf<T, true>(ctx, n, nrhs, L, ldl, sL, B, ldb, sB, batch);
  int n1 = split32(n);
  int rc = trsm_lower(ctx, n1, nrhs, L, ldl, sL, B, ldb, sB, batch);
  if (rc) return rc;
  rc = gemm_sub<T>(ctx, n - n1, nrhs, n1, L + n1, ldl, sL, B, ldb, sB, B + n1, ldb, sB, batch);
  if (rc) return rc;
  return trsm_lower(ctx, n - n1, nrhs, L + n1 + (int64_t)n1 * ldl, ldl, sL, B + n1, ldb, sB,
                    batch);
}

// X = U^{-1} B with U the upper n x n block.
template <typename T>
static int trsm_upper(qb_ctx* ctx, int n, int nrhs, const T* U, int64_t ldu, int64_t sU, T* B,
                      int64_t ldb, int64_t sB, int batch) {
  if (n <= 0 || nrhs <= 0) return QB_OK;
  if (n <= 32) return trsm_leaf<T, false>(ctx, n, nrhs, U, ldu, sU, B, ldb, sB, batch);
  int n1 = split32(n);
  int rc = trsm_upper(ctx, n - n1, nrhs, U + n1 + (int64_t)n1 * ldu, ldu, sU, B + n1, ldb, sB,
                      batch);
  if (rc) return rc;
  rc = gemm_sub<T>(ctx, n1, nrhs, n - n1, U + (int64_t)n1 * ldu, ldu, sU, B + n1, ldb, sB, B,
                   ldb, sB, batch);
  if (rc) return rc;
  return trsm_upper(ctx, n1, nrhs, U, ldu, sU, B, ldb, sB, batch);
}

// Threads of a leaf workgroup that runs beside the trailing update.  Two
// workgroups of the complex128 trailing-update tiles (70 KB of LDS, two waves
// per SIMD at 128 VGPRs) leave a CU 20 KB of LDS and two waves per SIMD: the
// complex128 leaf (18 KB, 94 VGPRs) fits there with 512 threads, and with
// 1024 it waits for the whole trailing update to drain.  QB_PANEL_THREADS (a
// multiple of 64 in [64, 1024], read once) overrides it.
template <typename T>
static int panel_threads_beside() {
  static const int env = [] {
    const char* e = getenv("QB_PANEL_THREADS");
    long v = e ? strtol(e, nullptr, 10) : 0;
    return (v >= 64 && v <= 1024 && v % 64 == 0) ? (int)v : 0;
  }();
  if (env) return env;
  return sizeof(T) == sizeof(cd) ? 512 : 1024;
}

// The lean kernels for a complex128 panel beside the trailing update
// (qb_ctx::panel_lean).  They leave LDS to the trailing tile, which pays only
// when that tile holds LDS: beside a register-streamed tile (no LDS) the
// ordinary panel kernels are faster (1009 ms against 1022 ms of LU on the
// 48 matrices of N = 8192).  QB_PANEL_LEAN (read once): 0 never, 1 always,
// unset: when the trailing update of the step takes an LDS tile.
static int panel_lean_mode() {
  static const int mode = [] {
    const char* e = getenv("QB_PANEL_LEAN");
    return !e || !e[0] ? -1 : e[0] == '0' ? 0 : 1;
  }();
  return mode;
}

template <typename T>
static bool panel_lean_beside(const qb_ctx* ctx, int M, int N, int K) {
  if (sizeof(T) != sizeof(cd)) return false;
  const int mode = panel_lean_mode();
  if (mode >= 0) return mode == 1;
  int v = ctx->gemm_pick_conc[1][qb_gemm_class(M, N, K)];
  if (v < 0 || v >= QB_GEMM_VARIANTS || !kGemmAvail[1][v]) v = gemm_default<T>(M, N, K);
  return !kGemmReg[1][v];
}

// Recursive panel: factor columns [c0, c0+w) (rows c0..n-1) of every matrix,
// leaves with `threads` threads per workgroup.
template <typename T>
static int panel_rec(qb_ctx* ctx, T* A, int64_t lda, int64_t sA, int n, int c0, int w, int* ipiv,
                     int64_t sP, int* info, int batch, int threads = 1024) {
  if (w <= 32) {
    int slot = qb_prof_begin(ctx, QB_K_PANEL, 8.0 * (double)(n - c0) * w * w / 2 * batch, 0);
    hipLaunchKernelGGL(k_getf2<T>, dim3(batch), dim3(threads), 0, ctx->stream, A, lda, sA, n, c0,
                       w, ipiv, sP, info);
    qb_prof_end(ctx, slot);
    QB_HIP(ctx, hipGetLastError());
    return QB_OK;
  }
  int w1 = split32(w);
  int rc = panel_rec(ctx, A, lda, sA, n, c0, w1, ipiv, sP, info, batch, threads);
  if (rc) return rc;
  rc = laswp(ctx, A, lda, sA, c0 + w1, c0 + w, c0, c0 + w1, ipiv, sP, batch);
  if (rc) return rc;
  T* A11 = A + c0 + (int64_t)c0 * lda;
  rc = trsm_lower(ctx, w1, w - w1, A11, lda, sA, A + c0 + (int64_t)(c0 + w1) * lda, lda, sA,
                  batch);
  if (rc) return rc;
  rc = gemm_sub<T>(ctx, n - c0 - w1, w - w1, w1, A + c0 + w1 + (int64_t)c0 * lda, lda, sA,
                   A + c0 + (int64_t)(c0 + w1) * lda, lda, sA,
                   A + c0 + w1 + (int64_t)(c0 + w1) * lda, lda, sA, batch);
  if (rc) return rc;
  rc = panel_rec(ctx, A, lda, sA, n, c0 + w1, w - w1, ipiv, sP, info, batch, threads);
  if (rc) return rc;
  return laswp(ctx, A, lda, sA, c0, c0 + w1, c0 + w1, c0 + w, ipiv, sP, batch);
}

// Panel width of the blocked LU of n x n matrices: 512 from n = 4096 (the
// trailing updates then read and write the trailing matrix half as often;
// measured at n = 8192, 48 matrices: complex64 LU 677 ms at 512 against 713
// at 256, complex128 window 1.94 s against 2.00), 256 below, where the
// factorisation is panel-bound.  QB_LU_NB (a multiple of 32 in [64, 1024],
// read once) overrides it.
int qb_lu_nb(int n) {
  static const int env = [] {
    const char* e = getenv("QB_LU_NB");
    long v = e ? strtol(e, nullptr, 10) : 0;
    return (v >= 64 && v <= 1024 && v % 32 == 0) ? (int)v : 0;
  }();
  if (env) return env;
  return n >= 4096 ? 512 : 256;
}

// QB_LASWP_AUX=0 (read once) keeps every row interchange on the main stream.
static bool laswp_left_aux() {
  static const bool on = [] {
    const char* e = getenv("QB_LASWP_AUX");
    return !(e && e[0] == '0');
  }();
  return on;
}

template <typename T>
int qb_getrf_first_panel(qb_ctx* ctx, T* A, int n, int64_t lda, int64_t sA, int* ipiv,
                         int64_t sP, int* info, int batch) {
  const int NB = qb_lu_nb(n);
  hipStream_t main = ctx->stream, aux = ctx->stream_aux;
  QB_HIP(ctx, hipEventRecord(ctx->ev_main, main));
  QB_HIP(ctx, hipStreamWaitEvent(aux, ctx->ev_main, 0));
  ctx->stream = aux;
  const int rc = panel_rec(ctx, A, lda, sA, n, 0, n < NB ? n : NB, ipiv, sP, info, batch);
  ctx->stream = main;
  if (rc) return rc;
  QB_HIP(ctx, hipEventRecord(ctx->ev_first, aux));
  return QB_OK;
}

template <typename T>
int qb_getrf_batched(qb_ctx* ctx, T* A, int n, int64_t lda, int64_t sA, int* ipiv, int64_t sP,
                     int* info, int batch, bool first_done) {
  // Lookahead: after the block row of step s is solved, the columns of panel
  // s+1 are updated first; panel s+1 is then factored on the auxiliary
  // stream while the rest of the trailing update runs on the main stream.
  // The two touch disjoint columns; the main stream waits for the panel
  // before the row swaps of step s+1.
  const int NB = qb_lu_nb(n);
  hipStream_t main = ctx->stream, aux = ctx->stream_aux;
  auto on_aux = [&](auto&& f) -> int {
    ctx->stream = aux;
    int rc = f();
    ctx->stream = main;
    return rc;
  };
  int rc = QB_OK;
  if (first_done)
    QB_HIP(ctx, hipStreamWaitEvent(main, ctx->ev_first, 0));
  else
    rc = panel_rec(ctx, A, lda, sA, n, 0, n < NB ? n : NB, ipiv, sP, info, batch);
  if (rc) return rc;
  // The interchanges of step s in the columns left of its panel (the L part
  // of earlier panels) run on the auxiliary stream, ahead of the next panel:
  // nothing on the main stream touches those columns any more once the
  // previous trailing update (which reads the last panel's L) is done.
  const bool left_aux = laswp_left_aux();
  for (int k0 = 0; k0 < n; k0 += NB) {
    int kb = n - k0 < NB ? n - k0 : NB;
    if (left_aux && k0 > 0) {
      QB_HIP(ctx, hipEventRecord(ctx->ev_main, main));
      QB_HIP(ctx, hipStreamWaitEvent(aux, ctx->ev_main, 0));
      rc = on_aux([&] {
        return laswp_rows(ctx, A, lda, sA, 0, k0, k0, k0 + kb, n, ipiv, sP, batch);
      });
    } else {
      rc = laswp_rows(ctx, A, lda, sA, 0, k0, k0, k0 + kb, n, ipiv, sP, batch);
    }
    if (rc) return rc;
    rc = laswp_rows(ctx, A, lda, sA, k0 + kb, n, k0, k0 + kb, n, ipiv, sP, batch);
    if (rc) return rc;
    if (k0 + kb >= n) break;
    const int k1 = k0 + kb, m = n - k1;
    const int kb1 = m < NB ? m : NB;
    T* A11 = A + k0 + (int64_t)k0 * lda;
    T* A12 = A + k0 + (int64_t)k1 * lda;
    T* A21 = A + k1 + (int64_t)k0 * lda;
    rc = trsm_lower(ctx, kb, m, A11, lda, sA, A12, lda, sA, batch);
    if (rc) return rc;
    // next panel's columns first
    rc = gemm_sub<T>(ctx, m, kb1, kb, A21, lda, sA, A12, lda, sA, A + k1 + (int64_t)k1 * lda, lda,
                     sA, batch);
    if (rc) return rc;
    QB_HIP(ctx, hipEventRecord(ctx->ev_main, main));
    QB_HIP(ctx, hipStreamWaitEvent(aux, ctx->ev_main, 0));
    rc = on_aux([&] {
      ctx->panel_lean = m > kb1 && panel_lean_beside<T>(ctx, m, m - kb1, kb);
      const int r = panel_rec(ctx, A, lda, sA, n, k1, kb1, ipiv, sP, info, batch,
                              m > kb1 ? panel_threads_beside<T>() : 1024);
      ctx->panel_lean = false;
      return r;
    });
    if (rc) return rc;
    QB_HIP(ctx, hipEventRecord(ctx->ev_aux, aux));
    if (m > kb1) {
      ctx->gemm_conc = true;  // beside the panel on the second stream
      rc = gemm_sub<T>(ctx, m, m - kb1, kb, A21, lda, sA, A12 + (int64_t)kb1 * lda, lda, sA,
                       A + k1 + (int64_t)(k1 + kb1) * lda, lda, sA, batch);
      ctx->gemm_conc = false;
      if (rc) return rc;
    }
    QB_HIP(ctx, hipStreamWaitEvent(main, ctx->ev_aux, 0));
  }
  if (left_aux && n > NB) {  // the last step's interchanges
    QB_HIP(ctx, hipEventRecord(ctx->ev_aux, aux));
    QB_HIP(ctx, hipStreamWaitEvent(main, ctx->ev_aux, 0));
  }
  return QB_OK;
}

void qb_graphs_clear(qb_ctx* ctx) {
  for (auto& g : ctx->graph_cache) (void)hipGraphExecDestroy(g.exec);
  ctx->graph_cache.clear();
}

template <typename T>
int qb_getrf_graph(qb_ctx* ctx, T* A, int n, int64_t lda, int64_t sA, int* ipiv, int64_t sP,
                   int* info, int batch, bool first_done) {
  if (!ctx->graphs || ctx->prof || ctx->bix || first_done)
    return qb_getrf_batched<T>(ctx, A, n, lda, sA, ipiv, sP, info, batch, first_done);
  const int ty = sizeof(T) == sizeof(cf) ? 0 : 1;
  for (auto& g : ctx->graph_cache)
    if (g.ty == ty && g.A == (const void*)A && g.n == n && g.lda == lda && g.sA == sA &&
        g.ipiv == ipiv && g.sP == sP && g.info == info && g.batch == batch &&
        g.epoch == ctx->gemm_epoch) {
      QB_HIP(ctx, hipGraphLaunch(g.exec, ctx->stream));
      return QB_OK;
    }
  QB_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = qb_getrf_batched<T>(ctx, A, n, lda, sA, ipiv, sP, info, batch);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
  if (rc || e != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    if (!rc) rc = qb_fail(ctx, QB_E_HIP, "LU graph capture failed");
    return rc;
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
  if (ctx->graph_cache.size() >= 16) {
    (void)hipGraphExecDestroy(ctx->graph_cache.front().exec);
    ctx->graph_cache.erase(ctx->graph_cache.begin());
  }
  ctx->graph_cache.push_back({ty, A, n, lda, sA, ipiv, sP, info, batch, ctx->gemm_epoch, exec});
  QB_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
  return QB_OK;
}

// B <- U^{-1} L^{-1} B with the factors of qb_getrf_batched; the row
// permutation of zgetrs is applied by the caller (fused into a gather).
template <typename T>
int qb_trsm_lu_batched(qb_ctx* ctx, const T* LU, int n, int64_t lda, int64_t sA, T* B, int nrhs,
                       int64_t ldb, int64_t sB, int batch) {
  int rc = trsm_lower(ctx, n, nrhs, LU, lda, sA, B, ldb, sB, batch);
  if (rc) return rc;
  return trsm_upper(ctx, n, nrhs, LU, lda, sA, B, ldb, sB, batch);
}

// perm[b][i]: the row of the original right-hand side that lands in row i
// after the interchanges ipiv[b][0..n) are applied in order.
__global__ void k_perm_from_ipiv(const int* __restrict__ ipiv, int* __restrict__ perm, int n,
                                 int batch) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const int* ip = ipiv + (int64_t)b * n;
  int* pm = perm + (int64_t)b * n;
  for (int i = 0; i < n; ++i) pm[i] = i;
  for (int s = 0; s < n; ++s) {
    int p = ip[s];
    if (p != s) {
      int t = pm[s];
      pm[s] = pm[p];
      pm[p] = t;
    }
  }
}

// The same with one workgroup per matrix and both arrays in LDS (8 n bytes):
// the interchanges stay sequential, but on LDS instead of global memory.
__global__ void __launch_bounds__(256) k_perm_from_ipiv_lds(const int* __restrict__ ipiv,
                                                            int* __restrict__ perm, int n) {
  extern __shared__ int sm_perm[];
  int* pm = sm_perm;
  int* ip = sm_perm + n;
  const int64_t b = blockIdx.x;
  for (int i = threadIdx.x; i < n; i += 256) {
    pm[i] = i;
    ip[i] = ipiv[b * n + i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int s = 0; s < n; ++s) {
      const int p = ip[s];
      if (p != s) {
        const int t = pm[s];
        pm[s] = pm[p];
        pm[p] = t;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 256) perm[b * n + i] = pm[i];
}

int qb_perm_from_ipiv(qb_ctx* ctx, const int* ipiv, int* perm, int n, int batch) {
  const size_t lds = (size_t)8 * n;
  if (n > 0 && lds <= (size_t)96 * 1024) {
    if (lds > (size_t)64 * 1024)
      (void)hipFuncSetAttribute((const void*)k_perm_from_ipiv_lds,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_perm_from_ipiv_lds, dim3(batch), dim3(256), lds, ctx->stream, ipiv, perm,
                       n);
  } else {
    hipLaunchKernelGGL(k_perm_from_ipiv, dim3((batch + 63) / 64), dim3(64), 0, ctx->stream, ipiv,
                       perm, n, batch);
  }
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

template <typename TO>
__global__ void k_gather_rows(const cd* __restrict__ src, int64_t s_src,
                              const int* __restrict__ perm, TO* __restrict__ dst, int n, int nrhs,
                              const int* __restrict__ bix) {
  const int64_t b = bix ? bix[blockIdx.z] : blockIdx.z;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int c = blockIdx.y;
  if (i >= n) return;
  int pi = perm[b * (int64_t)n + i];
  cd v = src[b * s_src + pi + (int64_t)c * n];
  TO* d = dst + b * (int64_t)n * nrhs + i + (int64_t)c * n;
  if constexpr (sizeof(TO) == sizeof(cf))
    *d = to_cf(v);
  else
    *d = v;
}

int qb_gather_rows_cd(qb_ctx* ctx, const cd* src, int64_t s_src, const int* perm, cd* dst, int n,
                      int nrhs, int batch) {
  dim3 g((n + 255) / 256, nrhs, batch);
  hipLaunchKernelGGL(k_gather_rows<cd>, g, dim3(256), 0, ctx->stream, src, s_src, perm, dst, n,
                     nrhs, ctx->bix);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
int qb_gather_rows_cd2cf(qb_ctx* ctx, const cd* src, int64_t s_src, const int* perm, cf* dst,
                         int n, int nrhs, int batch) {
  dim3 g((n + 255) / 256, nrhs, batch);
  hipLaunchKernelGGL(k_gather_rows<cf>, g, dim3(256), 0, ctx->stream, src, s_src, perm, dst, n,
                     nrhs, ctx->bix);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

// ---------------------------------------------------------------------------
// Elementwise helpers
// ---------------------------------------------------------------------------
__global__ void k_cd_to_cf(const cd* __restrict__ x, cf* __restrict__ y, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) y[i] = to_cf(x[i]);
}
__global__ void k_add_cf_to_cd(const cf* __restrict__ d, cd* __restrict__ x, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) {
    cd v = x[i];
    cf e = d[i];
    v.re += (double)e.re;
    v.im += (double)e.im;
    x[i] = v;
  }
}
// X_b += D_b for the nodes of the batch (slot stride ns).
__global__ void k_add_cf_to_cd_nodes(const cf* __restrict__ d, cd* __restrict__ x, int64_t n,
                                     int64_t ns, const int* __restrict__ bix) {
  const int64_t b = bix ? bix[blockIdx.y] : blockIdx.y;
  d += b * ns;
  x += b * ns;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    cd v = x[i];
    cf e = d[i];
    v.re += (double)e.re;
    v.im += (double)e.im;
    x[i] = v;
  }
}
// Copy the first nv entries of V into every node's slot of R (slot stride ns).
__global__ void k_bcast(const cd* __restrict__ V, cd* __restrict__ R, int64_t nv, int64_t ns,
                        int nb, const int* __restrict__ bix) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < nv * nb; i += st) {
    int64_t z = i / nv, e = i - z * nv;
    int64_t b = bix ? bix[z] : z;
    R[b * ns + e] = V[e];
  }
}
// A0 = sum_j w_j X_j, A1 = sum_j w_j z_j X_j over nq nodes (beyn.jl:282-283).
__global__ void k_moments(const cd* __restrict__ X, int64_t nv, int nq, const cd* __restrict__ w,
                          const cd* __restrict__ wz, cd* __restrict__ A0, cd* __restrict__ A1) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  cd a0 = mk(0, 0), a1 = mk(0, 0);
  for (int j = 0; j < nq; ++j) {
    cd x = X[(int64_t)j * nv + i];
    a0 = a0 + w[j] * x;
    a1 = a1 + wz[j] * x;
  }
  A0[i] = a0;
  A1[i] = a1;
}
// Per-node max |.|_inf of n_per entries (|re|+|im| bound not used: true modulus).
__global__ void k_node_maxabs(const cd* __restrict__ X, int64_t n_per, int64_t stride,
                              double* __restrict__ out, const int* __restrict__ bix) {
  const int64_t b = bix ? bix[blockIdx.y] : blockIdx.y;
  X += b * stride;
  double m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_per;
       i += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, cabsd(X[i]));
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0)
    atomicMax((unsigned long long*)(out + b), (unsigned long long)__double_as_longlong(m));
}
__global__ void k_node_maxabs32(const cf* __restrict__ X, int64_t n_per, double* __restrict__ out) {
  const int64_t b = blockIdx.y;
  X += b * n_per;
  double m = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_per;
       i += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, (double)hypotf(X[i].re, X[i].im));
  for (int off = 32; off >= 1; off >>= 1) m = fmax(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0)
    atomicMax((unsigned long long*)(out + b), (unsigned long long)__double_as_longlong(m));
}
// Per-node infinity norm of the matrix: max_i sum_j |A_ij|.  One thread per row.
__global__ void k_node_rownorm(const cd* __restrict__ A, int64_t n, int64_t sA,
                               double* __restrict__ out) {
  const int64_t b = blockIdx.y;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0;
  if (i < n) {
    const cd* a = A + b * sA + i;
    for (int64_t j = 0; j < n; ++j) s += cabsd(a[j * n]);
  }
  for (int off = 32; off >= 1; off >>= 1) s = fmax(s, __shfl_xor(s, off));
  if ((threadIdx.x & 63) == 0)
    atomicMax((unsigned long long*)(out + b), (unsigned long long)__double_as_longlong(s));
}

static unsigned grid1(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

int qb_cd_to_cf(qb_ctx* ctx, const cd* x, cf* y, int64_t n) {
  hipLaunchKernelGGL(k_cd_to_cf, dim3(grid1(n)), dim3(256), 0, ctx->stream, x, y, n);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
int qb_add_cf_to_cd(qb_ctx* ctx, const cf* d, cd* x, int64_t n) {
  hipLaunchKernelGGL(k_add_cf_to_cd, dim3(grid1(n)), dim3(256), 0, ctx->stream, d, x, n);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
int qb_add_cf_to_cd_nodes(qb_ctx* ctx, const cf* d, cd* x, int64_t n, int64_t ns, int nb) {
  int64_t g = (n + 255) / 256;
  dim3 grid((unsigned)(g > 4096 ? 4096 : g), nb);
  hipLaunchKernelGGL(k_add_cf_to_cd_nodes, grid, dim3(256), 0, ctx->stream, d, x, n, ns, ctx->bix);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
int qb_bcast(qb_ctx* ctx, const cd* V, cd* R, int64_t nv, int64_t ns, int nb) {
  hipLaunchKernelGGL(k_bcast, dim3(grid1(nv * nb)), dim3(256), 0, ctx->stream, V, R, nv, ns, nb,
                     ctx->bix);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
int qb_moments(qb_ctx* ctx, const cd* X, int64_t nv, int nq, const cd* w, const cd* wz, cd* A0,
               cd* A1) {
  hipLaunchKernelGGL(k_moments, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, ctx->stream, X,
                     nv, nq, w, wz, A0, A1);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
int qb_node_maxabs(qb_ctx* ctx, const cd* X, int64_t n_per, int64_t stride, int nb,
                   double* out_dev, int n_out) {
  // out_dev[b] for the physical node b of every batch entry; all n_out
  // entries are cleared first
  QB_HIP(ctx, hipMemsetAsync(out_dev, 0, (n_out > nb ? n_out : nb) * sizeof(double), ctx->stream));
  int64_t g = (n_per + 255) / 256;
  dim3 grid((unsigned)(g > 1024 ? 1024 : g), nb);
  hipLaunchKernelGGL(k_node_maxabs, grid, dim3(256), 0, ctx->stream, X, n_per, stride, out_dev,
                     ctx->bix);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
int qb_node_maxabs32(qb_ctx* ctx, const cf* X, int64_t n_per, int nb, double* out_dev) {
  QB_HIP(ctx, hipMemsetAsync(out_dev, 0, nb * sizeof(double), ctx->stream));
  int64_t g = (n_per + 255) / 256;
  dim3 grid((unsigned)(g > 1024 ? 1024 : g), nb);
  hipLaunchKernelGGL(k_node_maxabs32, grid, dim3(256), 0, ctx->stream, X, n_per, out_dev);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}
int qb_node_rownorm(qb_ctx* ctx, const cd* A, int64_t n, int nb, double* out_dev) {
  QB_HIP(ctx, hipMemsetAsync(out_dev, 0, nb * sizeof(double), ctx->stream));
  dim3 grid((unsigned)((n + 255) / 256), nb);
  hipLaunchKernelGGL(k_node_rownorm, grid, dim3(256), 0, ctx->stream, A, n, n * n, out_dev);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

// explicit instantiations
template int qb_getrf_batched<cf>(qb_ctx*, cf*, int, int64_t, int64_t, int*, int64_t, int*, int,
                                  bool);
template int qb_getrf_batched<cd>(qb_ctx*, cd*, int, int64_t, int64_t, int*, int64_t, int*, int,
                                  bool);
template int qb_getrf_first_panel<cf>(qb_ctx*, cf*, int, int64_t, int64_t, int*, int64_t, int*,
                                      int);
template int qb_getrf_first_panel<cd>(qb_ctx*, cd*, int, int64_t, int64_t, int*, int64_t, int*,
                                      int);
template int qb_getrf_graph<cf>(qb_ctx*, cf*, int, int64_t, int64_t, int*, int64_t, int*, int,
                                bool);
template int qb_getrf_graph<cd>(qb_ctx*, cd*, int, int64_t, int64_t, int*, int64_t, int*, int,
                                bool);
template int qb_trsm_lu_batched<cf>(qb_ctx*, const cf*, int, int64_t, int64_t, cf*, int, int64_t,
                                    int64_t, int);
template int qb_trsm_lu_batched<cd>(qb_ctx*, const cd*, int, int64_t, int64_t, cd*, int, int64_t,
                                    int64_t, int);
template int qb_gemm_tune<cf>(qb_ctx*, int, int, int, int, double*);
template int qb_gemm_tune<cd>(qb_ctx*, int, int, int, int, double*);
template int gemm_sub<cf>(qb_ctx*, int, int, int, const cf*, int64_t, int64_t, const cf*, int64_t,
                          int64_t, cf*, int64_t, int64_t, int);
template int gemm_sub<cd>(qb_ctx*, int, int, int, const cd*, int64_t, int64_t, const cd*, int64_t,
                          int64_t, cd*, int64_t, int64_t, int);
template int gemm_sub_v<cd>(qb_ctx*, int, int, int, int, const cd*, int64_t, int64_t, const cd*, int64_t,
                          int64_t, cd*, int64_t, int64_t, int);
