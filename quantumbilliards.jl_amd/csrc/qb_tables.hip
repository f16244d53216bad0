// K3: device construction of the per-wavenumber Chebyshev tables.
//
// Mirrors plan_h / plan_j (src/chebyshev/chebyshev_bessels.jl:461-504):
// uniform breaks (chebyshev_core.jl:95-103), samples at the M+1 Lobatto
// nodes r_j = ((b+a) + (b-a) cospi(j/M)) / 2 of every panel
// (_build_table_h! / _build_table_j!, bessels:331-372) and the DCT-I fit of
// _chebfit! (chebyshev_core.jl:67-80).  One thread builds one panel of one
// wavenumber for both orders (H0 and H1, or J0 and J1), so the two series of
// a panel are stored interleaved and read together by the assembly.
#include <math.h>
#include <string.h>

#include "qb_bessel.h"
#include "qb_cheb.h"
#include "qb_internal.h"

#define QB_MMAX 16

struct LobCoef {
  int M;
  double t[QB_MMAX + 1];                    // cospi(j/M)
  double C[(QB_MMAX + 1) * (QB_MMAX + 1)];  // C[m*(M+1)+j] = cospi(j*m/M)
};

static double cospi_host(double x) {
  x = fmod(fabs(x), 2.0);
  if (x > 1.0) x = 2.0 - x;
  if (x == 0.0) return 1.0;
  if (x == 1.0) return -1.0;
  if (x == 0.5) return 0.0;
  if (x > 0.5) return -(double)sinl(3.14159265358979323846264338327950288L * ((long double)x - 0.5L));
  return (double)cosl(3.14159265358979323846264338327950288L * (long double)x);
}

void qb_lobatto_host(int M, double* t, double* C) {
  for (int j = 0; j <= M; ++j) t[j] = cospi_host((double)j / (double)M);
  for (int m = 0; m <= M; ++m)
    for (int j = 0; j <= M; ++j) C[m * (M + 1) + j] = cospi_host((double)(j * m) / (double)M);
}

// kind 0: Hankel H0/H1 on [rmin, rmax]; kind 1: J0/J1 on [0, rmax].
template <int KIND>
__global__ void __launch_bounds__(64) k_build_tables(const cd* __restrict__ zk, int64_t nz,
                                                     double rmin, double rmax, int np,
                                                     LobCoef lc, cd* __restrict__ out) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nz * (int64_t)np) return;
  int64_t q = gid / np;
  int p = (int)(gid - q * np);
  const int M = lc.M;
  double h = (rmax - rmin) / (double)np;
  double a = fma((double)p, h, rmin);
  double b = (p == np - 1) ? rmax : fma((double)(p + 1), h, rmin);
  cd k = zk[q];
  cd f0[QB_MMAX + 1], f1[QB_MMAX + 1];
  for (int j = 0; j <= M; ++j) {
    double r = ((b + a) + (b - a) * lc.t[j]) / 2.0;
    cd z = mk(k.re * r, k.im * r);
    if (KIND == 0)
      qb_hankel1_01(z, &f0[j], &f1[j]);
    else
      qb_bessel_j01(z, &f0[j], &f1[j]);
  }
  cd* o = out + (q * (int64_t)np + p) * (int64_t)(2 * (M + 1));
  double s2 = 2.0 / (double)M;
  for (int m = 0; m <= M; ++m) {
    double sg = (m & 1) ? -1.0 : 1.0;
    cd s0 = 0.5 * f0[0];
    cd s1 = 0.5 * f1[0];
    s0 = s0 + 0.5 * (sg * f0[M]);
    s1 = s1 + 0.5 * (sg * f1[M]);
    for (int j = 1; j < M; ++j) {
      double c = lc.C[m * (M + 1) + j];
      s0 = s0 + f0[j] * c;
      s1 = s1 + f1[j] * c;
    }
    cd c0 = s2 * s0, c1 = s2 * s1;
    if (m == 0) {
      c0 = 0.5 * c0;
      c1 = 0.5 * c1;
    }
    o[2 * m] = c0;
    o[2 * m + 1] = c1;
  }
}

static LobCoef make_lob(int M) {
  LobCoef lc{};
  lc.M = M;
  double t[QB_MMAX + 1], C[(QB_MMAX + 1) * (QB_MMAX + 1)];
  qb_lobatto_host(M, t, C);
  for (int j = 0; j <= M; ++j) lc.t[j] = t[j];
  for (int i = 0; i < (M + 1) * (M + 1); ++i) lc.C[i] = C[i];
  return lc;
}

size_t qb_tables_bytes(int64_t nz, const qb_cheb_params& cp, size_t* hbytes) {
  const size_t hb = (size_t)nz * cp.n_panels_h * (cp.M_h + 1) * 2 * sizeof(cd);
  const size_t jb = (size_t)nz * cp.n_panels_j * (cp.M_j + 1) * 2 * sizeof(cd);
  if (hbytes) *hbytes = hb;
  return hb + jb;
}

static int tables_check(qb_ctx* ctx, double rmin_h, double rmax, const qb_cheb_params& cp) {
  if (cp.M_h < 2 || cp.M_h > QB_MMAX || cp.M_j < 2 || cp.M_j > QB_MMAX)
    return qb_fail(ctx, QB_E_ARG, "Chebyshev degree M must be in [2, 16]");
  if (cp.n_panels_h < 1 || cp.n_panels_j < 1)
    return qb_fail(ctx, QB_E_ARG, "n_panels must be >= 1");
  if (!(rmin_h > 0.0) || !(rmax > rmin_h))
    return qb_fail(ctx, QB_E_ARG, "table range needs 0 < rmin < rmax");
  return QB_OK;
}

static int tables_launch(qb_ctx* ctx, const double* z_dev, int64_t nz, double rmin_h,
                         double rmax, const qb_cheb_params& cp, DevTables* t) {
  size_t hb;
  const size_t tb = qb_tables_bytes(nz, cp, &hb);
  LobCoef lh = make_lob(cp.M_h), lj = make_lob(cp.M_j);
  const cd* zk = (const cd*)z_dev;
  int64_t nh = nz * (int64_t)cp.n_panels_h, nj = nz * (int64_t)cp.n_panels_j;
  int slot = qb_prof_begin(ctx, QB_K_TABLES, 0.0, (double)tb);
  hipLaunchKernelGGL(k_build_tables<0>, dim3((unsigned)((nh + 63) / 64)), dim3(64), 0, ctx->stream,
                     zk, nz, rmin_h, rmax, cp.n_panels_h, lh, t->H);
  QB_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(k_build_tables<1>, dim3((unsigned)((nj + 63) / 64)), dim3(64), 0, ctx->stream,
                     zk, nz, 0.0, rmax, cp.n_panels_j, lj, t->J);
  qb_prof_end(ctx, slot);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

static void tables_describe(int64_t nz, double rmin_h, double rmax, const qb_cheb_params& cp,
                            DevTables* t) {
  t->nz = nz;
  t->np_h = cp.n_panels_h;
  t->M_h = cp.M_h;
  t->np_j = cp.n_panels_j;
  t->M_j = cp.M_j;
  t->rmin_h = rmin_h;
  t->rmax = rmax;
}

int qb_tables_build_in(qb_ctx* ctx, const double* z_dev, int64_t nz, double rmin_h, double rmax,
                       const qb_cheb_params& cp, void* mem, size_t mem_bytes, DevTables* t) {
  int rc = tables_check(ctx, rmin_h, rmax, cp);
  if (rc) return rc;
  size_t hb;
  const size_t tb = qb_tables_bytes(nz, cp, &hb);
  if (tb > mem_bytes) return qb_fail(ctx, QB_E_ARG, "table memory too small");
  tables_describe(nz, rmin_h, rmax, cp, t);
  t->H = (cd*)mem;
  t->J = (cd*)((char*)mem + hb);
  // the memory belongs to the caller: these tables are not passed to qb_tables_free
  return tables_launch(ctx, z_dev, nz, rmin_h, rmax, cp, t);
}

int qb_tables_build(qb_ctx* ctx, const double* z_dev, int64_t nz, double rmin_h, double rmax,
                    const qb_cheb_params& cp, bool /*want_h0j0*/, DevTables* t) {
  if (cp.M_h < 2 || cp.M_h > QB_MMAX || cp.M_j < 2 || cp.M_j > QB_MMAX)
    return qb_fail(ctx, QB_E_ARG, "Chebyshev degree M must be in [2, 16]");
  if (cp.n_panels_h < 1 || cp.n_panels_j < 1)
    return qb_fail(ctx, QB_E_ARG, "n_panels must be >= 1");
  if (!(rmin_h > 0.0) || !(rmax > rmin_h))
    return qb_fail(ctx, QB_E_ARG, "table range needs 0 < rmin < rmax");
  t->nz = nz;
  t->np_h = cp.n_panels_h;
  t->M_h = cp.M_h;
  t->np_j = cp.n_panels_j;
  t->M_j = cp.M_j;
  t->rmin_h = rmin_h;
  t->rmax = rmax;
  size_t hb;
  const size_t jb = qb_tables_bytes(nz, cp, &hb) - hb;
  int rc = qb_dev_alloc(ctx, (void**)&t->H, hb);
  if (rc) return rc;
  rc = qb_dev_alloc(ctx, (void**)&t->J, jb);
  if (rc) return rc;
  return tables_launch(ctx, z_dev, nz, rmin_h, rmax, cp, t);
}

void qb_tables_free(qb_ctx* ctx, DevTables* t) {
  size_t hb = (size_t)t->nz * t->np_h * (t->M_h + 1) * 2 * sizeof(cd);
  size_t jb = (size_t)t->nz * t->np_j * (t->M_j + 1) * 2 * sizeof(cd);
  if (t->H) qb_dev_free(ctx, t->H, hb);
  if (t->J) qb_dev_free(ctx, t->J, jb);
  t->H = t->J = nullptr;
}

// Direct evaluation at arbitrary points: out4[4*i + {0,1,2,3}] = H0, H1, J0, J1.
__global__ void k_bessel_eval(const cd* __restrict__ z, int64_t n, cd* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cd zz = z[i], h0, h1, j0, j1;
  qb_bessel_j01(zz, &j0, &j1);
  if (zz.re > 0.0) {
    qb_hankel1_01(zz, &h0, &h1);
  } else {
    h0 = mk(NAN, NAN);
    h1 = mk(NAN, NAN);
  }
  out[4 * i + 0] = h0;
  out[4 * i + 1] = h1;
  out[4 * i + 2] = j0;
  out[4 * i + 3] = j1;
}

int qb_bessel_eval_dev(qb_ctx* ctx, const cd* z, int64_t n, cd* out4) {
  if (n <= 0) return QB_OK;
  hipLaunchKernelGGL(k_bessel_eval, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, ctx->stream,
                     z, n, out4);
  QB_HIP(ctx, hipGetLastError());
  return QB_OK;
}

// Table accuracy check of chebyshev_params (chebyshev_optimal_panelization.jl:41-77):
// max |table - direct| over `samples` uniform radii, for H0, H1 on
// [rmin_h, rmax] and J0, J1 on [0, rmax], over all wavenumbers.  Errors are
// non-negative doubles, so their bit patterns order like unsigned integers
// and a vector atomicMax on the bits gives the maximum.  The 4 wavefronts of
// a block reduce through LDS first, so one atomic per error kind and block
// meets on each address.  NaN errors count as +inf.
__global__ __launch_bounds__(256) void k_cheb_check(const cd* __restrict__ zk, int64_t nz, int samples, double rmin_h,
                             double rmax, int np_h, int M_h, const cd* __restrict__ H, int np_j,
                             int M_j, const cd* __restrict__ J,
                             unsigned long long* __restrict__ err4) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = nz * (int64_t)samples;
  double e[4] = {0, 0, 0, 0};
  if (gid < total) {
    int64_t q = gid / samples;
    int s = (int)(gid - q * samples);
    cd k = zk[q];
    double fr = samples > 1 ? (double)s / (double)(samples - 1) : 0.0;
    // range(a, b; length=n) points
    double rh = rmin_h + (rmax - rmin_h) * fr;
    double rj = rmax * fr;
    int p;
    double t;
    cd f0, f1, d0, d1;
    qb_panel(rh, rmin_h, rmax, np_h, &p, &t);
    qb_clenshaw2_dyn(H + (q * np_h + p) * (int64_t)(2 * (M_h + 1)), M_h, t, &f0, &f1);
    qb_hankel1_01(mk(k.re * rh, k.im * rh), &d0, &d1);
    e[0] = cabsd(f0 - d0);
    e[1] = cabsd(f1 - d1);
    qb_panel(rj, 0.0, rmax, np_j, &p, &t);
    qb_clenshaw2_dyn(J + (q * np_j + p) * (int64_t)(2 * (M_j + 1)), M_j, t, &f0, &f1);
    qb_bessel_j01(mk(k.re * rj, k.im * rj), &d0, &d1);
    e[2] = cabsd(f0 - d0);
    e[3] = cabsd(f1 - d1);
    // a NaN table or direct value must fail the tolerance (the reference's
    // max propagates NaN, optimal_panelization.jl:51); fmax and the bit-order
    // atomic below would drop it, +inf keeps it
    for (int c = 0; c < 4; ++c)
      if (isnan(e[c])) e[c] = INFINITY;
  }
  __shared__ double red[4][4];
  const int w = threadIdx.x >> 6;
  for (int c = 0; c < 4; ++c) {
    double v = e[c];
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) red[w][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int c = threadIdx.x;
    const double v = fmax(fmax(red[0][c], red[1][c]), fmax(red[2][c], red[3][c]));
    atomicMax(&err4[c], (unsigned long long)__double_as_longlong(v));
  }
}

int qb_cheb_check_dev(qb_ctx* ctx, const DevTables& t, const cd* z_dev, int64_t nz,
                      int32_t samples, double* err4_host) {
  unsigned long long* d = nullptr;
  int rc = qb_dev_alloc(ctx, (void**)&d, 4 * sizeof(unsigned long long));
  if (rc) return rc;
  QB_HIP(ctx, hipMemsetAsync(d, 0, 4 * sizeof(unsigned long long), ctx->stream));
  int64_t total = nz * (int64_t)samples;
  // k_cheb_check reduces over exactly 4 wavefronts of 256 threads per block
  hipLaunchKernelGGL(k_cheb_check, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     ctx->stream, z_dev, nz, samples, t.rmin_h, t.rmax, t.np_h, t.M_h, t.H, t.np_j,
                     t.M_j, t.J, d);
  QB_HIP(ctx, hipGetLastError());
  unsigned long long hbits[4];
  QB_HIP(ctx, hipMemcpyAsync(hbits, d, sizeof(hbits), hipMemcpyDeviceToHost, ctx->stream));
  QB_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int c = 0; c < 4; ++c) {
    long long b = (long long)hbits[c];
    double v;
    memcpy(&v, &b, sizeof(v));
    err4_host[c] = v;
  }
  qb_dev_free(ctx, d, 4 * sizeof(unsigned long long));
  return QB_OK;
}
