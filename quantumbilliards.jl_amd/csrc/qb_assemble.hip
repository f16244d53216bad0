// K1 / K2: batched CFIE-Kress and DLP-Kress assembly of T(z) = I - K(z).
//
// Entry formulas follow _all_k_nosymm_CFIE_chebyshev!
// (src/chebyshev/chebyshev_euclidian_helmholtz_cfie_kress.jl:543-665) and
// _construct_dlp_kress_matrices_chebyshev! + the I - D pass
// (chebyshev_euclidian_helmholtz_dlp_kress.jl:702-769, :1086-1095).
//
// Layout: out[q*N*N + i + j*N] (column-major per wavenumber, Julia's layout).
// A 64x4 thread block covers 64 consecutive rows (one wave per column, so
// every store of a wave is one contiguous 1 KiB run) and 16 columns.  The
// geometry of a pair (distance, 1/r, the DLP numerator, the circulant Kress
// R and log term) is computed in registers instead of being read from N x N
// caches as the reference does; grid.z walks the wavenumbers so that the
// blocks resident at one time read the tables of one wavenumber from L2.
#include "qb_bessel.h"
#include "qb_cheb.h"
#include "qb_internal.h"

#define QB_INV_TWO_PI 0.15915494309189533577
#define QB_EULER_OVER_PI (QB_EULER_GAMMA / QB_PI)

struct TabView {
  const cd* H;  // this wavenumber's Hankel table
  const cd* J;  // this wavenumber's J table
  int np_h, M_h, np_j, M_j;
  double rmin_h, rmax;
};

template <int MH, int MJ>
__device__ __forceinline__ void hank(const TabView& tv, double r, cd k, bool cutoff, cd* h0,
                                     cd* h1) {
  // h0_h1_multi_ks_at_r! (bessels:1190-1206): |kr| < 0.2 evaluates directly
  // (the |kr| < 1e-3 series is the same expansion), else interpolates.
  cd z = mk(k.re * r, k.im * r);
  if (cutoff && cabsd(z) < 0.2) {
    qb_hankel1_01(z, h0, h1);
    return;
  }
  int p;
  double t;
  qb_panel(r, tv.rmin_h, tv.rmax, tv.np_h, &p, &t);
  const cd* c = tv.H + (int64_t)p * (2 * (tv.M_h + 1));
  if (MH > 0)
    qb_clenshaw2<MH>(c, t, h0, h1);
  else
    qb_clenshaw2_dyn(c, tv.M_h, t, h0, h1);
}

template <int MH, int MJ>
__device__ __forceinline__ void bessj(const TabView& tv, double r, cd* j0, cd* j1) {
  int p;
  double t;
  qb_panel(r, 0.0, tv.rmax, tv.np_j, &p, &t);
  const cd* c = tv.J + (int64_t)p * (2 * (tv.M_j + 1));
  if (MJ > 0)
    qb_clenshaw2<MJ>(c, t, j0, j1);
  else
    qb_clenshaw2_dyn(c, tv.M_j, t, j0, j1);
}

// CFIE entry A(z)[i,j] (target i, source j).
template <int MH, int MJ>
__device__ cd cfie_entry(const DevGeom& g, const TabView& tv, int64_t i, int64_t j, cd k) {
  const cd ik = mk(-k.im, k.re);
  int ci = g.comp[i], cj = g.comp[j];
  double sj = g.speed[j], wj = g.ws[j];
  if (i == j) {
    // diagonal limits (cheb-CFIE :585-591)
    double kap = g.kappa[j];
    double rjj = g.rk[g.off[cj]];
    cd dval = mk(wj * kap, 0.0);
    double m1 = -QB_INV_TWO_PI * sj;
    cd k2 = k * k;
    cd lg = clog(mk(0.25 * k2.re * (sj * sj), 0.25 * k2.im * (sj * sj)));
    cd m2 = mk(-QB_EULER_OVER_PI - QB_INV_TWO_PI * lg.re, 0.5 - QB_INV_TWO_PI * lg.im) * sj;
    cd sval = mk(rjj * m1, 0.0) + wj * m2;
    cd v = dval + ik * sval;
    return mk(1.0 - v.re, -v.im);
  }
  double dx = g.x[i] - g.x[j], dy = g.y[i] - g.y[j];
  double r = hypot(dx, dy);
  double invr = 1.0 / r;
  double inn = g.dy[j] * dx - g.dx[j] * dy;
  cd h0, h1;
  hank<MH, MJ>(tv, r, k, true, &h0, &h1);
  const cd aL2 = mk(-0.5 * k.im, 0.5 * k.re);  // 0.5 i k
  const cd aM2 = mk(0.0, 0.5);
  if (ci != cj) {
    // smooth off-component block (cheb-CFIE :632-648)
    double cD = wj * inn * invr;
    double cS = wj * sj;
    cd dval = cD * (aL2 * h1);
    cd sval = cS * (aM2 * h0);
    cd v = dval + ik * sval;
    return -v;
  }
  cd j0, j1;
  bessj<MH, MJ>(tv, r, &j0, &j1);
  int64_t o = g.off[ci];
  int64_t nc = g.off[ci + 1] - o;
  int64_t d = (i - j) % nc;
  if (d < 0) d += nc;
  double Rij = g.rk[o + d], lt = g.lt[o + d];
  const cd aL1 = mk(-k.re * QB_INV_TWO_PI, -k.im * QB_INV_TWO_PI);
  const double aM1 = -QB_INV_TWO_PI;
  cd L1 = aL1 * j1, L2 = aL2 * h1;
  cd M1 = aM1 * j0, M2 = aM2 * h0;
  double cD1 = Rij * inn * invr, cD2 = wj * inn * invr, cD3 = wj * lt * inn * invr;
  double cS1 = Rij * sj, cS2 = wj * sj, cS3 = wj * sj * lt;
  cd dval = cD1 * L1 + cD2 * L2 - cD3 * L1;
  cd sval = cS1 * M1 + cS2 * M2 - cS3 * M1;
  cd v = dval + ik * sval;
  return -v;
}

// DLP entry F(z)[i,j] = delta_ij - D_ij (single component).
template <int MH, int MJ>
__device__ cd dlp_entry(const DevGeom& g, const TabView& tv, int64_t i, int64_t j, cd k) {
  double wj = g.ws[j];
  if (i == j) return mk(1.0 - wj * g.kappa[j], 0.0);
  double dx = g.x[i] - g.x[j], dy = g.y[i] - g.y[j];
  double r = hypot(dx, dy);
  double invr = 1.0 / r;
  double inn = g.dy[j] * dx - g.dx[j] * dy;
  cd h0, h1, j0, j1;
  // eval_h_multi_ks! (bessels:1051-1076) has no |kr| < 0.2 branch
  hank<MH, MJ>(tv, r, k, false, &h0, &h1);
  bessj<MH, MJ>(tv, r, &j0, &j1);
  int64_t nc = g.N;
  int64_t d = (i - j) % nc;
  if (d < 0) d += nc;
  double Rij = g.rk[d], lt = g.lt[d];
  const cd aL1 = mk(-k.re * QB_INV_TWO_PI, -k.im * QB_INV_TWO_PI);
  const cd aL2 = mk(-0.5 * k.im, 0.5 * k.re);
  double cD = inn * invr;
  double c1 = Rij * cD, c2 = wj * cD, c3 = wj * lt * cD;
  cd l1 = aL1 * j1, ht = aL2 * h1;
  cd D = c1 * l1 + c2 * ht - c3 * l1;
  return -D;
}

template <int OP, int MH, int MJ>
__global__ void __launch_bounds__(256) k_assemble(DevGeom g, const cd* __restrict__ zk,
                                                  const cd* __restrict__ H,
                                                  const cd* __restrict__ J, int np_h, int M_h,
                                                  int np_j, int M_j, double rmin_h, double rmax,
                                                  int64_t q0, cd* __restrict__ out64,
                                                  cf* __restrict__ out32) {
  const int64_t N = g.N;
  const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const int64_t q = q0 + blockIdx.z;
  TabView tv;
  tv.H = H + q * (int64_t)np_h * (2 * (M_h + 1));
  tv.J = J + q * (int64_t)np_j * (2 * (M_j + 1));
  tv.np_h = np_h;
  tv.M_h = M_h;
  tv.np_j = np_j;
  tv.M_j = M_j;
  tv.rmin_h = rmin_h;
  tv.rmax = rmax;
  const cd k = zk[q];
  if (i >= N) return;
  const int64_t base = (q - q0) * N * N;
#pragma unroll 1
  for (int c = 0; c < 4; ++c) {
    int64_t j = (int64_t)blockIdx.y * 16 + threadIdx.y + 4 * c;
    if (j >= N) break;
    cd v = (OP == QB_CFIE_KRESS) ? cfie_entry<MH, MJ>(g, tv, i, j, k)
                                 : dlp_entry<MH, MJ>(g, tv, i, j, k);
    int64_t idx = base + i + j * N;
    if (out64) out64[idx] = v;
    if (out32) out32[idx] = to_cf(v);
  }
}

template <int OP, int MH, int MJ>
static void launch_asm(qb_ctx* ctx, const qb_geom* g, const cd* z, int64_t q0, int64_t nzc,
                       const DevTables& t, cd* o64, cf* o32) {
  int64_t N = g->d.N;
  dim3 grid((unsigned)((N + 63) / 64), (unsigned)((N + 15) / 16), (unsigned)nzc);
  double ent = (double)N * N * nzc;
  int slot = qb_prof_begin(ctx, QB_K_ASSEMBLE, 0.0,
                           ent * ((o64 ? sizeof(cd) : 0) + (o32 ? sizeof(cf) : 0)));
  hipLaunchKernelGGL((k_assemble<OP, MH, MJ>), grid, dim3(64, 4), 0, ctx->stream, g->d, z, t.H,
                     t.J, t.np_h, t.M_h, t.np_j, t.M_j, t.rmin_h, t.rmax, q0, o64, o32);
  qb_prof_end(ctx, slot);
}

int qb_assemble_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_dev, int64_t nz,
                    const DevTables& t, cd* out64, cf* out32) {
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS)
    return qb_fail(ctx, QB_E_ARG, "unknown operator");
  if (op == QB_DLP_KRESS && g->d.ncomp != 1)
    return qb_fail(ctx, QB_E_ARG, "DLP_kress expects one boundary component");
  int64_t N = g->d.N;
  if ((N + 15) / 16 > 65535)
    return qb_fail(ctx, QB_E_ARG, "N too large for the assembly grid");
  // grid.z is limited to 65535; chunk the wavenumbers
  for (int64_t q0 = 0; q0 < nz; q0 += 1024) {
    int64_t nzc = nz - q0 < 1024 ? nz - q0 : 1024;
    cd* o64 = out64 ? out64 + q0 * N * N : nullptr;
    cf* o32 = out32 ? out32 + q0 * N * N : nullptr;
    // the kernel indexes its output from the chunk start; pass q0 for tables
    bool fast = (t.M_h == 5 && t.M_j == 5);
    if (op == QB_CFIE_KRESS) {
      if (fast)
        launch_asm<QB_CFIE_KRESS, 5, 5>(ctx, g, z_dev, q0, nzc, t, o64, o32);
      else
        launch_asm<QB_CFIE_KRESS, 0, 0>(ctx, g, z_dev, q0, nzc, t, o64, o32);
    } else {
      if (fast)
        launch_asm<QB_DLP_KRESS, 5, 5>(ctx, g, z_dev, q0, nzc, t, o64, o32);
      else
        launch_asm<QB_DLP_KRESS, 0, 0>(ctx, g, z_dev, q0, nzc, t, o64, o32);
    }
    QB_HIP(ctx, hipGetLastError());
  }
  return QB_OK;
}
