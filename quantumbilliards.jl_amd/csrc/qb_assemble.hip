// K1 / K2: batched CFIE-Kress and DLP-Kress assembly of T(z) = I - K(z).
//
// Entry formulas follow _all_k_nosymm_CFIE_chebyshev!
// (src/chebyshev/chebyshev_euclidian_helmholtz_cfie_kress.jl:543-665) and
// _construct_dlp_kress_matrices_chebyshev! + the I - D pass
// (chebyshev_euclidian_helmholtz_dlp_kress.jl:702-769, :1086-1095).
//
// Layout: out[q*N*N + i + j*N] (column-major per wavenumber, Julia's layout).
// The geometry of a pair (distance, 1/r, the DLP numerator, the circulant
// Kress R and log term) is computed in registers instead of being read from
// N x N caches as the reference does.  As in the reference, the specials of a
// distance are evaluated once for the pair (i, j), (j, i).
#include "qb_bessel.h"
#include "qb_cheb.h"
#include "qb_internal.h"
#include <functional>

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

// Special functions of one distance r at one wavenumber: H0, H1 (Hankel
// table, or the direct evaluation below |kr| = 0.2 for CFIE), J0, J1 (J table,
// only needed inside one component).
struct Spec {
  cd h0, h1, j0, j1;
};

template <int OP, int MH, int MJ>
__device__ __forceinline__ Spec specials(const TabView& tv, double r, cd k, bool same) {
  Spec s;
  hank<MH, MJ>(tv, r, k, OP == QB_CFIE_KRESS, &s.h0, &s.h1);
  if (same) {
    bessj<MH, MJ>(tv, r, &s.j0, &s.j1);
  } else {
    s.j0 = mk(0, 0);
    s.j1 = mk(0, 0);
  }
  return s;
}

// Off-diagonal entry A(z)[i,j] (target i, source j) from the shared
// specials of r = |x_i - x_j| (cheb-CFIE :594-612 and :636-646;
// cheb-DLP :737-765 with F = I - D).
template <int OP>
__device__ __forceinline__ cd offdiag_entry(const DevGeom& g, int64_t i, int64_t j, double dx,
                                            double dy, double invr, bool same, double Rij,
                                            double lt, cd k, const Spec& sp) {
  const cd ik = mk(-k.im, k.re);
  const cd aL2 = mk(-0.5 * k.im, 0.5 * k.re);  // i k / 2
  const cd aM2 = mk(0.0, 0.5);
  const cd aL1 = mk(-k.re * QB_INV_TWO_PI, -k.im * QB_INV_TWO_PI);
  const double aM1 = -QB_INV_TWO_PI;
  double wj = g.ws[j];
  double inn = g.dy[j] * dx - g.dx[j] * dy;  // t_j x (x_i - x_j)
  if (OP == QB_DLP_KRESS) {
    double cD = inn * invr;
    double c1 = Rij * cD, c2 = wj * cD, c3 = wj * lt * cD;
    cd l1 = aL1 * sp.j1, ht = aL2 * sp.h1;
    return -(c1 * l1 + c2 * ht - c3 * l1);
  }
  double sj = g.speed[j];
  if (!same) {
    double cD = wj * inn * invr;
    double cS = wj * sj;
    cd v = cD * (aL2 * sp.h1) + ik * (cS * (aM2 * sp.h0));
    return -v;
  }
  cd L1 = aL1 * sp.j1, L2 = aL2 * sp.h1;
  cd M1 = aM1 * sp.j0, M2 = aM2 * sp.h0;
  double cD1 = Rij * inn * invr, cD2 = wj * inn * invr, cD3 = wj * lt * inn * invr;
  double cS1 = Rij * sj, cS2 = wj * sj, cS3 = wj * sj * lt;
  cd dval = cD1 * L1 + cD2 * L2 - cD3 * L1;
  cd sval = cS1 * M1 + cS2 * M2 - cS3 * M1;
  return -(dval + ik * sval);
}

// Diagonal entry (cheb-CFIE :585-591; cheb-DLP :729-733 with F = I - D).
template <int OP>
__device__ __forceinline__ cd diag_entry(const DevGeom& g, int64_t j, cd k) {
  double wj = g.ws[j];
  if (OP == QB_DLP_KRESS) return mk(1.0 - wj * g.kappa[j], 0.0);
  const cd ik = mk(-k.im, k.re);
  double sj = g.speed[j];
  double rjj = g.rk[g.off[g.comp[j]]];
  cd dval = mk(wj * g.kappa[j], 0.0);
  double m1 = -QB_INV_TWO_PI * sj;
  cd k2 = k * k;
  cd lg = clog(mk(0.25 * k2.re * (sj * sj), 0.25 * k2.im * (sj * sj)));
  cd m2 = mk(-QB_EULER_OVER_PI - QB_INV_TWO_PI * lg.re, 0.5 - QB_INV_TWO_PI * lg.im) * sj;
  cd sval = mk(rjj * m1, 0.0) + wj * m2;
  cd v = dval + ik * sval;
  return mk(1.0 - v.re, -v.im);
}

// Tile-pair assembly.  Block b of grid.x is the pair of 64 x 64 tiles (I, J)
// and (J, I), I >= J, enumerated diagonal by diagonal (d = I - J), so the
// workgroups resident together work on similar distances and read nearby
// table panels.  Each thread computes the specials of 16 distances once and
// produces both A[i,j] (stored directly: a wave writes 64 consecutive rows)
// and A[j,i] (staged in LDS, then written column by column of the mirror
// tile).  grid.z walks the wavenumbers.
template <int OP, int MH, int MJ>
__global__ void __launch_bounds__(256) k_assemble(DevGeom g, const cd* __restrict__ zk,
                                                  const cd* __restrict__ H,
                                                  const cd* __restrict__ J, int np_h, int M_h,
                                                  int np_j, int M_j, double rmin_h, double rmax,
                                                  int64_t q0, int T, cd* __restrict__ out64,
                                                  cf* __restrict__ out32) {
  __shared__ cd mir[64][65];  // mir[i_local][j_local] = A[j, i]
  const int64_t N = g.N;
  const int64_t q = q0 + blockIdx.z;
  // pair index -> (d, J): diagonal d has T - d pairs
  int64_t p = blockIdx.x;
  int d = 0;
  while (p >= T - d) {
    p -= T - d;
    ++d;
  }
  const int Jt = (int)p, It = Jt + d;
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
  const int il = threadIdx.x & 63, jl0 = threadIdx.x >> 6;
  const int64_t i = (int64_t)It * 64 + il;
  const int64_t base = (q - q0) * N * N;
  const bool diag_tile = (It == Jt);
  const bool iok = i < N;
  double xi = 0, yi = 0;
  int ci = 0;
  if (iok) {
    xi = g.x[i];
    yi = g.y[i];
    ci = g.comp[i];
  }
#pragma unroll 1
  for (int c = 0; c < 16; ++c) {
    const int jl = jl0 + 4 * c;
    const int64_t j = (int64_t)Jt * 64 + jl;
    if (!iok || j >= N) continue;
    cd vij, vji = mk(0, 0);
    if (i == j) {
      vij = diag_entry<OP>(g, j, k);
    } else {
      double dx = xi - g.x[j], dy = yi - g.y[j];
      double r = hypot(dx, dy);
      double invr = 1.0 / r;
      int cj = g.comp[j];
      bool same = (ci == cj);
      double Rij = 0, lt = 0;
      if (same) {
        int64_t o = g.off[ci];
        int64_t nc = g.off[ci + 1] - o;
        int64_t dd = (i - j) % nc;
        if (dd < 0) dd += nc;
        Rij = g.rk[o + dd];
        lt = g.lt[o + dd];  // both circulant columns are even in (i - j)
      }
      Spec sp = specials<OP, MH, MJ>(tv, r, k, same || OP == QB_DLP_KRESS);
      vij = offdiag_entry<OP>(g, i, j, dx, dy, invr, same, Rij, lt, k, sp);
      if (!diag_tile) vji = offdiag_entry<OP>(g, j, i, -dx, -dy, invr, same, Rij, lt, k, sp);
    }
    int64_t idx = base + i + j * N;
    if (out64) out64[idx] = vij;
    if (out32) out32[idx] = to_cf(vij);
    if (!diag_tile) mir[il][jl] = vji;
  }
  if (diag_tile) return;
  __syncthreads();
  // mirror tile: column i (64 of them), rows j = Jt*64 + 0..63 contiguous
#pragma unroll 1
  for (int c = 0; c < 16; ++c) {
    const int il2 = jl0 + 4 * c;             // which column of the mirror tile
    const int jl2 = threadIdx.x & 63;        // row inside it
    const int64_t ii = (int64_t)It * 64 + il2, jj = (int64_t)Jt * 64 + jl2;
    if (ii >= N || jj >= N) continue;
    cd v = mir[il2][jl2];
    int64_t idx = base + jj + ii * N;
    if (out64) out64[idx] = v;
    if (out32) out32[idx] = to_cf(v);
  }
}

template <int OP, int MH, int MJ>
static void launch_asm(qb_ctx* ctx, const qb_geom* g, const cd* z, int64_t q0, int64_t nzc,
                       const DevTables& t, cd* o64, cf* o32) {
  int64_t N = g->d.N;
  int T = (int)((N + 63) / 64);
  int64_t pairs = (int64_t)T * (T + 1) / 2;
  dim3 grid((unsigned)pairs, 1, (unsigned)nzc);
  double ent = (double)N * N * nzc;
  int slot = qb_prof_begin(ctx, QB_K_ASSEMBLE, 0.0,
                           ent * ((o64 ? sizeof(cd) : 0) + (o32 ? sizeof(cf) : 0)));
  hipLaunchKernelGGL((k_assemble<OP, MH, MJ>), grid, dim3(256), 0, ctx->stream, g->d, z, t.H,
                     t.J, t.np_h, t.M_h, t.np_j, t.M_j, t.rmin_h, t.rmax, q0, T, o64, o32);
  qb_prof_end(ctx, slot);
}

// Conjugate-pair assembly for contours symmetric about the real axis (the
// Beyn windows with real k0): nodes a and b with k_b = conj(k_a) share the
// panel of every distance, and
//   J_n(k_a r) = (H_n(k_a r) + conj(H_n(k_b r))) / 2,  J_n(k_b r) = conj(J_n(k_a r))
// (H^(2)(z) = conj(H^(1)(conj z)) for real order), so one block assembles both
// matrices from the two nodes' Hankel panels alone: per entry pair and node,
// one panel gather instead of the H and J gathers of k_assemble.  A
// self-conjugate node (real k, a == b) gives J = Re H.  The mirror entries
// (j, i) of both nodes are staged in LDS 32 rows at a time and written as
// columns of the mirror tiles.  pairs[blockIdx.z] = (a, b); outputs indexed
// by node from out64 / out32.
template <int OP, int MH>
__global__ void __launch_bounds__(256)
k_assemble_conj(DevGeom g, const cd* __restrict__ zk, const cd* __restrict__ H, int np_h,
                double rmin_h, double rmax, const int* __restrict__ pairs, int T,
                cd* __restrict__ out64, cf* __restrict__ out32) {
  __shared__ cd mir[2][64][33];  // [node][i_local][j_local - 32 h] = A[j, i], half h
  constexpr int PH = 2 * (MH + 1);
  const int64_t N = g.N;
  const int qa = pairs[2 * blockIdx.z], qb = pairs[2 * blockIdx.z + 1];
  const bool two = qa != qb;
  int64_t p = blockIdx.x;
  int d = 0;
  while (p >= T - d) {
    p -= T - d;
    ++d;
  }
  const int Jt = (int)p, It = Jt + d;
  const cd* Ha = H + (int64_t)qa * np_h * PH;
  const cd* Hb = H + (int64_t)qb * np_h * PH;
  const cd ka = zk[qa], kb = zk[qb];
  const int il = threadIdx.x & 63, jl0 = threadIdx.x >> 6;
  const int64_t i = (int64_t)It * 64 + il;
  const bool iok = i < N;
  const int64_t ba = (int64_t)qa * N * N, bb = (int64_t)qb * N * N;
  const bool diag_tile = (It == Jt);
  const int64_t ic = iok ? i : N - 1;
  const double xi = g.x[ic], yi = g.y[ic];
  const int ci = g.comp[ic];
#pragma unroll 1
  for (int h = 0; h < 2; ++h) {
#pragma unroll 1
    for (int c = 8 * h; c < 8 * h + 8; ++c) {
      const int jl = jl0 + 4 * c;
      const int64_t j = (int64_t)Jt * 64 + jl;
      if (!iok || j >= N) continue;
      cd va, vb, wa = mk(0, 0), wb = mk(0, 0);  // (i, j) and (j, i) of nodes a and b
      if (i == j) {
        va = diag_entry<OP>(g, j, ka);
        vb = two ? diag_entry<OP>(g, j, kb) : va;
      } else {
        const double dx = xi - g.x[j], dy = yi - g.y[j];
        const double r = hypot(dx, dy);
        const double invr = 1.0 / r;
        const bool same = (ci == g.comp[j]);
        double Rij = 0, lt = 0;
        if (same) {
          const int64_t o = g.off[ci];
          const int64_t nc = g.off[ci + 1] - o;
          int64_t dd = (i - j) % nc;
          if (dd < 0) dd += nc;
          Rij = g.rk[o + dd];
          lt = g.lt[o + dd];  // both circulant columns are even in (i - j)
        }
        int pp;
        double t;
        qb_panel(r, rmin_h, rmax, np_h, &pp, &t);
        Spec sa, sb;
        qb_clenshaw2<MH>(Ha + (int64_t)pp * PH, t, &sa.h0, &sa.h1);
        if (two)
          qb_clenshaw2<MH>(Hb + (int64_t)pp * PH, t, &sb.h0, &sb.h1);
        else
          sb = sa;
        if (OP == QB_CFIE_KRESS) {
          // h0_h1_multi_ks_at_r! (bessels:1190-1206): |kr| < 0.2 directly
          const cd z = mk(ka.re * r, ka.im * r);
          if (cabsd(z) < 0.2) {
            qb_hankel1_01(z, &sa.h0, &sa.h1);
            if (two)
              qb_hankel1_01(mk(kb.re * r, kb.im * r), &sb.h0, &sb.h1);
            else
              sb = sa;
          }
        }
        if (same || OP == QB_DLP_KRESS) {
          sa.j0 = 0.5 * (sa.h0 + cconj(sb.h0));
          sa.j1 = 0.5 * (sa.h1 + cconj(sb.h1));
        } else {
          sa.j0 = mk(0, 0);
          sa.j1 = mk(0, 0);
        }
        sb.j0 = cconj(sa.j0);
        sb.j1 = cconj(sa.j1);
        va = offdiag_entry<OP>(g, i, j, dx, dy, invr, same, Rij, lt, ka, sa);
        vb = two ? offdiag_entry<OP>(g, i, j, dx, dy, invr, same, Rij, lt, kb, sb) : va;
        if (!diag_tile) {
          wa = offdiag_entry<OP>(g, j, i, -dx, -dy, invr, same, Rij, lt, ka, sa);
          wb = two ? offdiag_entry<OP>(g, j, i, -dx, -dy, invr, same, Rij, lt, kb, sb) : wa;
        }
      }
      const int64_t e1 = i + j * N;
      if (out64) {
        out64[ba + e1] = va;
        if (two) out64[bb + e1] = vb;
      }
      if (out32) {
        out32[ba + e1] = to_cf(va);
        if (two) out32[bb + e1] = to_cf(vb);
      }
      if (!diag_tile) {
        mir[0][il][jl - 32 * h] = wa;
        mir[1][il][jl - 32 * h] = wb;
      }
    }
    if (diag_tile) continue;  // uniform over the block
    __syncthreads();
    // mirror rows j = Jt*64 + 32h + 0..31 of the 64 columns i: lanes 0..31
    // write one column's 32 contiguous rows, lanes 32..63 the next column
#pragma unroll 1
    for (int c = 0; c < 8; ++c) {
      const int il2 = 2 * (jl0 + 4 * c) + (threadIdx.x >> 5 & 1);
      const int jl2 = threadIdx.x & 31;
      const int64_t ii = (int64_t)It * 64 + il2, jj = (int64_t)Jt * 64 + 32 * h + jl2;
      if (ii >= N || jj >= N) continue;
      const int64_t e2 = jj + ii * N;
      const cd xa = mir[0][il2][jl2];
      if (out64) out64[ba + e2] = xa;
      if (out32) out32[ba + e2] = to_cf(xa);
      if (two) {
        const cd xb = mir[1][il2][jl2];
        if (out64) out64[bb + e2] = xb;
        if (out32) out32[bb + e2] = to_cf(xb);
      }
    }
    __syncthreads();
  }
}

// One-to-one pairing of the wavenumbers with their conjugates (several
// disks of one window can share a real node): a real node pairs with itself,
// any other node with an unused conjugate.  False when some node has none.
bool qb_conj_pairs(const cd* z_host, int64_t nz, std::vector<int>* pr) {
  pr->clear();
  std::vector<char> used(nz, 0);
  for (int64_t a = 0; a < nz; ++a) {
    if (used[a]) continue;
    const cd za = z_host[a];
    const double sc = fabs(za.re) + fabs(za.im);
    int64_t b = -1;
    if (fabs(za.im) <= 1e-14 * sc) {
      b = a;
    } else {
      for (int64_t c = a + 1; c < nz; ++c) {
        const cd zc = z_host[c];
        if (!used[c] && fabs(zc.re - za.re) <= 1e-14 * sc && fabs(zc.im + za.im) <= 1e-14 * sc) {
          b = c;
          break;
        }
      }
    }
    if (b < 0) return false;
    used[a] = used[b] = 1;
    pr->push_back((int)a);
    pr->push_back((int)b);
  }
  const int64_t npair = (int64_t)pr->size() / 2;
  return npair >= 1 && npair <= 65535;
}

int qb_assemble_conj_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_host,
                         const cd* z_dev, int64_t nz, const DevTables& t, cd* out64,
                         cf* out32) {
  // usable: plain geometry, tables of static degree 5 or 7, and every node's
  // conjugate among the nodes (to rounding); returns 1 when not usable
  if (g->sym.m > 0 || t.M_h != t.M_j || (t.M_h != 5 && t.M_h != 7)) return 1;
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS) return 1;
  if (op == QB_DLP_KRESS && g->d.ncomp != 1) return 1;
  std::vector<int> pr;
  if (!qb_conj_pairs(z_host, nz, &pr)) return 1;
  int npair = (int)pr.size() / 2;
  int* pd = nullptr;
  int rc = qb_dev_alloc(ctx, (void**)&pd, pr.size() * sizeof(int));
  if (rc) return rc;
  if (hipMemcpyAsync(pd, pr.data(), pr.size() * sizeof(int), hipMemcpyHostToDevice,
                     ctx->stream) != hipSuccess) {
    qb_dev_free(ctx, pd, pr.size() * sizeof(int));
    return qb_fail(ctx, QB_E_HIP, "pair upload failed");
  }
  const int64_t N = g->d.N;
  const int T = (int)((N + 63) / 64);
  const int64_t tpairs = (int64_t)T * (T + 1) / 2;
  const double ent = (double)N * N * nz;
  int slot = qb_prof_begin(ctx, QB_K_ASSEMBLE, 0.0,
                           ent * ((out64 ? sizeof(cd) : 0) + (out32 ? sizeof(cf) : 0)));
  const dim3 grid((unsigned)tpairs, 1, (unsigned)npair);
  if (op == QB_CFIE_KRESS) {
    if (t.M_h == 5)
      hipLaunchKernelGGL((k_assemble_conj<QB_CFIE_KRESS, 5>), grid, dim3(256), 0, ctx->stream,
                         g->d, z_dev, t.H, t.np_h, t.rmin_h, t.rmax, pd, T, out64, out32);
    else
      hipLaunchKernelGGL((k_assemble_conj<QB_CFIE_KRESS, 7>), grid, dim3(256), 0, ctx->stream,
                         g->d, z_dev, t.H, t.np_h, t.rmin_h, t.rmax, pd, T, out64, out32);
  } else {
    if (t.M_h == 5)
      hipLaunchKernelGGL((k_assemble_conj<QB_DLP_KRESS, 5>), grid, dim3(256), 0, ctx->stream,
                         g->d, z_dev, t.H, t.np_h, t.rmin_h, t.rmax, pd, T, out64, out32);
    else
      hipLaunchKernelGGL((k_assemble_conj<QB_DLP_KRESS, 7>), grid, dim3(256), 0, ctx->stream,
                         g->d, z_dev, t.H, t.np_h, t.rmin_h, t.rmax, pd, T, out64, out32);
  }
  qb_prof_end(ctx, slot);
  hipError_t e = hipGetLastError();
  // the pair list must outlive the launch: wait before handing it back
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  qb_dev_free(ctx, pd, pr.size() * sizeof(int));
  if (e != hipSuccess) return qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
  return QB_OK;
}

// ---------------------------------------------------------------------------
// Conjugate-pair assembly from envelope tables (the kernel of Beyn windows
// with real k0)
// ---------------------------------------------------------------------------
// H_n^(1)(z) = e^{iz} E_n(z) with E_n smooth and non-oscillatory (~ (2 / pi
// z)^{1/2} (1 + O(1/z)) for large z; the log / pole of H_n at z = 0 for
// small z).  At fixed wavenumber k the envelope E_n(k r) is a Chebyshev
// series of low degree on geometric panels: quarter octaves
// [2^o (1 + q/4), 2^o (1 + (q+1)/4)], degree 12, about 1e-15 relative on
// every panel (E_n is analytic in the Bernstein ellipse of a quarter octave,
// whose nearest singularity, r = 0, lies 9 half-widths from its centre).
// From rmin ~ 1e-3 to rmax ~ 2.6 that is 52 panels and 21 KB per wavenumber
// for E_0 and E_1, against the ~1.2e5 panels of the uniform tables of H_n
// itself (their oscillation e^{ikr} at k = 800 needs them), so both nodes'
// tables of a conjugate pair sit in LDS.  Per distance: the panel from the
// exponent and the top mantissa bits of r (exact), one Clenshaw sweep per
// node, one sincos(Re k r) and one exp(Im k r) shared by the pair
// (e^{i conj(k) r} = e^{i Re k r} e^{Im k r}).  H comes out to ~1e-14
// relative, at least as accurate as the reference's uniform tables at their
// 1e-12 tolerance; J, the |k r| < 0.2 direct branch and the entry formulas
// are those of k_assemble_conj.
#define QB_ENV_M 12
#define QB_ENV_NC (QB_ENV_M + 1)
#define QB_ENV_PMAX 128  // 32 octaves

struct EnvLob {
  double t[QB_ENV_NC];                 // cospi(j / M)
  double C[QB_ENV_NC * QB_ENV_NC];     // C[m (M+1) + j] = cospi(j m / M)
};

// Quarter-octave panel p (0-based from octave o0) and local coordinate t of r.
__device__ __forceinline__ void env_panel(double r, int o0, int P, int* p, double* t) {
  const int e = ilogb(r);
  const double m = ldexp(r, -e);  // [1, 2)
  int q = (int)((m - 1.0) * 4.0);
  q = q > 3 ? 3 : q;
  int pp = 4 * (e - o0) + q;
  pp = pp < 0 ? 0 : (pp > P - 1 ? P - 1 : pp);
  *p = pp;
  *t = 8.0 * m - (9.0 + 2.0 * q);
}

// Envelope tables: one thread per (wavenumber, panel); coefficients of the
// interpolant at the M+1 Lobatto points (DCT-I with c_0 and c_M halved),
// E_0 and E_1 interleaved: env[((q P + p) (M+1) + m) 2 + {0, 1}].
__global__ __launch_bounds__(64) void k_env_tables(const cd* __restrict__ zk, int64_t nz, int o0,
                                                   int P, EnvLob lc, cd* __restrict__ env) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nz * (int64_t)P) return;
  const int64_t q = gid / P;
  const int p = (int)(gid - q * P);
  const int oct = o0 + p / 4, qq = p % 4;
  const double a = ldexp(1.0 + 0.25 * qq, oct), b = ldexp(1.25 + 0.25 * qq, oct);
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  const cd k = zk[q];
  cd f0[QB_ENV_NC], f1[QB_ENV_NC];
#pragma unroll
  for (int j = 0; j < QB_ENV_NC; ++j) {
    const double r = fma(h, lc.t[j], c);
    const cd z = mk(k.re * r, k.im * r);
    cd h0, h1;
    qb_hankel1_01(z, &h0, &h1);
    double sn, cs;
    sincos(z.re, &sn, &cs);
    const double g = exp(z.im);
    const cd e = mk(cs * g, -sn * g);  // e^{-i z}
    f0[j] = h0 * e;
    f1[j] = h1 * e;
  }
  cd* o = env + (q * P + p) * (int64_t)(2 * QB_ENV_NC);
  const double s2 = 2.0 / QB_ENV_M;
#pragma unroll
  for (int m = 0; m <= QB_ENV_M; ++m) {
    const double sg = (m & 1) ? -1.0 : 1.0;
    cd s0 = 0.5 * f0[0] + 0.5 * (sg * f0[QB_ENV_M]);
    cd s1 = 0.5 * f1[0] + 0.5 * (sg * f1[QB_ENV_M]);
    for (int j = 1; j < QB_ENV_M; ++j) {
      const double cc = lc.C[m * QB_ENV_NC + j];
      s0 = s0 + f0[j] * cc;
      s1 = s1 + f1[j] * cc;
    }
    const double w = (m == 0 || m == QB_ENV_M) ? 0.5 * s2 : s2;
    o[2 * m] = w * s0;
    o[2 * m + 1] = w * s1;
  }
}

// MW: columns of the mirror tile staged through LDS at a time (16, or 4 for
// the variant that fits three workgroups per CU: 2 x 64 x 5 instead of
// 2 x 64 x 17 entries); OCC: waves per SIMD the registers must allow.
template <int OP, int MW, int OCC>
__global__ void __launch_bounds__(256, OCC)
k_assemble_env(DevGeom g, const cd* __restrict__ zk, const cd* __restrict__ env, int o0, int P,
               const int* __restrict__ pairs, int T, int J0, int first, cd* __restrict__ out64,
               cf* __restrict__ out32) {
  static_assert(MW == 4 || MW == 8 || MW == 16, "mirror width");
  extern __shared__ __align__(16) cd smem[];
  const int tsz = P * 2 * QB_ENV_NC;  // complex coefficients per node
  cd* Ta = smem;
  cd* Tb = smem + tsz;
  cd (*mir)[64][MW + 1] = (cd (*)[64][MW + 1])(smem + 2 * tsz);  // [node][i_local][j_local - MW h]
  const int64_t N = g.N;
  const int qa = pairs[2 * blockIdx.z], qb = pairs[2 * blockIdx.z + 1];
  const bool two = qa != qb;
  for (int e = threadIdx.x; e < tsz; e += 256) {
    Ta[e] = env[(int64_t)qa * tsz + e];
    Tb[e] = env[(int64_t)qb * tsz + e];
  }
  // tile pair (It, Jt), It >= Jt: first = 1 walks the pairs with Jt < J0
  // column by column (the first J0 tile columns, and by the mirror the first
  // J0 tile rows, complete); first = 0 the pairs with Jt >= J0 diagonal by
  // diagonal (J0 = 0: every pair)
  int64_t p = blockIdx.x;
  int It, Jt;
  if (first) {
    int j = 0;
    while (p >= T - j) {
      p -= T - j;
      ++j;
    }
    Jt = j;
    It = j + (int)p;
  } else {
    const int Tr = T - J0;
    int d = 0;
    while (p >= Tr - d) {
      p -= Tr - d;
      ++d;
    }
    Jt = J0 + (int)p;
    It = Jt + d;
  }
  const cd ka = zk[qa], kb = zk[qb];
  const int il = threadIdx.x & 63, jl0 = threadIdx.x >> 6;
  const int64_t i = (int64_t)It * 64 + il;
  const bool iok = i < N;
  const int64_t ba = (int64_t)qa * N * N, bb = (int64_t)qb * N * N;
  const bool diag_tile = (It == Jt);
  const int64_t ic = iok ? i : N - 1;
  const double xi = g.x[ic], yi = g.y[ic];
  const int ci = g.comp[ic];
  __syncthreads();
  constexpr int CPS = MW / 4;  // column groups (4 columns each) per mirror stage
#pragma unroll 1
  for (int h = 0; h < 64 / MW; ++h) {
#pragma unroll 1
    for (int c = CPS * h; c < CPS * h + CPS; ++c) {
      const int jl = jl0 + 4 * c;
      const int64_t j = (int64_t)Jt * 64 + jl;
      if (!iok || j >= N) continue;
      cd va, vb, wa = mk(0, 0), wb = mk(0, 0);  // (i, j) and (j, i) of nodes a and b
      if (i == j) {
        va = diag_entry<OP>(g, j, ka);
        vb = two ? diag_entry<OP>(g, j, kb) : va;
      } else {
        const double dx = xi - g.x[j], dy = yi - g.y[j];
        const double r = hypot(dx, dy);
        const double invr = 1.0 / r;
        const bool same = (ci == g.comp[j]);
        double Rij = 0, lt = 0;
        if (same) {
          const int64_t o = g.off[ci];
          const int64_t nc = g.off[ci + 1] - o;
          int64_t dd = (i - j) % nc;
          if (dd < 0) dd += nc;
          Rij = g.rk[o + dd];
          lt = g.lt[o + dd];  // both circulant columns are even in (i - j)
        }
        Spec sa, sb;
        const cd za = mk(ka.re * r, ka.im * r);
        if (OP == QB_CFIE_KRESS && cabsd(za) < 0.2) {
          // h0_h1_multi_ks_at_r! (bessels:1190-1206): |kr| < 0.2 directly
          qb_hankel1_01(za, &sa.h0, &sa.h1);
          if (two)
            qb_hankel1_01(mk(kb.re * r, kb.im * r), &sb.h0, &sb.h1);
          else
            sb = sa;
        } else {
          int pp;
          double t;
          env_panel(r, o0, P, &pp, &t);
          double sn, cs;
          sincos(za.re, &sn, &cs);
          const double ga = exp(-za.im);  // |e^{i k_a r}|; node b has k_b = conj(k_a)
          cd e0, e1;
          qb_clenshaw2<QB_ENV_M>(Ta + pp * (2 * QB_ENV_NC), t, &e0, &e1);
          const cd ea = mk(cs * ga, sn * ga);
          sa.h0 = e0 * ea;
          sa.h1 = e1 * ea;
          if (two) {
            qb_clenshaw2<QB_ENV_M>(Tb + pp * (2 * QB_ENV_NC), t, &e0, &e1);
            const double gb = 1.0 / ga;
            const cd eb = mk(cs * gb, sn * gb);
            sb.h0 = e0 * eb;
            sb.h1 = e1 * eb;
          } else {
            sb = sa;
          }
        }
        if (same || OP == QB_DLP_KRESS) {
          sa.j0 = 0.5 * (sa.h0 + cconj(sb.h0));
          sa.j1 = 0.5 * (sa.h1 + cconj(sb.h1));
        } else {
          sa.j0 = mk(0, 0);
          sa.j1 = mk(0, 0);
        }
        sb.j0 = cconj(sa.j0);
        sb.j1 = cconj(sa.j1);
        va = offdiag_entry<OP>(g, i, j, dx, dy, invr, same, Rij, lt, ka, sa);
        vb = two ? offdiag_entry<OP>(g, i, j, dx, dy, invr, same, Rij, lt, kb, sb) : va;
        if (!diag_tile) {
          wa = offdiag_entry<OP>(g, j, i, -dx, -dy, invr, same, Rij, lt, ka, sa);
          wb = two ? offdiag_entry<OP>(g, j, i, -dx, -dy, invr, same, Rij, lt, kb, sb) : wa;
        }
      }
      const int64_t e1 = i + j * N;
      if (out64) {
        out64[ba + e1] = va;
        if (two) out64[bb + e1] = vb;
      }
      if (out32) {
        out32[ba + e1] = to_cf(va);
        if (two) out32[bb + e1] = to_cf(vb);
      }
      if (!diag_tile) {
        mir[0][il][jl - MW * h] = wa;
        mir[1][il][jl - MW * h] = wb;
      }
    }
    if (diag_tile) continue;  // uniform over the block
    __syncthreads();
    // mirror rows j = Jt*64 + MW h + 0..MW-1 of the 64 columns i: each group
    // of MW lanes writes one column's MW contiguous rows
#pragma unroll 1
    for (int c = 0; c < MW / 4; ++c) {
      const int il2 = (int)(threadIdx.x / MW) + (256 / MW) * c;
      const int jl2 = (int)(threadIdx.x % MW);
      const int64_t ii = (int64_t)It * 64 + il2, jj = (int64_t)Jt * 64 + MW * h + jl2;
      if (ii >= N || jj >= N) continue;
      const int64_t e2 = jj + ii * N;
      const cd xa = mir[0][il2][jl2];
      if (out64) out64[ba + e2] = xa;
      if (out32) out32[ba + e2] = to_cf(xa);
      if (two) {
        const cd xb = mir[1][il2][jl2];
        if (out64) out64[bb + e2] = xb;
        if (out32) out32[bb + e2] = to_cf(xb);
      }
    }
    __syncthreads();
  }
}

// Default variant of the envelope assembly per operator, chosen by measurement
// (three workgroups per CU for DLP, whose registers fit without spilling).
static bool env_occ3_default(int op) { return op == QB_DLP_KRESS; }

int qb_assemble_env_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_host,
                        const cd* z_dev, int64_t nz, double rmin, double rmax, cd* out64,
                        cf* out32, int cols_first, const std::function<int()>& between) {
  // usable: plain geometry and every node's conjugate among the nodes;
  // returns 1 when not usable
  if (g->sym.m > 0) return 1;
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS) return 1;
  if (op == QB_DLP_KRESS && g->d.ncomp != 1) return 1;
  if (!(rmin > 0.0) || !(rmax > rmin)) return 1;
  std::vector<int> pr;
  if (!qb_conj_pairs(z_host, nz, &pr)) return 1;
  const int o0 = ilogb(rmin), o1 = ilogb(rmax);
  const int P = 4 * (o1 - o0 + 1);
  if (P > QB_ENV_PMAX) return 1;
  const int npair = (int)pr.size() / 2;
  const size_t ebytes = (size_t)nz * P * 2 * QB_ENV_NC * sizeof(cd);
  cd* env = nullptr;
  int* pd = nullptr;
  int rc = qb_dev_alloc(ctx, (void**)&env, ebytes);
  if (!rc) rc = qb_dev_alloc(ctx, (void**)&pd, pr.size() * sizeof(int));
  if (rc) {
    qb_dev_free(ctx, env, ebytes);
    return rc;
  }
  EnvLob lc;
  qb_lobatto_host(QB_ENV_M, lc.t, lc.C);
  const int64_t nt = nz * (int64_t)P;
  int slot = qb_prof_begin(ctx, QB_K_TABLES, 0.0, (double)ebytes);
  hipLaunchKernelGGL(k_env_tables, dim3((unsigned)((nt + 63) / 64)), dim3(64), 0, ctx->stream,
                     z_dev, nz, o0, P, lc, env);
  qb_prof_end(ctx, slot);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(pd, pr.data(), pr.size() * sizeof(int), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess) {
    const int64_t N = g->d.N;
    const int T = (int)((N + 63) / 64);
    const int64_t tpairs = (int64_t)T * (T + 1) / 2;
    // the first cols_first columns (whole 64-column tiles) in a launch of
    // their own, then `between` (enqueued while the rest assembles)
    const int J0 = between && cols_first > 0 && cols_first % 64 == 0 && cols_first / 64 < T
                       ? cols_first / 64
                       : 0;
    const int64_t pairs_first = (int64_t)J0 * T - (int64_t)J0 * (J0 - 1) / 2;
    // the three-workgroups-per-CU variant (4-column mirror stages) where its
    // LDS fits three times in a CU (QB_ASM_ENV_OCC=2 or 3 forces a variant)
    const size_t tab = (size_t)2 * P * 2 * QB_ENV_NC * sizeof(cd);
    const size_t lds16 = tab + 2 * 64 * 17 * sizeof(cd), lds4 = tab + 2 * 64 * 5 * sizeof(cd);
    bool occ3 = lds4 <= (size_t)(160 * 1024) / 3 && env_occ3_default(op);
    if (const char* e = getenv("QB_ASM_ENV_OCC"))
      occ3 = e[0] == '3' && lds4 <= (size_t)(160 * 1024) / 3;
    const size_t lds = occ3 ? lds4 : lds16;
    const double ent = (double)N * N * nz;
    slot = qb_prof_begin(ctx, QB_K_ASSEMBLE, 0.0,
                         ent * ((out64 ? sizeof(cd) : 0) + (out32 ? sizeof(cf) : 0)));
    auto launch = [&](auto kern, int64_t npairs, int first) {
      (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
      hipLaunchKernelGGL(kern, dim3((unsigned)npairs, 1, (unsigned)npair), dim3(256), lds,
                         ctx->stream, g->d, z_dev, env, o0, P, pd, T, J0, first, out64, out32);
    };
    auto launch_all = [&](int64_t npairs, int first) {
      if (op == QB_CFIE_KRESS) {
        if (occ3)
          launch(k_assemble_env<QB_CFIE_KRESS, 4, 3>, npairs, first);
        else
          launch(k_assemble_env<QB_CFIE_KRESS, 16, 1>, npairs, first);
      } else {
        if (occ3)
          launch(k_assemble_env<QB_DLP_KRESS, 4, 3>, npairs, first);
        else
          launch(k_assemble_env<QB_DLP_KRESS, 16, 1>, npairs, first);
      }
    };
    int brc = 0;
    if (J0 > 0) {
      launch_all(pairs_first, 1);
      e = hipGetLastError();
      if (e == hipSuccess) brc = between();
      if (e == hipSuccess && brc == 0) launch_all(tpairs - pairs_first, 0);
    } else {
      launch_all(tpairs, 0);
    }
    qb_prof_end(ctx, slot);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess && brc != 0) {
      (void)hipStreamSynchronize(ctx->stream);
      qb_dev_free(ctx, pd, pr.size() * sizeof(int));
      qb_dev_free(ctx, env, ebytes);
      return brc;
    }
  }
  // the tables and the pair list must outlive the launch
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  else (void)hipStreamSynchronize(ctx->stream);
  qb_dev_free(ctx, pd, pr.size() * sizeof(int));
  qb_dev_free(ctx, env, ebytes);
  if (e != hipSuccess) return qb_fail(ctx, QB_E_HIP, hipGetErrorString(e));
  return QB_OK;
}

int qb_assemble_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_dev, int64_t nz,
                    const DevTables& t, cd* out64, cf* out32) {
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS)
    return qb_fail(ctx, QB_E_ARG, "unknown operator");
  if (op == QB_DLP_KRESS && g->d.ncomp != 1)
    return qb_fail(ctx, QB_E_ARG, "DLP_kress expects one boundary component");
  if (g->sym.m > 0) return qb_assemble_red_dev(ctx, g, op, z_dev, nz, t, out64, out32);
  int64_t N = g->d.N;
  // grid.x = number of 64 x 64 tile pairs (2^31 - 1 limit: N < 2.9e6)
  // grid.z is limited to 65535; chunk the wavenumbers
  for (int64_t q0 = 0; q0 < nz; q0 += 1024) {
    int64_t nzc = nz - q0 < 1024 ? nz - q0 : 1024;
    cd* o64 = out64 ? out64 + q0 * N * N : nullptr;
    cf* o32 = out32 ? out32 + q0 * N * N : nullptr;
    // the kernel indexes its output from the chunk start; pass q0 for tables
    const int mh = t.M_h, mj = t.M_j;
    if (op == QB_CFIE_KRESS) {
      if (mh == 5 && mj == 5)
        launch_asm<QB_CFIE_KRESS, 5, 5>(ctx, g, z_dev, q0, nzc, t, o64, o32);
      else if (mh == 7 && mj == 7)
        launch_asm<QB_CFIE_KRESS, 7, 7>(ctx, g, z_dev, q0, nzc, t, o64, o32);
      else
        launch_asm<QB_CFIE_KRESS, 0, 0>(ctx, g, z_dev, q0, nzc, t, o64, o32);
    } else {
      if (mh == 5 && mj == 5)
        launch_asm<QB_DLP_KRESS, 5, 5>(ctx, g, z_dev, q0, nzc, t, o64, o32);
      else if (mh == 7 && mj == 7)
        launch_asm<QB_DLP_KRESS, 7, 7>(ctx, g, z_dev, q0, nzc, t, o64, o32);
      else
        launch_asm<QB_DLP_KRESS, 0, 0>(ctx, g, z_dev, q0, nzc, t, o64, o32);
    }
    QB_HIP(ctx, hipGetLastError());
  }
  return QB_OK;
}

// ---------------------------------------------------------------------------
// Assembly with the first and second wavenumber derivatives, T, dT/dk and
// d2T/dk2, for the expanded boundary integral method:
//   CFIE: _all_k_nosymm_CFIE_chebyshev_deriv! (cheb-CFIE:881-1059) behind
//         construct_boundary_matrices_with_derivatives! (cheb-CFIE:1507-1530);
//   DLP:  _construct_dlp_kress_matrices_derivatives_chebyshev! and the sign
//         flip + identity of its caller (cheb-DLP:881-952, :1097-1108), behind
//         construct_boundary_matrices_with_derivatives! (cheb-DLP:1126-1144).
// Both read H0, H1, J0, J1 with the |kr| < 0.2 direct branch
// (h0_h1_j0_j1_multi_ks_at_r!; the plain DLP path has no such branch).
// ---------------------------------------------------------------------------
struct D3 {
  cd a, a1, a2;
};

template <int MH, int MJ>
__device__ __forceinline__ Spec specials_all(const TabView& tv, double r, cd k, bool want_j) {
  Spec s;
  hank<MH, MJ>(tv, r, k, true, &s.h0, &s.h1);
  if (want_j) {
    bessj<MH, MJ>(tv, r, &s.j0, &s.j1);
  } else {
    s.j0 = mk(0, 0);
    s.j1 = mk(0, 0);
  }
  return s;
}

// Off-diagonal entry (target i, source j) of T, T', T''.
// Between components the derivative of the double-layer term
// w_j (inn/r)(i k/2) H1(kr) is +(i/2) w_j inn k H0 (and its k-derivative);
// the reference writes both with a minus sign (cheb-CFIE:1031-1032).
// off_sign = -1 reproduces that, +1 (the default) is the derivative.
template <int OP>
__device__ __forceinline__ D3 offdiag_deriv(const DevGeom& g, int64_t j, double dx, double dy,
                                            double r, double invr, bool same, double Rij,
                                            double lt, cd k, const Spec& sp, double off_sign) {
  const cd ik = mk(-k.im, k.re);
  const cd I = mk(0.0, 1.0);
  const cd aL1 = mk(-k.re * QB_INV_TWO_PI, -k.im * QB_INV_TWO_PI);
  const cd aL2 = mk(-0.5 * k.im, 0.5 * k.re);
  const cd aM2 = mk(0.0, 0.5);
  const double aM1 = -QB_INV_TWO_PI;
  const double wj = g.ws[j];
  const double inn = g.dy[j] * dx - g.dx[j] * dy;
  const cd kr = k * r;
  D3 o;
  if (OP == QB_DLP_KRESS) {
    const double cD = inn * invr;
    const double c1 = Rij * cD, c2 = wj * cD, c3 = wj * lt * cD;
    const cd l1 = aL1 * sp.j1;
    const cd D = c1 * l1 + c2 * (aL2 * sp.h1) - c3 * l1;
    const cd kc = k * QB_INV_TWO_PI;
    const cd D1 = kc * ((-Rij * inn) * sp.j0 + wj * ((lt * inn) * sp.j0 + (QB_PI * inn) * (I * sp.h0)));
    const cd u = sp.j0 - kr * sp.j1, v = sp.h0 - kr * sp.h1;
    const cd D2 = (inn * QB_INV_TWO_PI) * ((-Rij) * u + wj * (lt * u + QB_PI * (I * v)));
    o.a = -D;
    o.a1 = -D1;
    o.a2 = -D2;
    return o;
  }
  const double sj = g.speed[j];
  const cd mh = mk(0.0, -0.5), ph = mk(0.0, 0.5);
  if (!same) {
    const double cD = wj * inn * invr, cD1 = wj * inn, cS = wj * sj;
    const cd dval = cD * (aL2 * sp.h1);
    const cd dval1 = off_sign * (ph * (cD1 * (k * sp.h0)));
    const cd dval2 = off_sign * (ph * (cD1 * (sp.h0 - kr * sp.h1)));
    const cd sval = cS * (aM2 * sp.h0);
    const cd sval1 = mh * ((cS * r) * sp.h1);
    const cd sval2 = ph * ((cS * r) * cdiv(sp.h1 - kr * sp.h0, k));
    o.a = -(dval + ik * sval);
    o.a1 = -(dval1 + I * sval + ik * sval1);
    o.a2 = -(dval2 + 2.0 * (I * sval1) + ik * sval2);
    return o;
  }
  const cd L1 = aL1 * sp.j1, L2 = aL2 * sp.h1;
  const cd M1 = aM1 * sp.j0, M2 = aM2 * sp.h0;
  const double cDR = Rij * inn * invr, cDW = wj * inn * invr, cDL = wj * lt * inn * invr;
  const double cR = Rij * inn * QB_INV_TWO_PI, cW = wj * inn * QB_INV_TWO_PI;
  const double cSR = Rij * sj, cS = wj * sj;
  const cd dval = cDR * L1 + cDW * L2 - cDL * L1;
  const cd dval1 = -(cR * (k * sp.j0)) + cW * (k * (lt * sp.j0 + QB_PI * (I * sp.h0)));
  const cd dval2 = cR * (kr * sp.j1 - sp.j0) +
                   cW * (lt * (sp.j0 - kr * sp.j1) + QB_PI * (I * (sp.h0 - kr * sp.h1)));
  const cd sval = cSR * M1 + cS * M2 - (cS * lt) * M1;
  const cd sval1 = (r * sj * QB_INV_TWO_PI) * (Rij * sp.j1 - wj * (lt * sp.j1 + QB_PI * (I * sp.h1)));
  const cd sval2 = (r * sj * QB_INV_TWO_PI) *
                   cdiv(Rij * (kr * sp.j0 - sp.j1) +
                            wj * (lt * (sp.j1 - kr * sp.j0) + QB_PI * (I * (sp.h1 - kr * sp.h0))),
                        k);
  o.a = -(dval + ik * sval);
  o.a1 = -(dval1 + I * sval + ik * sval1);
  o.a2 = -(dval2 + 2.0 * (I * sval1) + ik * sval2);
  return o;
}

// Diagonal entries (cheb-CFIE:922-941; cheb-DLP: D_ii = w_i kappa_i, no
// derivative terms, then F = I - D).
template <int OP>
__device__ __forceinline__ D3 diag_deriv(const DevGeom& g, int64_t j, cd k) {
  D3 o;
  const double wj = g.ws[j];
  if (OP == QB_DLP_KRESS) {
    o.a = mk(1.0 - wj * g.kappa[j], 0.0);
    o.a1 = mk(0, 0);
    o.a2 = mk(0, 0);
    return o;
  }
  const cd I = mk(0.0, 1.0);
  const cd ik = mk(-k.im, k.re);
  const double sj = g.speed[j];
  const double rjj = g.rk[g.off[g.comp[j]]];
  const cd dval = mk(wj * g.kappa[j], 0.0);
  const double m1 = -QB_INV_TWO_PI * sj;
  const cd k2 = k * k;
  const cd lg = clog(mk(0.25 * k2.re * (sj * sj), 0.25 * k2.im * (sj * sj)));
  const cd m2 = mk(-QB_EULER_OVER_PI - QB_INV_TWO_PI * lg.re, 0.5 - QB_INV_TWO_PI * lg.im) * sj;
  const cd sval = mk(rjj * m1, 0.0) + wj * m2;
  const cd sval1 = (-wj * sj / QB_PI) * crecip(k);
  const cd sval2 = (wj * sj / QB_PI) * crecip(k2);
  const cd v = dval + ik * sval;
  o.a = mk(1.0 - v.re, -v.im);
  o.a1 = -(I * sval + ik * sval1);
  o.a2 = -(2.0 * (I * sval1) + ik * sval2);
  return o;
}

// Tile pairs as in k_assemble; the mirrored tile of each of the three
// outputs is staged in LDS 16 source columns at a time (3 x 16 x 65 complex
// = 50 KB) and written 16 contiguous rows per column.
template <int OP, int MH, int MJ>
__global__ void __launch_bounds__(256) k_assemble_deriv(DevGeom g, const cd* __restrict__ zk,
                                                        const cd* __restrict__ H,
                                                        const cd* __restrict__ J, int np_h,
                                                        int M_h, int np_j, int M_j,
                                                        double rmin_h, double rmax, int64_t q0,
                                                        int T, cd* __restrict__ o0,
                                                        cd* __restrict__ o1,
                                                        cd* __restrict__ o2, double off_sign) {
  __shared__ cd mir[3][16][65];  // mir[out][j_local - 16 ch][i_local] = A[j, i]
  const int64_t N = g.N;
  const int64_t q = q0 + blockIdx.z;
  int64_t p = blockIdx.x;
  int d = 0;
  while (p >= T - d) {
    p -= T - d;
    ++d;
  }
  const int Jt = (int)p, It = Jt + d;
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
  const int il = threadIdx.x & 63, jl0 = threadIdx.x >> 6;
  const int64_t i = (int64_t)It * 64 + il;
  const int64_t base = (q - q0) * N * N;
  const bool diag_tile = (It == Jt);
  const bool iok = i < N;
  double xi = 0, yi = 0;
  int ci = 0;
  if (iok) {
    xi = g.x[i];
    yi = g.y[i];
    ci = g.comp[i];
  }
#pragma unroll 1
  for (int ch = 0; ch < 4; ++ch) {
#pragma unroll 1
    for (int c = 0; c < 4; ++c) {
      const int jc = jl0 + 4 * c;  // 0..15 inside the chunk
      const int64_t j = (int64_t)Jt * 64 + 16 * ch + jc;
      if (!iok || j >= N) continue;
      D3 vij, vji;
      if (i == j) {
        vij = diag_deriv<OP>(g, j, k);
      } else {
        const double dx = xi - g.x[j], dy = yi - g.y[j];
        const double r = hypot(dx, dy);
        const double invr = 1.0 / r;
        const int cj = g.comp[j];
        const bool same = (ci == cj);
        double Rij = 0, lt = 0;
        if (same) {
          const int64_t o = g.off[ci];
          const int64_t nc = g.off[ci + 1] - o;
          int64_t dd = (i - j) % nc;
          if (dd < 0) dd += nc;
          Rij = g.rk[o + dd];
          lt = g.lt[o + dd];
        }
        const Spec sp = specials_all<MH, MJ>(tv, r, k, same);
        vij = offdiag_deriv<OP>(g, j, dx, dy, r, invr, same, Rij, lt, k, sp, off_sign);
        if (!diag_tile)
          vji = offdiag_deriv<OP>(g, i, -dx, -dy, r, invr, same, Rij, lt, k, sp, off_sign);
      }
      const int64_t idx = base + i + j * N;
      o0[idx] = vij.a;
      o1[idx] = vij.a1;
      o2[idx] = vij.a2;
      if (!diag_tile) {
        mir[0][jc][il] = vji.a;
        mir[1][jc][il] = vji.a1;
        mir[2][jc][il] = vji.a2;
      }
    }
    if (diag_tile) continue;  // uniform over the workgroup
    __syncthreads();
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int jr = e & 15, il2 = e >> 4;
      const int64_t ii = (int64_t)It * 64 + il2, jj = (int64_t)Jt * 64 + 16 * ch + jr;
      if (ii >= N || jj >= N) continue;
      const int64_t idx = base + jj + ii * N;
      o0[idx] = mir[0][jr][il2];
      o1[idx] = mir[1][jr][il2];
      o2[idx] = mir[2][jr][il2];
    }
    __syncthreads();
  }
}

template <int OP, int MH, int MJ>
static void launch_asm_deriv(qb_ctx* ctx, const qb_geom* g, const cd* z, int64_t q0, int64_t nzc,
                             const DevTables& t, cd* o0, cd* o1, cd* o2, double off_sign) {
  int64_t N = g->d.N;
  int T = (int)((N + 63) / 64);
  int64_t pairs = (int64_t)T * (T + 1) / 2;
  dim3 grid((unsigned)pairs, 1, (unsigned)nzc);
  int slot = qb_prof_begin(ctx, QB_K_ASSEMBLE, 0.0, 3.0 * (double)N * N * nzc * sizeof(cd));
  hipLaunchKernelGGL((k_assemble_deriv<OP, MH, MJ>), grid, dim3(256), 0, ctx->stream, g->d, z,
                     t.H, t.J, t.np_h, t.M_h, t.np_j, t.M_j, t.rmin_h, t.rmax, q0, T, o0, o1, o2,
                     off_sign);
  qb_prof_end(ctx, slot);
}

int qb_assemble_deriv_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_dev, int64_t nz,
                          const DevTables& t, cd* out, cd* dout, cd* ddout, bool ref_sign) {
  const double off_sign = ref_sign ? -1.0 : 1.0;
  if (op != QB_CFIE_KRESS && op != QB_DLP_KRESS)
    return qb_fail(ctx, QB_E_ARG, "unknown operator");
  if (op == QB_DLP_KRESS && g->d.ncomp != 1)
    return qb_fail(ctx, QB_E_ARG, "DLP_kress expects one boundary component");
  if (!t.J) return qb_fail(ctx, QB_E_ARG, "derivative assembly needs the J tables");
  if (g->sym.m > 0 && g->sym.nirr != 1)
    return qb_fail(ctx, QB_E_ARG, "derivative assembly takes one irrep");
  if (g->sym.m > 0)
    return qb_assemble_red_deriv_dev(ctx, g, op, z_dev, nz, t, out, dout, ddout, off_sign);
  int64_t N = g->d.N;
  for (int64_t q0 = 0; q0 < nz; q0 += 1024) {
    int64_t nzc = nz - q0 < 1024 ? nz - q0 : 1024;
    const int64_t o = q0 * N * N;
    const int mh = t.M_h, mj = t.M_j;
    if (op == QB_CFIE_KRESS) {
      if (mh == 7 && mj == 7)
        launch_asm_deriv<QB_CFIE_KRESS, 7, 7>(ctx, g, z_dev, q0, nzc, t, out + o, dout + o,
                                              ddout + o, off_sign);
      else
        launch_asm_deriv<QB_CFIE_KRESS, 0, 0>(ctx, g, z_dev, q0, nzc, t, out + o, dout + o,
                                              ddout + o, off_sign);
    } else {
      if (mh == 7 && mj == 7)
        launch_asm_deriv<QB_DLP_KRESS, 7, 7>(ctx, g, z_dev, q0, nzc, t, out + o, dout + o,
                                             ddout + o, off_sign);
      else
        launch_asm_deriv<QB_DLP_KRESS, 0, 0>(ctx, g, z_dev, q0, nzc, t, out + o, dout + o,
                                             ddout + o, off_sign);
    }
    QB_HIP(ctx, hipGetLastError());
  }
  return QB_OK;
}

// ---------------------------------------------------------------------------
// Symmetry-reduced assembly (geometry with orbits, qb_geom_set_orbits):
//   CFIE: _all_k_symm_CFIE_chebyshev! (cheb-CFIE:698-824) and its derivative
//         twin _all_k_symm_CFIE_chebyshev_deriv! (cheb-CFIE:1087-1243);
//   DLP:  the reduced _construct_dlp_kress_matrices_chebyshev! (cheb-DLP:771-880)
//         and its derivative twin (cheb-DLP:955-1050), then F = I - D.
// Entry (a, b) of the m x m reduced matrix is
//   sum_l scale[b, l] * A(rep_a, img[b, l]),
// A the full operator entry computed exactly as k_assemble / k_assemble_deriv
// compute it (same specials, same Kress terms, identity on the diagonal).
// One thread per reduced entry and wavenumber: a wave holds 64 consecutive
// rows a of one column b, so the images of b are wave-uniform and the stores
// are coalesced.  The workgroup covers a 64 x 64 tile; grid.z walks the
// wavenumbers.
// ---------------------------------------------------------------------------
struct PairGeo {
  double dx, dy, r, invr, Rij, lt;
  bool same;
};

__device__ __forceinline__ PairGeo pair_geo(const DevGeom& g, int64_t i, int64_t j, double xi,
                                            double yi, int ci) {
  PairGeo p;
  p.dx = xi - g.x[j];
  p.dy = yi - g.y[j];
  p.r = hypot(p.dx, p.dy);
  p.invr = 1.0 / p.r;
  p.same = (ci == g.comp[j]);
  p.Rij = 0;
  p.lt = 0;
  if (p.same) {
    const int64_t o = g.off[ci];
    const int64_t nc = g.off[ci + 1] - o;
    int64_t dd = (i - j) % nc;
    if (dd < 0) dd += nc;
    p.Rij = g.rk[o + dd];
    p.lt = g.lt[o + dd];
  }
  return p;
}

__device__ __forceinline__ TabView tab_view(const cd* H, const cd* J, int64_t q, int np_h,
                                            int M_h, int np_j, int M_j, double rmin_h,
                                            double rmax) {
  TabView tv;
  tv.H = H + q * (int64_t)np_h * (2 * (M_h + 1));
  tv.J = J ? J + q * (int64_t)np_j * (2 * (M_j + 1)) : nullptr;
  tv.np_h = np_h;
  tv.M_h = M_h;
  tv.np_j = np_j;
  tv.M_j = M_j;
  tv.rmin_h = rmin_h;
  tv.rmax = rmax;
  return tv;
}

// One thread per reduced entry (a, b); with several irreps (s.nirr, at most
// MAXI) every image term is evaluated once and summed into the entry of each
// irrep with that irrep's weight.  Output matrix e * nz + q (irrep-major).
template <int OP, int MH, int MJ, int MAXI>
__global__ void __launch_bounds__(256) k_assemble_red(DevGeom g, DevSym s, const cd* __restrict__ zk,
                                                      const cd* __restrict__ H,
                                                      const cd* __restrict__ J, int np_h, int M_h,
                                                      int np_j, int M_j, double rmin_h,
                                                      double rmax, int64_t q0, int64_t nz,
                                                      cd* __restrict__ out64,
                                                      cf* __restrict__ out32) {
  const int64_t m = s.m;
  const int G = s.G;
  const int nirr = s.nirr;
  const int64_t q = q0 + blockIdx.z;
  const TabView tv = tab_view(H, J, q, np_h, M_h, np_j, M_j, rmin_h, rmax);
  const cd k = zk[q];
  const int il = threadIdx.x & 63, jl0 = threadIdx.x >> 6;
  const int64_t a = (int64_t)blockIdx.x * 64 + il;
  if (a >= m) return;  // no barrier in this kernel
  const int64_t i = s.img[a * G];
  const double xi = g.x[i], yi = g.y[i];
  const int ci = g.comp[i];
#pragma unroll 1
  for (int c = 0; c < 16; ++c) {
    const int64_t b = (int64_t)blockIdx.y * 64 + jl0 + 4 * c;
    if (b >= m) break;
    cd acc[MAXI];
#pragma unroll
    for (int e = 0; e < MAXI; ++e) acc[e] = mk(0, 0);
#pragma unroll 1
    for (int l = 0; l < G; ++l) {
      const int64_t j = s.img[b * G + l];
      cd v;
      if (i == j) {
        v = diag_entry<OP>(g, j, k);
      } else {
        const PairGeo p = pair_geo(g, i, j, xi, yi, ci);
        const Spec sp = specials<OP, MH, MJ>(tv, p.r, k, p.same || OP == QB_DLP_KRESS);
        v = offdiag_entry<OP>(g, i, j, p.dx, p.dy, p.invr, p.same, p.Rij, p.lt, k, sp);
      }
#pragma unroll
      for (int e = 0; e < MAXI; ++e)
        if (e < nirr) acc[e] = acc[e] + s.scale[((int64_t)e * m + b) * G + l] * v;
    }
#pragma unroll
    for (int e = 0; e < MAXI; ++e) {
      if (e >= nirr) break;
      const int64_t idx = ((int64_t)e * nz + q) * m * m + a + b * m;
      if (out64) out64[idx] = acc[e];
      if (out32) out32[idx] = to_cf(acc[e]);
    }
  }
}

// Pair-shared reduced assembly (geometry with verified isometric orbits,
// DevSym::pair).  Workgroups walk 64 x 64 tile pairs (I, J), I >= J, as
// k_assemble does.  Off the diagonal a thread evaluates the specials of
// r = |x_{rep_a} - x_{img[b,l]}| once per l and uses them for the image term
// of entry (a, b) and for the image term inv[l] of entry (b, a), whose
// distance |x_{rep_b} - x_{img[a,inv[l]]}| is the same; every other factor
// (direction, tangent, component, Kress R and log term) is computed for each
// entry from its own nodes.  Entry (b, a) is staged in LDS and written as
// columns of the mirror tile.  Diagonal tiles are assembled entry by entry.
template <int OP, int MH, int MJ>
__global__ void __launch_bounds__(256) k_assemble_red_pair(DevGeom g, DevSym s,
                                                           const cd* __restrict__ zk,
                                                           const cd* __restrict__ H,
                                                           const cd* __restrict__ J, int np_h,
                                                           int M_h, int np_j, int M_j,
                                                           double rmin_h, double rmax, int64_t q0,
                                                           int T, cd* __restrict__ out64,
                                                           cf* __restrict__ out32) {
  __shared__ cd mir[64][65];  // mir[a_local][b_local] = A_red[b, a]
  const int64_t m = s.m;
  const int G = s.G;
  const int64_t q = q0 + blockIdx.z;
  int64_t p = blockIdx.x;
  int d = 0;
  while (p >= T - d) {
    p -= T - d;
    ++d;
  }
  const int Jt = (int)p, It = Jt + d;
  const TabView tv = tab_view(H, J, q, np_h, M_h, np_j, M_j, rmin_h, rmax);
  const cd k = zk[q];
  const int il = threadIdx.x & 63, jl0 = threadIdx.x >> 6;
  const int64_t a = (int64_t)It * 64 + il;
  const bool aok = a < m;
  const bool diag_tile = (It == Jt);
  const int64_t base = (q - q0) * m * m;
  int64_t i = 0;
  double xi = 0, yi = 0;
  int ci = 0;
  if (aok) {
    i = s.img[a * G];
    xi = g.x[i];
    yi = g.y[i];
    ci = g.comp[i];
  }
#pragma unroll 1
  for (int c = 0; c < 16; ++c) {
    const int jl = jl0 + 4 * c;
    const int64_t b = (int64_t)Jt * 64 + jl;
    if (!aok || b >= m) continue;
    cd acc = mk(0, 0), accm = mk(0, 0);
    if (diag_tile) {
#pragma unroll 1
      for (int l = 0; l < G; ++l) {
        const int64_t j = s.img[b * G + l];
        cd v;
        if (i == j) {
          v = diag_entry<OP>(g, j, k);
        } else {
          const PairGeo pg = pair_geo(g, i, j, xi, yi, ci);
          const Spec sp = specials<OP, MH, MJ>(tv, pg.r, k, pg.same || OP == QB_DLP_KRESS);
          v = offdiag_entry<OP>(g, i, j, pg.dx, pg.dy, pg.invr, pg.same, pg.Rij, pg.lt, k, sp);
        }
        acc = acc + s.scale[b * G + l] * v;
      }
    } else {
      const int64_t ib = s.img[b * G];
      const double xb = g.x[ib], yb = g.y[ib];
      const int cb = g.comp[ib];
#pragma unroll 1
      for (int l = 0; l < G; ++l) {
        const int64_t j = s.img[b * G + l];
        const int ls = s.inv[l];
        const int64_t jm = s.img[a * G + ls];
        const PairGeo p1 = pair_geo(g, i, j, xi, yi, ci);
        const PairGeo p2 = pair_geo(g, ib, jm, xb, yb, cb);  // its r equals p1.r
        const Spec sp = specials<OP, MH, MJ>(tv, p1.r, k,
                                             p1.same || p2.same || OP == QB_DLP_KRESS);
        const cd v1 = offdiag_entry<OP>(g, i, j, p1.dx, p1.dy, p1.invr, p1.same, p1.Rij, p1.lt,
                                        k, sp);
        const cd v2 = offdiag_entry<OP>(g, ib, jm, p2.dx, p2.dy, p1.invr, p2.same, p2.Rij,
                                        p2.lt, k, sp);
        acc = acc + s.scale[b * G + l] * v1;
        accm = accm + s.scale[a * G + ls] * v2;
      }
      mir[il][jl] = accm;
    }
    const int64_t idx = base + a + b * m;
    if (out64) out64[idx] = acc;
    if (out32) out32[idx] = to_cf(acc);
  }
  if (diag_tile) return;  // uniform over the workgroup
  __syncthreads();
#pragma unroll 1
  for (int c = 0; c < 16; ++c) {
    const int il2 = jl0 + 4 * c;       // column a of the mirror tile
    const int jl2 = threadIdx.x & 63;  // row b inside it
    const int64_t aa = (int64_t)It * 64 + il2, bb = (int64_t)Jt * 64 + jl2;
    if (aa >= m || bb >= m) continue;
    const cd v = mir[il2][jl2];
    const int64_t idx = base + bb + aa * m;
    if (out64) out64[idx] = v;
    if (out32) out32[idx] = to_cf(v);
  }
}

// Largest mismatch | |x_{rep_a} - x_{img[b,l]}| - |x_{rep_b} - x_{img[a,inv[l]]}| |
// / (1e-13 r + floor_abs) over one 64 x 64 tile of (a, b) and all l, one value
// per workgroup (<= 1: the distances mirror).  floor_abs is a few ulps of the
// coordinate scale, the rounding every distance carries from its nodes.
__global__ void __launch_bounds__(256) k_orbit_check(DevGeom g, DevSym s, double floor_abs,
                                                     double* out) {
  __shared__ double red[256];
  const int64_t m = s.m;
  const int G = s.G;
  const int il = threadIdx.x & 63, jl0 = threadIdx.x >> 6;
  const int64_t a = (int64_t)blockIdx.x * 64 + il;
  double worst = 0.0;
  if (a < m) {
    const int64_t i = s.img[a * G];
    for (int c = 0; c < 16; ++c) {
      const int64_t b = (int64_t)blockIdx.y * 64 + jl0 + 4 * c;
      if (b >= m) break;
      const int64_t ib = s.img[b * G];
      for (int l = 0; l < G; ++l) {
        const int64_t j = s.img[b * G + l];
        const int64_t jm = s.img[a * G + s.inv[l]];
        const double r1 = hypot(g.x[i] - g.x[j], g.y[i] - g.y[j]);
        const double r2 = hypot(g.x[ib] - g.x[jm], g.y[ib] - g.y[jm]);
        const double dd = fabs(r1 - r2) / (1e-13 * r1 + floor_abs);
        worst = fmax(worst, (i == j) != (ib == jm) ? 1e300 : dd);
      }
    }
  }
  red[threadIdx.x] = worst;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

int qb_orbit_pairs_ok(qb_ctx* ctx, const DevGeom& g, const DevSym& s, double floor_abs, int* ok) {
  *ok = 0;
  const unsigned Tm = (unsigned)((s.m + 63) / 64);
  const size_t nb = (size_t)Tm * Tm;
  double* d = nullptr;
  int rc = qb_dev_alloc(ctx, (void**)&d, nb * sizeof(double));
  if (rc) return rc;
  hipLaunchKernelGGL(k_orbit_check, dim3(Tm, Tm), dim3(256), 0, ctx->stream, g, s, floor_abs, d);
  std::vector<double> h(nb);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), d, nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  qb_dev_free(ctx, d, nb * sizeof(double));
  if (e != hipSuccess) return qb_fail(ctx, QB_E_HIP, std::string("orbit check: ") + hipGetErrorString(e));
  double worst = 0.0;
  for (double v : h) worst = v > worst ? v : worst;
  *ok = worst <= 1.0 ? 1 : 0;
  return QB_OK;
}

template <int OP, int MH, int MJ>
__global__ void __launch_bounds__(256) k_assemble_red_deriv(
    DevGeom g, DevSym s, const cd* __restrict__ zk, const cd* __restrict__ H,
    const cd* __restrict__ J, int np_h, int M_h, int np_j, int M_j, double rmin_h, double rmax,
    int64_t q0, cd* __restrict__ o0, cd* __restrict__ o1, cd* __restrict__ o2, double off_sign) {
  const int64_t m = s.m;
  const int G = s.G;
  const int64_t q = q0 + blockIdx.z;
  const TabView tv = tab_view(H, J, q, np_h, M_h, np_j, M_j, rmin_h, rmax);
  const cd k = zk[q];
  const int il = threadIdx.x & 63, jl0 = threadIdx.x >> 6;
  const int64_t a = (int64_t)blockIdx.x * 64 + il;
  if (a >= m) return;
  const int64_t i = s.img[a * G];
  const double xi = g.x[i], yi = g.y[i];
  const int ci = g.comp[i];
  const int64_t base = (q - q0) * m * m;
#pragma unroll 1
  for (int c = 0; c < 16; ++c) {
    const int64_t b = (int64_t)blockIdx.y * 64 + jl0 + 4 * c;
    if (b >= m) break;
    cd a0 = mk(0, 0), a1 = mk(0, 0), a2 = mk(0, 0);
#pragma unroll 1
    for (int l = 0; l < G; ++l) {
      const int64_t j = s.img[b * G + l];
      const cd sc = s.scale[b * G + l];
      D3 v;
      if (i == j) {
        v = diag_deriv<OP>(g, j, k);
      } else {
        const PairGeo p = pair_geo(g, i, j, xi, yi, ci);
        const Spec sp = specials_all<MH, MJ>(tv, p.r, k, p.same);
        v = offdiag_deriv<OP>(g, j, p.dx, p.dy, p.r, p.invr, p.same, p.Rij, p.lt, k, sp,
                              off_sign);
      }
      a0 = a0 + sc * v.a;
      a1 = a1 + sc * v.a1;
      a2 = a2 + sc * v.a2;
    }
    const int64_t idx = base + a + b * m;
    o0[idx] = a0;
    o1[idx] = a1;
    o2[idx] = a2;
  }
}

// Multi-irrep twin of k_assemble_red_pair: one specials evaluation per image
// term serves entry (a, b) and entry (b, a) of all nirr irreps at once.  The
// nirr mirror values of a thread are written directly as rows of the mirror
// tiles: 64 lanes store 64 scattered 16-byte values per instruction, and the
// L2 cache merges the 64 rows of every mirror column that the 4 waves of the
// block write into whole lines.  Diagonal tiles are assembled entry by entry.
// Outputs indexed as in k_assemble_red (matrix e * nz + q from the base).
template <int OP, int MH, int MJ, int MAXI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_assemble_red_pair_multi(DevGeom g, DevSym s, const cd* __restrict__ zk,
                          const cd* __restrict__ H, const cd* __restrict__ J, int np_h, int M_h,
                          int np_j, int M_j, double rmin_h, double rmax, int64_t q0, int64_t nz,
                          int T, cd* __restrict__ out64, cf* __restrict__ out32) {
  const int64_t m = s.m;
  const int G = s.G;
  const int nirr = s.nirr;
  const int64_t q = q0 + blockIdx.z;
  int64_t p = blockIdx.x;
  int d = 0;
  while (p >= T - d) {
    p -= T - d;
    ++d;
  }
  const int Jt = (int)p, It = Jt + d;
  const TabView tv = tab_view(H, J, q, np_h, M_h, np_j, M_j, rmin_h, rmax);
  const cd k = zk[q];
  const int il = threadIdx.x & 63, jl0 = threadIdx.x >> 6;
  const int64_t a = (int64_t)It * 64 + il;
  if (a >= m) return;  // no barrier in this kernel
  const bool diag_tile = (It == Jt);
  const int64_t mm = m * m;
  const int64_t i = s.img[a * G];
  const double xi = g.x[i], yi = g.y[i];
  const int ci = g.comp[i];
#pragma unroll 1
  for (int c = 0; c < 16; ++c) {
    const int64_t b = (int64_t)Jt * 64 + jl0 + 4 * c;
    if (b >= m) break;
    cd acc[MAXI], accm[MAXI];
#pragma unroll
    for (int e = 0; e < MAXI; ++e) acc[e] = accm[e] = mk(0, 0);
    if (diag_tile) {
#pragma unroll 1
      for (int l = 0; l < G; ++l) {
        const int64_t j = s.img[b * G + l];
        cd v;
        if (i == j) {
          v = diag_entry<OP>(g, j, k);
        } else {
          const PairGeo pg = pair_geo(g, i, j, xi, yi, ci);
          const Spec sp = specials<OP, MH, MJ>(tv, pg.r, k, pg.same || OP == QB_DLP_KRESS);
          v = offdiag_entry<OP>(g, i, j, pg.dx, pg.dy, pg.invr, pg.same, pg.Rij, pg.lt, k, sp);
        }
#pragma unroll
        for (int e = 0; e < MAXI; ++e)
          if (e < nirr) acc[e] = acc[e] + s.scale[((int64_t)e * m + b) * G + l] * v;
      }
    } else {
      const int64_t ib = s.img[b * G];
      const double xb = g.x[ib], yb = g.y[ib];
      const int cb = g.comp[ib];
#pragma unroll 1
      for (int l = 0; l < G; ++l) {
        const int64_t j = s.img[b * G + l];
        const int ls = s.inv[l];
        const int64_t jm = s.img[a * G + ls];
        const PairGeo p1 = pair_geo(g, i, j, xi, yi, ci);
        const PairGeo p2 = pair_geo(g, ib, jm, xb, yb, cb);  // its r equals p1.r
        const Spec sp = specials<OP, MH, MJ>(tv, p1.r, k,
                                             p1.same || p2.same || OP == QB_DLP_KRESS);
        const cd v1 = offdiag_entry<OP>(g, i, j, p1.dx, p1.dy, p1.invr, p1.same, p1.Rij, p1.lt,
                                        k, sp);
        const cd v2 = offdiag_entry<OP>(g, ib, jm, p2.dx, p2.dy, p1.invr, p2.same, p2.Rij,
                                        p2.lt, k, sp);
#pragma unroll
        for (int e = 0; e < MAXI; ++e)
          if (e < nirr) {
            acc[e] = acc[e] + s.scale[((int64_t)e * m + b) * G + l] * v1;
            accm[e] = accm[e] + s.scale[((int64_t)e * m + a) * G + ls] * v2;
          }
      }
    }
#pragma unroll
    for (int e = 0; e < MAXI; ++e) {
      if (e >= nirr) break;
      const int64_t idx = ((int64_t)e * nz + q) * mm + a + b * m;
      if (out64) out64[idx] = acc[e];
      if (out32) out32[idx] = to_cf(acc[e]);
      if (!diag_tile) {
        const int64_t idt = ((int64_t)e * nz + q) * mm + b + a * m;
        if (out64) out64[idt] = accm[e];
        if (out32) out32[idt] = to_cf(accm[e]);
      }
    }
  }
}

// o64 / o32: the output base (matrix 0); the kernels index matrix e * nz + q.
template <int OP, int MH, int MJ>
static void launch_red(qb_ctx* ctx, const qb_geom* g, const cd* z, int64_t q0, int64_t nzc,
                       int64_t nz, const DevTables& t, cd* o64, cf* o32) {
  const int64_t m = g->sym.m;
  const int nirr = g->sym.nirr;
  const unsigned Tm = (unsigned)((m + 63) / 64);
  const double ent = (double)m * m * nzc * nirr;
  int slot = qb_prof_begin(ctx, QB_K_ASSEMBLE, 0.0,
                           ent * ((o64 ? sizeof(cd) : 0) + (o32 ? sizeof(cf) : 0)));
  const dim3 grid(Tm, Tm, (unsigned)nzc);
  if (g->sym.pair && nirr == 1) {
    const int64_t pairs = (int64_t)Tm * (Tm + 1) / 2;
    hipLaunchKernelGGL((k_assemble_red_pair<OP, MH, MJ>), dim3((unsigned)pairs, 1, (unsigned)nzc),
                       dim3(256), 0, ctx->stream, g->d, g->sym, z, t.H, t.J, t.np_h, t.M_h,
                       t.np_j, t.M_j, t.rmin_h, t.rmax, q0, (int)Tm,
                       o64 ? o64 + q0 * m * m : nullptr, o32 ? o32 + q0 * m * m : nullptr);
  } else if (nirr == 1) {
    hipLaunchKernelGGL((k_assemble_red<OP, MH, MJ, 1>), grid, dim3(256), 0, ctx->stream, g->d,
                       g->sym, z, t.H, t.J, t.np_h, t.M_h, t.np_j, t.M_j, t.rmin_h, t.rmax, q0,
                       nz, o64, o32);
  } else if (g->sym.pair && nirr <= 8) {
    const int64_t pairs = (int64_t)Tm * (Tm + 1) / 2;
    hipLaunchKernelGGL((k_assemble_red_pair_multi<OP, MH, MJ, 8>),
                       dim3((unsigned)pairs, 1, (unsigned)nzc), dim3(256), 0, ctx->stream, g->d,
                       g->sym, z, t.H, t.J, t.np_h, t.M_h, t.np_j, t.M_j, t.rmin_h, t.rmax, q0,
                       nz, (int)Tm, o64, o32);
  } else if (nirr <= 8) {
    hipLaunchKernelGGL((k_assemble_red<OP, MH, MJ, 8>), grid, dim3(256), 0, ctx->stream, g->d,
                       g->sym, z, t.H, t.J, t.np_h, t.M_h, t.np_j, t.M_j, t.rmin_h, t.rmax, q0,
                       nz, o64, o32);
  } else {
    hipLaunchKernelGGL((k_assemble_red<OP, MH, MJ, QB_MAX_IRREPS>), grid, dim3(256), 0,
                       ctx->stream, g->d, g->sym, z, t.H, t.J, t.np_h, t.M_h, t.np_j, t.M_j,
                       t.rmin_h, t.rmax, q0, nz, o64, o32);
  }
  qb_prof_end(ctx, slot);
}

template <int OP, int MH, int MJ>
static void launch_red_deriv(qb_ctx* ctx, const qb_geom* g, const cd* z, int64_t q0,
                             int64_t nzc, const DevTables& t, cd* o0, cd* o1, cd* o2,
                             double off_sign) {
  const int64_t m = g->sym.m;
  const unsigned Tm = (unsigned)((m + 63) / 64);
  int slot = qb_prof_begin(ctx, QB_K_ASSEMBLE, 0.0, 3.0 * (double)m * m * nzc * sizeof(cd));
  hipLaunchKernelGGL((k_assemble_red_deriv<OP, MH, MJ>), dim3(Tm, Tm, (unsigned)nzc), dim3(256),
                     0, ctx->stream, g->d, g->sym, z, t.H, t.J, t.np_h, t.M_h, t.np_j, t.M_j,
                     t.rmin_h, t.rmax, q0, o0, o1, o2, off_sign);
  qb_prof_end(ctx, slot);
}

int qb_assemble_red_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_dev, int64_t nz,
                        const DevTables& t, cd* out64, cf* out32) {
  for (int64_t q0 = 0; q0 < nz; q0 += 1024) {
    const int64_t nzc = nz - q0 < 1024 ? nz - q0 : 1024;
    const int mh = t.M_h, mj = t.M_j;
    if (op == QB_CFIE_KRESS) {
      if (mh == 7 && mj == 7)
        launch_red<QB_CFIE_KRESS, 7, 7>(ctx, g, z_dev, q0, nzc, nz, t, out64, out32);
      else if (mh == 5 && mj == 5)
        launch_red<QB_CFIE_KRESS, 5, 5>(ctx, g, z_dev, q0, nzc, nz, t, out64, out32);
      else
        launch_red<QB_CFIE_KRESS, 0, 0>(ctx, g, z_dev, q0, nzc, nz, t, out64, out32);
    } else {
      if (mh == 7 && mj == 7)
        launch_red<QB_DLP_KRESS, 7, 7>(ctx, g, z_dev, q0, nzc, nz, t, out64, out32);
      else if (mh == 5 && mj == 5)
        launch_red<QB_DLP_KRESS, 5, 5>(ctx, g, z_dev, q0, nzc, nz, t, out64, out32);
      else
        launch_red<QB_DLP_KRESS, 0, 0>(ctx, g, z_dev, q0, nzc, nz, t, out64, out32);
    }
    QB_HIP(ctx, hipGetLastError());
  }
  return QB_OK;
}

int qb_assemble_red_deriv_dev(qb_ctx* ctx, const qb_geom* g, int op, const cd* z_dev,
                              int64_t nz, const DevTables& t, cd* out, cd* dout, cd* ddout,
                              double off_sign) {
  const int64_t m = g->sym.m;
  for (int64_t q0 = 0; q0 < nz; q0 += 1024) {
    const int64_t nzc = nz - q0 < 1024 ? nz - q0 : 1024;
    const int64_t o = q0 * m * m;
    const int mh = t.M_h, mj = t.M_j;
    if (op == QB_CFIE_KRESS) {
      if (mh == 7 && mj == 7)
        launch_red_deriv<QB_CFIE_KRESS, 7, 7>(ctx, g, z_dev, q0, nzc, t, out + o, dout + o,
                                              ddout + o, off_sign);
      else
        launch_red_deriv<QB_CFIE_KRESS, 0, 0>(ctx, g, z_dev, q0, nzc, t, out + o, dout + o,
                                              ddout + o, off_sign);
    } else {
      if (mh == 7 && mj == 7)
        launch_red_deriv<QB_DLP_KRESS, 7, 7>(ctx, g, z_dev, q0, nzc, t, out + o, dout + o,
                                             ddout + o, off_sign);
      else
        launch_red_deriv<QB_DLP_KRESS, 0, 0>(ctx, g, z_dev, q0, nzc, t, out + o, dout + o,
                                             ddout + o, off_sign);
    }
    QB_HIP(ctx, hipGetLastError());
  }
  return QB_OK;
}
