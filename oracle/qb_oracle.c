/* CPU oracle, C part: single-component CFIE-Kress assembly of a block of
 * columns with OpenMP, given the Chebyshev tables.  TEST INFRASTRUCTURE ONLY
 * (the cpu_baseline leg of bench.py and tests); the product never links it.
 *
 * Restates _all_k_nosymm_CFIE_chebyshev! / same_block_col!
 * (src/chebyshev/chebyshev_euclidian_helmholtz_cfie_kress.jl:543-615) for one
 * wavenumber: panel lookup (chebyshev_bessels.jl:574-577), Clenshaw
 * (chebyshev_core.jl:287-314), the Kress split, the diagonal limits.
 * Tables: H[p][m][f][re,im] with f = 0 (H0), 1 (H1); J likewise (J0, J1).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>

static const double INV_TWO_PI = 0.15915494309189533577;

static void panel(double r, double rmin, double rmax, int np, int *p, double *t) {
  double h = (rmax - rmin) / np;
  int q = (int)floor((r - rmin) * (1.0 / h));
  q = q < 0 ? 0 : (q > np - 1 ? np - 1 : q);
  double a = q * h + rmin;
  double b = (q == np - 1) ? rmax : (q + 1) * h + rmin;
  *p = q;
  *t = (2.0 * r - (b + a)) / (b - a);
}

static double complex clen(const double *c, int M, int f, double t) {
  double complex b1 = 0, b2 = 0;
  double u = 2.0 * t;
  for (int m = M; m >= 1; --m) {
    const double *e = c + 4 * m + 2 * f;
    double complex b0 = u * b1 - b2 + (e[0] + I * e[1]);
    b2 = b1;
    b1 = b0;
  }
  return t * b1 - b2 + (c[2 * f] + I * c[2 * f + 1]);
}

void qbo_cfie_cols(int64_t n, const double *x, const double *y, const double *dx,
                   const double *dy, const double *ddx, const double *ddy, const double *ws,
                   const double *rkcol, const double *ltcol, int64_t j0, int64_t j1, double zre,
                   double zim, const double *H, int np_h, int M_h, double rmin_h, double rmax,
                   const double *J, int np_j, int M_j, double *out) {
  const double complex k = zre + I * zim;
  const double complex ik = I * k;
  const double complex aL1 = -k * INV_TWO_PI, aL2 = 0.5 * I * k, aM2 = 0.5 * I;
  const double aM1 = -INV_TWO_PI;
  const double euler_over_pi = 0.57721566490153286061 / 3.14159265358979323846;
#pragma omp parallel for schedule(static)
  for (int64_t j = j0; j < j1; ++j) {
    double sj = sqrt(dx[j] * dx[j] + dy[j] * dy[j]);
    double wj = ws[j];
    double complex *col = (double complex *)(out + 2 * (j - j0) * n);
    for (int64_t i = 0; i < n; ++i) {
      if (i == j) {
        double kap = INV_TWO_PI * (-(dx[j] * ddy[j] - dy[j] * ddx[j]) / (sj * sj));
        double complex m2 = ((0.5 * I - euler_over_pi) -
                             INV_TWO_PI * clog((k * k / 4.0) * (sj * sj))) * sj;
        double complex sval = rkcol[0] * (aM1 * sj) + wj * m2;
        col[i] = 1.0 - (wj * kap + ik * sval);
        continue;
      }
      double ddx_ = x[i] - x[j], ddy_ = y[i] - y[j];
      double r = hypot(ddx_, ddy_), invr = 1.0 / r;
      double inn = dy[j] * ddx_ - dx[j] * ddy_;
      int64_t d = ((i - j) % n + n) % n;
      double Rij = rkcol[d], lt = ltcol[d];
      int p;
      double t;
      double complex h0, h1, jj0, jj1;
      if (cabs(k * r) < 0.2) {
        /* not reached by the timed workloads (k r >= 0.5 there) */
        h0 = h1 = NAN;
      } else {
        panel(r, rmin_h, rmax, np_h, &p, &t);
        const double *c = H + (int64_t)p * 4 * (M_h + 1);
        h0 = clen(c, M_h, 0, t);
        h1 = clen(c, M_h, 1, t);
      }
      panel(r, 0.0, rmax, np_j, &p, &t);
      const double *cj = J + (int64_t)p * 4 * (M_j + 1);
      jj0 = clen(cj, M_j, 0, t);
      jj1 = clen(cj, M_j, 1, t);
      double complex L1 = aL1 * jj1, L2 = aL2 * h1, M1 = aM1 * jj0, M2 = aM2 * h0;
      double cD = inn * invr;
      double complex D = (Rij * cD) * L1 + (wj * cD) * L2 - (wj * lt * cD) * L1;
      double complex S = (Rij * sj) * M1 + (wj * sj) * M2 - (wj * sj * lt) * M1;
      col[i] = -(D + ik * S);
    }
  }
}
