// Device-side Chebyshev panel lookup and Clenshaw evaluation.
#pragma once
#include "qb_complex.h"

// Uniform panel lookup (_find_panel_uniform, chebyshev_bessels.jl:574-577)
// and local coordinate t = (2r - (b+a)) / (b-a) with the breaks of
// _breaks_uniform (chebyshev_core.jl:95-103).  p is 0-based here.
QB_HD void qb_panel(double r, double rmin, double rmax, int np, int* p_out,
                    double* t_out) {
  double h = (rmax - rmin) / (double)np;
  double invdr = 1.0 / h;
  int p = (int)floor((r - rmin) * invdr);
  p = p < 0 ? 0 : (p > np - 1 ? np - 1 : p);
  double a = fma((double)p, h, rmin);
  double b = (p == np - 1) ? rmax : fma((double)(p + 1), h, rmin);
  *p_out = p;
  *t_out = (2.0 * r - (b + a)) / (b - a);
}

// Clenshaw for two interleaved coefficient series (the {f0, f1} pairs of one
// panel): c[2m] belongs to the first function, c[2m+1] to the second
// (_cheb_clenshaw, chebyshev_core.jl:287-314).
template <int M>
QB_HD void qb_clenshaw2(const cd* __restrict__ c, double t, cd* f0, cd* f1) {
  cd a1 = mk(0, 0), a2 = mk(0, 0), b1 = mk(0, 0), b2 = mk(0, 0);
  double u = 2.0 * t;
#pragma unroll
  for (int m = M; m >= 1; --m) {
    cd ca = c[2 * m], cb = c[2 * m + 1];
    cd a0 = mk(fma(u, a1.re, -a2.re) + ca.re, fma(u, a1.im, -a2.im) + ca.im);
    cd b0 = mk(fma(u, b1.re, -b2.re) + cb.re, fma(u, b1.im, -b2.im) + cb.im);
    a2 = a1; a1 = a0;
    b2 = b1; b1 = b0;
  }
  cd ca = c[0], cb = c[1];
  *f0 = mk(fma(t, a1.re, -a2.re) + ca.re, fma(t, a1.im, -a2.im) + ca.im);
  *f1 = mk(fma(t, b1.re, -b2.re) + cb.re, fma(t, b1.im, -b2.im) + cb.im);
}

QB_HD void qb_clenshaw2_dyn(const cd* __restrict__ c, int M, double t, cd* f0, cd* f1) {
  cd a1 = mk(0, 0), a2 = mk(0, 0), b1 = mk(0, 0), b2 = mk(0, 0);
  double u = 2.0 * t;
  for (int m = M; m >= 1; --m) {
    cd ca = c[2 * m], cb = c[2 * m + 1];
    cd a0 = mk(fma(u, a1.re, -a2.re) + ca.re, fma(u, a1.im, -a2.im) + ca.im);
    cd b0 = mk(fma(u, b1.re, -b2.re) + cb.re, fma(u, b1.im, -b2.im) + cb.im);
    a2 = a1; a1 = a0;
    b2 = b1; b1 = b0;
  }
  cd ca = c[0], cb = c[1];
  *f0 = mk(fma(t, a1.re, -a2.re) + ca.re, fma(t, a1.im, -a2.im) + ca.im);
  *f1 = mk(fma(t, b1.re, -b2.re) + cb.re, fma(t, b1.im, -b2.im) + cb.im);
}
