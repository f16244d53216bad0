// Host build of the device Bessel evaluator (quantumbilliards.jl_amd/csrc/qb_bessel.h)
// so the CPU test suite can compare it with SciPy/mpmath values.
#include "../../quantumbilliards.jl_amd/csrc/qb_bessel.h"
#include <stdint.h>

extern "C" void qbt_bessel_host(const double* z, int64_t n, double* j0, double* j1,
                                double* h0, double* h1) {
  for (int64_t i = 0; i < n; ++i) {
    cd zz = mk(z[2 * i], z[2 * i + 1]);
    cd a, b, c, d;
    qb_bessel_j01(zz, &a, &b);
    j0[2 * i] = a.re; j0[2 * i + 1] = a.im;
    j1[2 * i] = b.re; j1[2 * i + 1] = b.im;
    if (zz.re > 0.0) {
      qb_hankel1_01(zz, &c, &d);
    } else {
      c = mk(NAN, NAN); d = mk(NAN, NAN);
    }
    h0[2 * i] = c.re; h0[2 * i + 1] = c.im;
    h1[2 * i] = d.re; h1[2 * i + 1] = d.im;
  }
}
