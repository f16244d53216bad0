// Complex-argument Bessel J0, J1 and Hankel H0^(1), H1^(1) in FP64.
//
// Replaces the AMOS calls (SpecialFunctions.besselh / besselj) that the
// reference makes when it samples its Chebyshev tables
// (src/chebyshev/chebyshev_bessels.jl:331-344, :360-372) and in its direct
// small-|z| branch (chebyshev_bessels.jl:1197-1199).
//
// Three regions in rho = |z|, each a published method:
//  * rho <= QB_SERIES_MAX: ascending power series (DLMF 10.2.2, 10.8.2;
//    Abramowitz & Stegun 9.1.10-9.1.13).  For |z| < 1e-3 this is the same
//    truncated expansion the reference uses there (bessels:729-756).
//  * QB_SERIES_MAX < rho < QB_ASYM_MIN: Miller backward recurrence for J_n
//    normalised by J0 + 2 sum J_2k = 1 (DLMF 10.12.4, 10.74(iv)) and
//    Neumann expansions for Y0, Y1 (DLMF 10.23.16 with n = 0, 1).
//  * rho >= QB_ASYM_MIN: Hankel asymptotic expansions (DLMF 10.17.5-6),
//    J = (H1 + H2) / 2.
// Valid for Re z > 0 for the Hankel functions (the z = k r of a contour in
// the right half plane); J is entire and valid everywhere the regions apply.
#pragma once
#include "qb_complex.h"

#define QB_SERIES_MAX 2.0
#define QB_ASYM_MIN 20.0
#define QB_EULER_GAMMA 0.57721566490153286061
#define QB_PI 3.14159265358979323846
#define QB_2_OVER_PI 0.63661977236758134308
#define QB_SQRT_2_OVER_PI 0.79788456080286535588

// Ascending series for J0, J1 (and optionally Y0, Y1).
QB_HD void qb_bessel_series(cd z, bool want_y, cd* J0, cd* J1, cd* Y0, cd* Y1) {
  cd w = 0.5 * z;          // z/2
  cd q = -(w * w);         // -z^2/4
  cd t0 = mk(1.0, 0.0);    // q^k / (k!)^2
  cd s1 = mk(1.0, 0.0);    // q^k / (k! (k+1)!)
  cd j0 = t0, j1s = s1;
  cd y0s = mk(0.0, 0.0);   // sum_{k>=1} H_k q^k/(k!)^2
  cd y1s = s1;             // sum_{k>=0} (H_k + H_{k+1}) q^k/(k!(k+1)!), k=0 term = 1
  double Hk = 0.0;
  for (int k = 1; k <= 30; ++k) {
    double dk = (double)k;
    t0 = t0 * q * (1.0 / (dk * dk));
    s1 = s1 * q * (1.0 / (dk * (dk + 1.0)));
    Hk += 1.0 / dk;
    j0 = j0 + t0;
    j1s = j1s + s1;
    y0s = y0s + Hk * t0;
    y1s = y1s + (2.0 * Hk + 1.0 / (dk + 1.0)) * s1;
    if (cabs2(t0) < 1e-36 * cabs2(j0) && cabs2(s1) < 1e-36 * cabs2(j1s)) break;
  }
  *J0 = j0;
  *J1 = w * j1s;
  if (want_y) {
    cd lw = clog(w) + mk(QB_EULER_GAMMA, 0.0);  // ln(z/2) + gamma
    *Y0 = QB_2_OVER_PI * (lw * j0 - y0s);
    cd invw = crecip(w);
    *Y1 = (-1.0 / QB_PI) * invw + QB_2_OVER_PI * (lw * (*J1)) - (1.0 / QB_PI) * (w * y1s);
  }
}

// Miller backward recurrence for J0, J1 (and Neumann series for Y0, Y1).
QB_HD void qb_bessel_miller(cd z, bool want_y, cd* J0, cd* J1, cd* Y0, cd* Y1) {
  double rho = cabsd(z);
  int n0 = (int)(rho + 9.0 * cbrt(rho) + 20.0);
  n0 += (n0 & 1);  // even start
  cd iz = crecip(z);
  cd fp1 = mk(0.0, 0.0);      // f_{n+1}
  cd f = mk(1e-30, 0.0);      // f_n, n = n0
  cd norm = mk(0.0, 0.0);     // f_0 + 2 sum f_2k
  cd s0 = mk(0.0, 0.0);       // sum_{k>=1} (-1)^k f_2k / k
  cd s1 = mk(0.0, 0.0);       // sum_{k>=1} (-1)^k (2k+1) f_{2k+1} / (k (k+1))
  cd f1 = mk(0.0, 0.0);
  for (int n = n0; n >= 1; --n) {
    // f holds f_n; account for it, then step to f_{n-1}.
    if ((n & 1) == 0) {
      int k = n >> 1;
      norm = norm + 2.0 * f;
      double sg = (k & 1) ? -1.0 : 1.0;
      s0 = s0 + (sg / (double)k) * f;
    } else if (n >= 3) {
      int k = (n - 1) >> 1;
      double sg = (k & 1) ? -1.0 : 1.0;
      s1 = s1 + (sg * (double)n / ((double)k * (double)(k + 1))) * f;
    }
    if (n == 1) f1 = f;
    cd fm1 = (2.0 * (double)n) * (iz * f) - fp1;
    fp1 = f;
    f = fm1;
  }
  norm = norm + f;  // f_0
  cd inv = crecip(norm);
  cd j0 = f * inv, j1 = f1 * inv;
  *J0 = j0;
  *J1 = j1;
  if (want_y) {
    cd lw = clog(0.5 * z) + mk(QB_EULER_GAMMA, 0.0);
    *Y0 = QB_2_OVER_PI * (lw * j0) - (2.0 * QB_2_OVER_PI) * (s0 * inv);
    *Y1 = (-QB_2_OVER_PI) * (j0 * iz) + QB_2_OVER_PI * ((lw - mk(1.0, 0.0)) * j1) -
          QB_2_OVER_PI * (s1 * inv);
  }
}

// Hankel asymptotic expansion for nu = 0, 1.  Returns H^(1)_nu and, if
// wanted, H^(2)_nu.
QB_HD void qb_hankel_asym(cd z, bool want_h2, cd* H10, cd* H11, cd* H20, cd* H21) {
  cd u = cdiv(mk(0.0, 1.0), z);  // i / z
  // even/odd partial sums of sum a_k(nu) u^k for nu = 0, 1
  cd e0 = mk(1.0, 0.0), o0 = mk(0.0, 0.0), e1 = mk(1.0, 0.0), o1 = mk(0.0, 0.0);
  cd up = mk(1.0, 0.0);
  double a0 = 1.0, a1 = 1.0;
  for (int k = 1; k <= 40; ++k) {
    double m = (double)(2 * k - 1);
    a0 *= (0.0 - m * m) / (8.0 * (double)k);
    a1 *= (4.0 - m * m) / (8.0 * (double)k);
    up = up * u;
    cd t0 = a0 * up, t1 = a1 * up;
    if (k & 1) {
      o0 = o0 + t0;
      o1 = o1 + t1;
    } else {
      e0 = e0 + t0;
      e1 = e1 + t1;
    }
    if (cabs2(t0) < 1e-36 && cabs2(t1) < 1e-36) break;
  }
  // sqrt(2/(pi z)) e^{i(z - pi/4)}; e^{-i pi/4} applied as a constant so the
  // large real part of z is reduced exactly by sincos.
  cd pre = cdiv(mk(QB_SQRT_2_OVER_PI, 0.0), csqrt(z));
  double sr, cr;
  sincos(z.re, &sr, &cr);
  const double c4 = 0.70710678118654752440;
  double dn = exp(-z.im);
  cd ph = mk(dn * cr, dn * sr) * mk(c4, -c4);  // e^{iz} e^{-i pi/4}
  cd p1 = pre * ph;
  *H10 = p1 * (e0 + o0);
  *H11 = mk(0.0, -1.0) * (p1 * (e1 + o1));  // extra e^{-i pi/2}
  if (want_h2) {
    double up_ = exp(z.im);
    cd ph2 = mk(up_ * cr, -up_ * sr) * mk(c4, c4);  // e^{-iz} e^{i pi/4}
    cd p2 = pre * ph2;
    *H20 = p2 * (e0 - o0);
    *H21 = mk(0.0, 1.0) * (p2 * (e1 - o1));
  }
}

// J0(z), J1(z) for any z (regions by |z|).
QB_HD void qb_bessel_j01(cd z, cd* J0, cd* J1) {
  double rho = cabsd(z);
  cd d0, d1;
  if (rho <= QB_SERIES_MAX) {
    qb_bessel_series(z, false, J0, J1, &d0, &d1);
  } else if (rho < QB_ASYM_MIN) {
    qb_bessel_miller(z, false, J0, J1, &d0, &d1);
  } else {
    cd h10, h11, h20, h21;
    qb_hankel_asym(z, true, &h10, &h11, &h20, &h21);
    *J0 = 0.5 * (h10 + h20);
    *J1 = 0.5 * (h11 + h21);
  }
}

// H0^(1)(z), H1^(1)(z) for z != 0, Re z > 0.
QB_HD void qb_hankel1_01(cd z, cd* H0, cd* H1) {
  double rho = cabsd(z);
  if (rho < QB_ASYM_MIN) {
    cd j0, j1, y0, y1;
    if (rho <= QB_SERIES_MAX)
      qb_bessel_series(z, true, &j0, &j1, &y0, &y1);
    else
      qb_bessel_miller(z, true, &j0, &j1, &y0, &y1);
    *H0 = j0 + mk(-y0.im, y0.re);  // J + iY
    *H1 = j1 + mk(-y1.im, y1.re);
  } else {
    cd h20, h21;
    qb_hankel_asym(z, false, H0, H1, &h20, &h21);
  }
}
