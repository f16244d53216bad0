// Minimal complex arithmetic shared by host and device code.
//
// The layout (re, im) of cd/cf matches Julia's ComplexF64/ComplexF32 and C's
// double _Complex / float _Complex, so buffers cross the C ABI unchanged.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define QB_HD __host__ __device__ __forceinline__

struct cd {
  double re, im;
};
struct cf {
  float re, im;
};

QB_HD cd mk(double r, double i) { return cd{r, i}; }
QB_HD cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
QB_HD cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
QB_HD cd operator-(cd a) { return cd{-a.re, -a.im}; }
QB_HD cd operator*(cd a, cd b) {
  return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
QB_HD cd operator*(double s, cd a) { return cd{s * a.re, s * a.im}; }
QB_HD cd operator*(cd a, double s) { return cd{s * a.re, s * a.im}; }
QB_HD cd cconj(cd a) { return cd{a.re, -a.im}; }
QB_HD double cabs2(cd a) { return a.re * a.re + a.im * a.im; }
QB_HD double cabsd(cd a) { return hypot(a.re, a.im); }
// Smith's algorithm for a / b (avoids overflow in |b|^2).
QB_HD cd cdiv(cd a, cd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return cd{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    double r = b.re / b.im, d = b.im + b.re * r;
    return cd{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
QB_HD cd crecip(cd b) { return cdiv(cd{1.0, 0.0}, b); }
QB_HD cd clog(cd a) { return cd{log(hypot(a.re, a.im)), atan2(a.im, a.re)}; }
// Principal square root (branch cut on the negative real axis).
QB_HD cd csqrt(cd a) {
  double m = hypot(a.re, a.im);
  if (m == 0.0) return cd{0.0, 0.0};
  if (a.re >= 0.0) {
    double t = sqrt(0.5 * (m + a.re));
    return cd{t, a.im / (2.0 * t)};
  } else {
    double t = sqrt(0.5 * (m - a.re));
    return cd{fabs(a.im) / (2.0 * t), copysign(t, a.im)};
  }
}

QB_HD cf operator+(cf a, cf b) { return cf{a.re + b.re, a.im + b.im}; }
QB_HD cf operator-(cf a, cf b) { return cf{a.re - b.re, a.im - b.im}; }
QB_HD cf operator*(cf a, cf b) {
  return cf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
QB_HD cf to_cf(cd a) { return cf{(float)a.re, (float)a.im}; }
QB_HD cd to_cd(cf a) { return cd{(double)a.re, (double)a.im}; }
