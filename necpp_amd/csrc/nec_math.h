// nec_math.h -- constants and complex arithmetic shared by host and device.
//
// The constants are nec2++'s truncated literals (src/math_util.h:72-107,
// src/electromag.cpp:22-47); matrix entries depend on them at the 1e-11
// level.  Complex operations are written out so that host and device evaluate
// them in the same order as g++ does for std::complex<double> without FMA:
//   a*b = (ar*br - ai*bi, ar*bi + ai*br), complex*real / complex/real per
//   component, complex/complex = Smith's method (libgcc __divdc3).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#define NEC_HD __host__ __device__ __forceinline__

namespace necgpu {

struct zc {
  double re, im;
};

NEC_HD zc zmk(double r, double i) {
  zc z;
  z.re = r;
  z.im = i;
  return z;
}
NEC_HD zc zadd(zc a, zc b) { return zmk(a.re + b.re, a.im + b.im); }
NEC_HD zc zsub(zc a, zc b) { return zmk(a.re - b.re, a.im - b.im); }
NEC_HD zc zneg(zc a) { return zmk(-a.re, -a.im); }
NEC_HD zc zmul(zc a, zc b) { return zmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
NEC_HD zc zmulr(zc a, double s) { return zmk(a.re * s, a.im * s); }
NEC_HD zc zdivr(zc a, double s) { return zmk(a.re / s, a.im / s); }
NEC_HD zc zconj(zc a) { return zmk(a.re, -a.im); }
NEC_HD zc zdiv(zc a, zc b) {
  zc r;
  if (fabs(b.re) < fabs(b.im)) {
    double ratio = b.re / b.im;
    double denom = (b.re * ratio) + b.im;
    r.re = ((a.re * ratio) + a.im) / denom;
    r.im = ((a.im * ratio) - a.re) / denom;
  } else {
    double ratio = b.im / b.re;
    double denom = (b.im * ratio) + b.re;
    r.re = ((a.im * ratio) + a.re) / denom;
    r.im = (a.im - (a.re * ratio)) / denom;
  }
  return r;
}

// pi() and friends with nec2++'s truncated value.
constexpr double kPi = 3.1415926536;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kFourPi = 4.0 * kPi;
constexpr double kSmin = 1.e-3;        // src/common.h:57
constexpr int kPchcon = 100000;        // src/c_geometry.h:26
constexpr double kEuler = .577215664;  // src/c_geometry.cpp:2810
constexpr double kNearConst = 4.771341189;  // src/nec_context.cpp:3055

// em::constants (src/electromag.cpp:22-47).
// Principal complex square root (the branch libm's csqrt takes: cut on the
// negative real axis, Im of the result with the sign of Im z).  For Im z != 0
// it uses 2 Re(w) Im(w) = Im(z) for the smaller part, so neither part loses
// digits to the cancellation in |z| -/+ Re z.  No rescaling: arguments here
// are of order one (ground impedance ratios).
NEC_HD zc zsqrt(zc z) {
  if (z.im == 0.0) {
    if (z.re < 0.0) return zmk(0.0, copysign(sqrt(-z.re), z.im));
    return zmk(fabs(sqrt(z.re)), copysign(0.0, z.im));
  }
  const double d = hypot(z.re, z.im);
  double r, s;
  if (z.re > 0.0) {
    r = sqrt(0.5 * (d + z.re));
    s = 0.5 * (z.im / r);
  } else {
    s = sqrt(0.5 * (d - z.re));
    r = fabs(0.5 * (z.im / s));
  }
  return zmk(r, copysign(s, z.im));
}

NEC_HD double em_permittivity() { return 8.854e-12; }
NEC_HD double em_permeability() { return kFourPi * 1.0e-7; }
NEC_HD double em_speed_of_light() { return 1.0 / sqrt(em_permeability() * em_permittivity()); }
NEC_HD double em_impedance() { return sqrt(em_permeability() / em_permittivity()); }

NEC_HD double deg_to_rad(double d) { return d * ((2 * kPi) / 360.0); }
NEC_HD double rad_to_deg(double r) { return r * (360.0 / (2 * kPi)); }
// sin(x) / x.  For |x| <= 0.5 its Taylor series to x^16 (truncation below
// 1e-19 relative), which replaces the library sine in the far-field shape
// factors of short segments; the reference's own small-argument branch
// (|x| < 1e-7: 1 - x^2 / 6) is the series' first two terms.
NEC_HD double sinc(double x) {
  if (fabs(x) > 0.5) return sin(x) / x;
  const double x2 = x * x;
  double p = 1.0 / 355687428096000.0;        // 1 / 17!
  p = p * x2 - 1.0 / 1307674368000.0;        // 1 / 15!
  p = p * x2 + 1.0 / 6227020800.0;           // 1 / 13!
  p = p * x2 - 1.0 / 39916800.0;             // 1 / 11!
  p = p * x2 + 1.0 / 362880.0;               // 1 / 9!
  p = p * x2 - 1.0 / 5040.0;                 // 1 / 7!
  p = p * x2 + 1.0 / 120.0;                  // 1 / 5!
  p = p * x2 - 1.0 / 6.0;                    // 1 / 3!
  return p * x2 + 1.0;
}

NEC_HD double db10(double x) { return (x < 1.0e-20) ? -999.99 : 10.0 * log10(x); }
NEC_HD double from_db10(double x) { return (x < -99.9) ? 0.0 : pow(10.0, x / 10.0); }

}  // namespace necgpu
