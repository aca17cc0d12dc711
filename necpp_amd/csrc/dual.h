// dual.h -- forward-mode derivative numbers for the adjoint path.
//
// Dn<NV> carries a value and NV partial derivatives.  The value part of
// every operation is computed exactly as the plain double code computes it
// (same operations, same order), so a kernel instantiated on Dn<NV> returns
// the plain result in .v and the derivatives of that computed result (with
// the same branch decisions) in .d.  cx<T> is the complex type the templated
// fill code uses; for T = double its operations are those of zc.
#pragma once

#include <cmath>

#include "nec_math.h"

namespace necgpu {

template <int NV>
struct Dn {
  double v;
  double d[NV];
  Dn() = default;
  // a constant: no derivative
  NEC_HD Dn(double x) : v(x) {
    for (int q = 0; q < NV; ++q) d[q] = 0.0;
  }
};

// ---- value access / construction, for double and Dn alike
NEC_HD double val(double x) { return x; }
template <int NV>
NEC_HD double val(const Dn<NV>& x) {
  return x.v;
}
// ---- arithmetic
template <int NV>
NEC_HD Dn<NV> operator+(const Dn<NV>& a, const Dn<NV>& b) {
  Dn<NV> r;
  r.v = a.v + b.v;
  for (int q = 0; q < NV; ++q) r.d[q] = a.d[q] + b.d[q];
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator-(const Dn<NV>& a, const Dn<NV>& b) {
  Dn<NV> r;
  r.v = a.v - b.v;
  for (int q = 0; q < NV; ++q) r.d[q] = a.d[q] - b.d[q];
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator-(const Dn<NV>& a) {
  Dn<NV> r;
  r.v = -a.v;
  for (int q = 0; q < NV; ++q) r.d[q] = -a.d[q];
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator*(const Dn<NV>& a, const Dn<NV>& b) {
  Dn<NV> r;
  r.v = a.v * b.v;
  for (int q = 0; q < NV; ++q) r.d[q] = a.d[q] * b.v + a.v * b.d[q];
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator/(const Dn<NV>& a, const Dn<NV>& b) {
  Dn<NV> r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
  for (int q = 0; q < NV; ++q) r.d[q] = (a.d[q] - r.v * b.d[q]) * ib;
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator+(const Dn<NV>& a, double b) {
  Dn<NV> r = a;
  r.v = a.v + b;
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator+(double a, const Dn<NV>& b) {
  Dn<NV> r = b;
  r.v = a + b.v;
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator-(const Dn<NV>& a, double b) {
  Dn<NV> r = a;
  r.v = a.v - b;
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator-(double a, const Dn<NV>& b) {
  Dn<NV> r;
  r.v = a - b.v;
  for (int q = 0; q < NV; ++q) r.d[q] = -b.d[q];
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator*(const Dn<NV>& a, double b) {
  Dn<NV> r;
  r.v = a.v * b;
  for (int q = 0; q < NV; ++q) r.d[q] = a.d[q] * b;
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator*(double a, const Dn<NV>& b) {
  Dn<NV> r;
  r.v = a * b.v;
  for (int q = 0; q < NV; ++q) r.d[q] = a * b.d[q];
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator/(const Dn<NV>& a, double b) {
  Dn<NV> r;
  r.v = a.v / b;
  for (int q = 0; q < NV; ++q) r.d[q] = a.d[q] / b;
  return r;
}
template <int NV>
NEC_HD Dn<NV> operator/(double a, const Dn<NV>& b) {
  Dn<NV> r;
  r.v = a / b.v;
  const double f = -r.v / b.v;
  for (int q = 0; q < NV; ++q) r.d[q] = f * b.d[q];
  return r;
}

// ---- functions (values through the same library calls as the double code);
// the double overloads stay visible next to the Dn ones
using ::cos;
using ::fabs;
using ::log;
using ::sin;
using ::sincos;
using ::sqrt;
template <int NV>
NEC_HD Dn<NV> sqrt(const Dn<NV>& a) {
  Dn<NV> r;
  r.v = ::sqrt(a.v);
  // at zero (a coincident point) the derivative seeds are zero: keep them so
  const double f = (r.v > 0.0) ? 0.5 / r.v : 0.0;
  for (int q = 0; q < NV; ++q) r.d[q] = f * a.d[q];
  return r;
}
template <int NV>
NEC_HD Dn<NV> log(const Dn<NV>& a) {
  Dn<NV> r;
  r.v = ::log(a.v);
  for (int q = 0; q < NV; ++q) r.d[q] = a.d[q] / a.v;
  return r;
}
template <int NV>
NEC_HD Dn<NV> sin(const Dn<NV>& a) {
  Dn<NV> r;
  r.v = ::sin(a.v);
  const double c = ::cos(a.v);
  for (int q = 0; q < NV; ++q) r.d[q] = c * a.d[q];
  return r;
}
template <int NV>
NEC_HD Dn<NV> cos(const Dn<NV>& a) {
  Dn<NV> r;
  r.v = ::cos(a.v);
  const double s = ::sin(a.v);
  for (int q = 0; q < NV; ++q) r.d[q] = -s * a.d[q];
  return r;
}
template <int NV>
NEC_HD void sincos(const Dn<NV>& a, Dn<NV>* s, Dn<NV>* c) {
  double sv, cv;
  ::sincos(a.v, &sv, &cv);
  s->v = sv;
  c->v = cv;
  for (int q = 0; q < NV; ++q) {
    s->d[q] = cv * a.d[q];
    c->d[q] = -sv * a.d[q];
  }
}
template <int NV>
NEC_HD Dn<NV> fabs(const Dn<NV>& a) {
  return (a.v < 0.0) ? -a : a;
}

// ---- complex numbers over T
template <class T>
struct cx {
  T re, im;
};
template <class T>
NEC_HD cx<T> cmk(const T& r, const T& i) {
  cx<T> z;
  z.re = r;
  z.im = i;
  return z;
}
template <class T>
NEC_HD cx<T> cadd(const cx<T>& a, const cx<T>& b) {
  return cmk<T>(a.re + b.re, a.im + b.im);
}
template <class T>
NEC_HD cx<T> csub(const cx<T>& a, const cx<T>& b) {
  return cmk<T>(a.re - b.re, a.im - b.im);
}
template <class T>
NEC_HD cx<T> cneg(const cx<T>& a) {
  return cmk<T>(-a.re, -a.im);
}
template <class T>
NEC_HD cx<T> cmul(const cx<T>& a, const cx<T>& b) {
  return cmk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class T>
NEC_HD cx<T> cmulr(const cx<T>& a, const T& s) {
  return cmk<T>(a.re * s, a.im * s);
}
template <class T>
NEC_HD cx<T> cdivr(const cx<T>& a, const T& s) {
  return cmk<T>(a.re / s, a.im / s);
}
// multiplication by a complex constant (no derivative)
template <class T>
NEC_HD cx<T> cmulc(const zc& a, const cx<T>& b) {
  return cmk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}

}  // namespace necgpu
