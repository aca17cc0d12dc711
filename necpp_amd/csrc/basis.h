// basis.h -- current expansion (basis) functions, host and device.
//
// Restates c_geometry::sbf (src/c_geometry.cpp:3102-3289), trio
// (:2982-3087) and tbf (:2763-2977) over plain arrays so the same code can
// build the basis tables on the host and on the GPU (batched sweeps).  The
// arithmetic order is the reference's, so coefficients are bit-identical
// when the same libm is used (host) and agree to an ulp on the device.
//
// Basis function i (1-based) is the sum of a constant, a sine and a cosine
// term on segment i and on every segment joined to either end of i; the
// coefficient triple (A, B, C) of basis i on segment j is what sbf(i, j)
// returns.
#pragma once

#include "dual.h"
#include "nec_math.h"

namespace necgpu {

template <class T>
struct SegViewT {
  const int* icon1;
  const int* icon2;
  const T* s;  // lengths in wavelengths
  const T* b;  // radii in wavelengths
};
using SegView = SegViewT<double>;

namespace basis_detail {
template <class T>
NEC_HD T omc_series_or(const T& d, const T& exact) {
  if (val(d) <= 0.015) {
    T omc = 4. * d * d;
    return ((1.3888889e-3 * omc - 4.1666666667e-2) * omc + .5) * omc;
  }
  return exact;
}
template <class T>
NEC_HD T log_weight(const T& b) {
  return 1. / (log(1. / (kPi * b)) - kEuler);
}
template <class T>
NEC_HD T end_cap(const T& b) {
  T q = kPi * b;
  T x = q * q;
  return q * (1. - .5 * x) / (1. - x);
}
}  // namespace basis_detail

// sbf(i, is): coefficients of basis i on segment is (1-based).  Returns 0,
// or 1 on a connection error.
template <class T>
NEC_HD int sbf(const SegViewT<T>& g, int i, int is, T* aa, T* bb, T* cc) {
  using namespace basis_detail;
  T pp = T(0.), pm = T(0.);
  double sig = -1.;
  int count = 0, njun1 = 0, june = 0;
  *aa = T(0.);
  *bb = T(0.);
  *cc = T(0.);
  const int ix = i - 1;
  int jcox = g.icon1[ix];
  if (jcox > kPchcon) jcox = i;
  int jend = -1, iend = -1;
  do {
    if (jcox != 0) {
      if (jcox < 0) jcox = -jcox;
      else {
        sig = -sig;
        jend = -jend;
      }
      const int k = jcox - 1;
      ++count;
      const T d = kPi * g.s[k];
      const T sdh = sin(d), cdh = cos(d);
      const T sd = 2. * sdh * cdh;
      const T omc = omc_series_or(d, 1. - cdh * cdh + sdh * sdh);
      const T aj = log_weight(g.b[k]);
      pp = pp - omc / sd * aj;
      if (jcox == is) {
        *aa = aj / sd * sig;
        *bb = aj / (2. * cdh);
        *cc = -aj / (2. * sdh) * sig;
        june = iend;
      }
      if (jcox != i) {
        jcox = (jend != 1) ? g.icon1[k] : g.icon2[k];
        if ((jcox < 0 ? -jcox : jcox) != i) {
          if (jcox == 0) return 1;
          continue;
        }
      } else if (jcox == is) {
        *bb = -*bb;
      }
      if (iend == 1) break;
    }
    pm = -pp;
    pp = T(0.);
    njun1 = count;
    jcox = g.icon2[ix];
    if (jcox > kPchcon) jcox = i;
    jend = 1;
    iend = 1;
    sig = -1.;
  } while (jcox != 0);
  const int njun2 = count - njun1;
  const T d = kPi * g.s[ix];
  const T sdh = sin(d), cdh = cos(d);
  const T sd = 2. * sdh * cdh;
  const T cd = cdh * cdh - sdh * sdh;
  const T omc = omc_series_or(d, 1. - cd);
  const T ap = log_weight(g.b[ix]);
  const T aj = ap;
  if (njun1 == 0) {
    const T xxi = end_cap(g.b[ix]);
    if (njun2 == 0) {
      *aa = T(-1.);
      *cc = 1. / (cdh - xxi * sdh);
      return 0;
    }
    const T qp = -(omc + xxi * sd) / (sd * (ap + xxi * pp) + cd * (xxi * ap - pp));
    if (june == 1) {
      *aa = -*aa * qp;
      *bb = *bb * qp;
      *cc = -*cc * qp;
      if (i != is) return 0;
    }
    *aa = *aa - 1.;
    const T den = cd - xxi * sd;
    *bb = *bb + (sdh + ap * qp * (cdh - xxi * sdh)) / den;
    *cc = *cc + (cdh + ap * qp * (sdh + xxi * cdh)) / den;
    return 0;
  }
  if (njun2 == 0) {
    const T xxi = end_cap(g.b[ix]);
    const T qm = (omc + xxi * sd) / (sd * (aj - xxi * pm) + cd * (pm + xxi * aj));
    if (june == -1) {
      *aa = *aa * qm;
      *bb = *bb * qm;
      *cc = *cc * qm;
      if (i != is) return 0;
    }
    *aa = *aa - 1.;
    const T den = cd - xxi * sd;
    *bb = *bb + (aj * qm * (cdh - xxi * sdh) - sdh) / den;
    *cc = *cc + (cdh - aj * qm * (sdh + xxi * cdh)) / den;
    return 0;
  }
  T qp = sd * (pm * pp + aj * ap) + cd * (pm * ap - pp * aj);
  const T qm = (ap * omc - pp * sd) / qp;
  qp = -(aj * omc + pm * sd) / qp;
  if (june != 0) {
    if (june < 0) {
      *aa = *aa * qm;
      *bb = *bb * qm;
      *cc = *cc * qm;
    } else {
      *aa = -*aa * qp;
      *bb = *bb * qp;
      *cc = -*cc * qp;
    }
    if (i != is) return 0;
  }
  *aa = *aa - 1.;
  *bb = *bb + (aj * qm + ap * qp) * sdh / sd;
  *cc = *cc + (aj * qm - ap * qp) * cdh / sd;
  return 0;
}

// trio(j): the basis functions with support on segment j (1-based), in the
// reference's order (segments around end 1, then around end 2, then j
// itself).  Returns the count, or -1 on a connection error / overflow.
template <class T>
NEC_HD int trio(const SegViewT<T>& g, int j, int* jco, T* ax, T* bx, T* cx, int cap) {
  int count = 0, jend = 0, iend = 0;
  const int jx = j - 1;
  int jcox = g.icon1[jx];
  if (jcox <= kPchcon) {
    jend = -1;
    iend = -1;
  }
  if (jcox == 0 || jcox > kPchcon) {
    jcox = g.icon2[jx];
    if (jcox <= kPchcon) {
      jend = 1;
      iend = 1;
    }
    if (jcox == 0 || jcox > kPchcon) {
      if (cap < 1 || sbf(g, j, j, &ax[0], &bx[0], &cx[0])) return -1;
      jco[0] = j;
      return 1;
    }
  }
  do {
    if (jcox < 0) jcox = -jcox;
    else jend = -jend;
    const int k = jcox - 1;
    if (jcox != j) {
      if (count >= cap || sbf(g, jcox, j, &ax[count], &bx[count], &cx[count])) return -1;
      jco[count++] = jcox;
      jcox = (jend != 1) ? g.icon1[k] : g.icon2[k];
      if (jcox == 0) return -1;
      continue;
    }
    if (iend == 1) break;
    jcox = g.icon2[jx];
    if (jcox > kPchcon) break;
    jend = 1;
    iend = 1;
  } while (jcox != 0);
  if (count >= cap || sbf(g, j, j, &ax[count], &bx[count], &cx[count])) return -1;
  jco[count++] = j;
  return count;
}

// tbf(i, icap): every coefficient of basis i (1-based): entries for the
// segments around end 1, then end 2, then segment i itself.  Returns the
// count, or -1 on a connection error / overflow.
NEC_HD int tbf(const SegView& g, int i, int icap, int* jco, double* ax, double* bx, double* cx, int cap) {
  using namespace basis_detail;
  int count = 0, njun1 = 0;
  double pp = 0.0, pm = 0.0, sig = -1.0;
  const int ix = i - 1;
  int jcox = g.icon1[ix];
  if (jcox > kPchcon) jcox = i;
  int jend = -1, iend = -1;
  do {
    if (jcox != 0) {
      if (jcox < 0) jcox = -jcox;
      else {
        sig = -sig;
        jend = -jend;
      }
      const int k = jcox - 1;
      if (count >= cap) return -1;
      const int e = count++;
      jco[e] = jcox;
      const double d = kPi * g.s[k];
      const double sdh = sin(d), cdh = cos(d);
      const double sd = 2.0 * sdh * cdh;
      const double omc = omc_series_or(d, 1. - cdh * cdh + sdh * sdh);
      const double aj = log_weight(g.b[k]);
      pp = pp - omc / sd * aj;
      ax[e] = aj / sd * sig;
      bx[e] = aj / (2. * cdh);
      cx[e] = -aj / (2. * sdh) * sig;
      if (jcox != i) {
        jcox = (jend == 1) ? g.icon2[k] : g.icon1[k];
        if ((jcox < 0 ? -jcox : jcox) != i) {
          if (jcox != 0) continue;
          return -1;
        }
      } else {
        bx[e] = -bx[e];
      }
      if (iend == 1) break;
    }
    pm = -pp;
    pp = 0.;
    njun1 = count;
    jcox = g.icon2[ix];
    if (jcox > kPchcon) jcox = i;
    jend = 1;
    iend = 1;
    sig = -1.;
  } while (jcox != 0);
  const int njun2 = count - njun1;
  const int self = count;
  if (self >= cap) return -1;
  jco[self] = i;
  const double d = kPi * g.s[ix];
  const double sdh = sin(d), cdh = cos(d);
  const double sd = 2.0 * sdh * cdh;
  const double cd = cdh * cdh - sdh * sdh;
  const double omc = omc_series_or(d, 1.0 - cd);
  const double ap = log_weight(g.b[ix]);
  const double aj = ap;
  const double xxi = (icap == 0) ? 0. : end_cap(g.b[ix]);
  if (njun1 == 0) {
    if (njun2 == 0) {
      bx[self] = 0.;
      cx[self] = 1. / (cdh - xxi * sdh);
      ax[self] = -1.;
      return self + 1;
    }
    const double qp = -(omc + xxi * sd) / (sd * (ap + xxi * pp) + cd * (xxi * ap - pp));
    const double den = cd - xxi * sd;
    bx[self] = (sdh + ap * qp * (cdh - xxi * sdh)) / den;
    cx[self] = (cdh + ap * qp * (sdh + xxi * cdh)) / den;
    for (int e = 0; e < njun2; ++e) {
      ax[e] = -ax[e] * qp;
      bx[e] = bx[e] * qp;
      cx[e] = -cx[e] * qp;
    }
    ax[self] = -1.;
    return self + 1;
  }
  if (njun2 == 0) {
    const double qm = (omc + xxi * sd) / (sd * (aj - xxi * pm) + cd * (pm + xxi * aj));
    const double den = cd - xxi * sd;
    bx[self] = (aj * qm * (cdh - xxi * sdh) - sdh) / den;
    cx[self] = (cdh - aj * qm * (sdh + xxi * cdh)) / den;
    for (int e = 0; e < njun1; ++e) {
      ax[e] = ax[e] * qm;
      bx[e] = bx[e] * qm;
      cx[e] = cx[e] * qm;
    }
    ax[self] = -1.;
    return self + 1;
  }
  double qp = sd * (pm * pp + aj * ap) + cd * (pm * ap - pp * aj);
  const double qm = (ap * omc - pp * sd) / qp;
  qp = -(aj * omc + pm * sd) / qp;
  bx[self] = (aj * qm + ap * qp) * sdh / sd;
  cx[self] = (aj * qm - ap * qp) * cdh / sd;
  for (int e = 0; e < njun1; ++e) {
    ax[e] = ax[e] * qm;
    bx[e] = bx[e] * qm;
    cx[e] = cx[e] * qm;
  }
  for (int e = njun1; e < count; ++e) {
    ax[e] = -ax[e] * qp;
    bx[e] = bx[e] * qp;
    cx[e] = -cx[e] * qp;
  }
  ax[self] = -1.;
  return self + 1;
}

}  // namespace necgpu
