// fill.hip -- FP64 impedance-matrix fill for thin wires in free space.
//
// The matrix entry Z[i][k] (observation segment i, basis function k) is
//   sum over source segments j in the support of basis k, ascending j, of
//   etk(i,j)*A(k,j) + ets(i,j)*B(k,j) + etc(i,j)*C(k,j)
// where (etk, ets, etc) are the fields at segment i of unit constant / sine /
// cosine currents on segment j projected on i's direction
// (nec_context::cmset / cmww, src/nec_context.cpp:2077-2203, 2514-2654), and
// (A, B, C) the coefficients of basis k on segment j (c_geometry::trio).
//
// Two kernels:
//   fill_fields    one thread per (i, j) pair: efld_compute -> eksc / ekscx
//                  (adaptive Romberg intx over gf) or the lumped-element
//                  far branch, projected; writes E[i][j] = 3 complex.
//   fill_assemble  one thread per (i, k): gathers E[i][j] over the support
//                  list of column k in ascending j, which is the order the
//                  reference's scatter adds them in, starting from zero.
// Both are compiled with -ffp-contract=off and follow the reference's
// operation order, so entries agree with nec2++ to the ulp-level differences
// of the device sin/cos/log.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "dual.h"

#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

// The near-field routines below are templates over the scalar type T:
// double for the fill, Dn<NV> (forward-mode derivatives) for the adjoint.
// The double instantiation performs exactly the operations of the reference
// routine it cites, in its order.
template <class T>
struct GfStateT {
  int ija;
  T zpk, rkb2;
};
using GfState = GfStateT<double>;

// gf, src/nec_context.cpp:3553-3576
template <class T>
__device__ __forceinline__ void gf(const GfStateT<T>& st, const T& zk, T* co, T* si) {
  const T zdk = zk - st.zpk;
  const T rk = sqrt(st.rkb2 + zdk * zdk);
  T s, c;
  sincos(rk, &s, &c);
  *si = s / rk;
  if (st.ija != 0) {
    *co = c / rk;
    return;
  }
  if (val(rk) >= .2) {
    *co = (c - 1.0) / rk;
    return;
  }
  const T rks = rk * rk;
  *co = ((-1.38888889e-3 * rks + 4.16666667e-2) * rks - 0.5) * rk;
}

// test, src/c_evlcom.cpp:845-868 (dmin = 0)
template <class T>
__device__ __forceinline__ void conv_test(const T& f1r, const T& f2r, T* tr, const T& f1i, const T& f2i, T* ti) {
  T den = fabs(f2r);
  const T t = fabs(f2i);
  if (val(den) < val(t)) den = t;
  if (val(den) < 1.0e-37) {
    *tr = T(0.0);
    *ti = T(0.0);
    return;
  }
  *tr = fabs((f1r - f2r) / den);
  *ti = fabs((f1i - f2i) / den);
}

// intx, src/nec_context.cpp:4135-4297: variable-step Romberg integration of
// the gf integrand; for the self term (ij == 0) half the segment plus the
// analytic log term.  The step decisions use values only, so a Dn
// instantiation differentiates the quadrature the double one computes.
template <class T>
__device__ void intx(const GfStateT<T>& st, const T& el1, const T& el2, const T& b, int ij, T* sgr, T* sgi) {
  const int nx = 1, nma = 65536, nts = 4;
  const double rx = 1.0e-4;
  T z = el1;
  T ze = el2;
  if (ij == 0) ze = T(0.);
  const T s = ze - z;
  const double fnm = nma;
  const T ep = s / (10. * fnm);
  const T zend = ze - ep;
  T accr = T(0.), acci = T(0.);
  int ns = nx, nt = 0;
  bool fresh = true;
  T dz = T(0.), dzot = T(0.);
  T g1r, g1i, g2r, g2i, g3r = T(0), g3i = T(0), g4r, g4i, g5r = T(0), g5i = T(0);
  gf(st, z, &g1r, &g1i);
  // The step count is bounded by 65536 subdivisions of a finite interval;
  // the guard only stops a non-finite interval from spinning forever.
  long guard = 0;
  if (!(fabs(val(s)) < 1.0e300)) {
    *sgr = s;
    *sgi = s;
    return;
  }
  for (; guard < (1L << 22); ++guard) {
    if (fresh) {
      const double fns = ns;
      dz = s / fns;
      T zp = z + dz;
      if (val(zp) > val(ze)) {
        dz = ze - z;
        if (fabs(val(dz)) <= val(ep)) break;
      }
      dzot = dz * .5;
      zp = z + dzot;
      gf(st, zp, &g3r, &g3i);
      zp = z + dz;
      gf(st, zp, &g5r, &g5i);
    }
    const T t00r = (g1r + g5r) * dzot;
    const T t00i = (g1i + g5i) * dzot;
    const T t01r = (t00r + dz * g3r) * 0.5;
    const T t01i = (t00i + dz * g3i) * 0.5;
    const T t10r = (4.0 * t01r - t00r) / 3.0;
    const T t10i = (4.0 * t01i - t00i) / 3.0;
    T te1r, te1i;
    conv_test(t01r, t10r, &te1r, t01i, t10i, &te1i);
    if ((val(te1i) <= rx) && (val(te1r) <= rx)) {
      accr = accr + t10r;
      acci = acci + t10i;
      nt += 2;
      z = z + dz;
      if (val(z) >= val(zend)) break;
      g1r = g5r;
      g1i = g5i;
      if (nt >= nts && ns > nx) {
        ns = ns / 2;
        nt = 1;
      }
      fresh = true;
      continue;
    }
    T zp = z + dz * 0.25;
    gf(st, zp, &g2r, &g2i);
    zp = z + dz * 0.75;
    gf(st, zp, &g4r, &g4i);
    const T t02r = (t01r + dzot * (g2r + g4r)) * 0.5;
    const T t02i = (t01i + dzot * (g2i + g4i)) * 0.5;
    const T t11r = (4.0 * t02r - t01r) / 3.0;
    const T t11i = (4.0 * t02i - t01i) / 3.0;
    const T t20r = (16.0 * t11r - t10r) / 15.0;
    const T t20i = (16.0 * t11i - t10i) / 15.0;
    T te2r, te2i;
    conv_test(t11r, t20r, &te2r, t11i, t20i, &te2i);
    if ((val(te2i) > rx) || (val(te2r) > rx)) {
      nt = 0;
      if (ns < nma) {
        ns = ns * 2;
        const double fns = ns;
        dz = s / fns;
        dzot = dz * 0.5;
        g5r = g3r;
        g5i = g3i;
        g3r = g2r;
        g3i = g2i;
        fresh = false;
        continue;
      }
      // step size limited: the reference accepts the 5-point value
    }
    accr = accr + t20r;
    acci = acci + t20i;
    nt++;
    z = z + dz;
    if (val(z) >= val(zend)) break;
    g1r = g5r;
    g1i = g5i;
    if (nt >= nts && ns > nx) {
      ns = ns / 2;
      nt = 1;
    }
    fresh = true;
  }
  if (ij == 0) {
    accr = 2. * (accr + log((sqrt(b * b + s * s) + s) / b));
    acci = 2. * acci;
  }
  *sgr = accr;
  *sgi = acci;
}

// gx, src/nec_context.cpp:3605-3616
template <class T>
__device__ __forceinline__ void gx_t(const T& zz, const T& rh, double xk, cx<T>* gz, cx<T>* gzp) {
  const T r2 = zz * zz + rh * rh;
  const T r = sqrt(r2);
  const T rkz = xk * r;
  T s, c;
  sincos(rkz, &s, &c);
  const cx<T> temp = cdivr(cmk<T>(c, -s), r);
  *gz = temp;
  *gzp = cdivr(cmul(cneg(cmk<T>(T(1.0), rkz)), temp), r2);
}
__device__ __forceinline__ void gx(double zz, double rh, double xk, zc* gz, zc* gzp) {
  cx<double> a, b;
  gx_t<double>(zz, rh, xk, &a, &b);
  *gz = zmk(a.re, a.im);
  *gzp = zmk(b.re, b.im);
}

// gxx, src/nec_context.cpp:3623-3673
__device__ void gxx(double zz, double rh, double a, double a2, double xk, int ira, zc* g1, zc* g1p, zc* g2, zc* g2p,
                    zc* g3, zc* gzp) {
  const double r2 = zz * zz + rh * rh;
  const double r = sqrt(r2);
  const double r4 = r2 * r2;
  const double rk = xk * r;
  const double rk2 = rk * rk;
  const double rh2 = rh * rh;
  const double t1 = .25 * a2 * rh2 / r4;
  double t2 = .5 * a2 / r2;
  const zc c1 = zmk(1.0, rk);
  zc c2 = zmulr(c1, 3.);
  c2.re = c2.re - rk2;
  const zc c3 = zsub(zmulr(zmk(6.0, rk), rk2), zmulr(c1, 15.));
  double s, c;
  sincos(rk, &s, &c);
  zc gz = zdivr(zmk(c, -s), r);
  const zc tc = zmulr(c2, t1);
  *g2 = zmul(gz, zmk(1. + tc.re, tc.im));
  *g1 = zsub(*g2, zmul(zmulr(c1, t2), gz));
  gz = zdivr(gz, r2);
  *g2p = zmul(gz, zsub(zmulr(c3, t1), c1));
  *gzp = zmul(zmulr(c2, t2), gz);
  *g3 = zadd(*g2p, *gzp);
  *g1p = zmulr(*g3, zz);
  if (ira != 1) {
    *g3 = zmulr(zadd(*g3, *gzp), rh);
    *gzp = zmul(zmulr(c1, -zz), gz);
    if (rh <= 1.0e-10) {
      *g2 = zmk(0., 0.);
      *g2p = zmk(0., 0.);
      return;
    }
    *g2 = zdivr(*g2, rh);
    *g2p = zdivr(zmulr(*g2p, zz), rh);
    return;
  }
  t2 = .5 * a;
  *g2 = zmul(zmulr(c1, -t2), gz);
  *g2p = zdivr(zmul(zmulr(gz, t2), c2), r2);
  *g3 = zsub(zmulr(*g2p, rh2), zmul(zmulr(gz, a), c1));
  *g2p = zmulr(*g2p, zz);
  *gzp = zmul(zmulr(c1, -zz), gz);
}

struct NearOut {
  zc ezs, ers, ezc, erc, ezk, erk;
};
template <class T>
struct NearOutT {
  cx<T> ezs, ers, ezc, erc, ezk, erk;
};

// eksc, src/nec_context.cpp:3051-3100
template <class T>
__device__ void eksc(const T& s, const T& z, const T& rh, double xk, int ij, NearOutT<T>* o) {
  const zc k1 = zmk(0.0, kNearConst), mk1 = zneg(k1);
  GfStateT<T> st;
  st.ija = ij;
  st.zpk = xk * z;
  const T rhk = xk * rh;
  st.rkb2 = rhk * rhk;
  const T sh = .5 * s;
  const T shk = xk * sh;
  T ss, cs;
  sincos(shk, &ss, &cs);
  const T z2a = sh - z;
  const T z1a = -(sh + z);
  cx<T> gz1, gz2, gp1, gp2;
  gx_t(z1a, rh, xk, &gz1, &gp1);
  gx_t(z2a, rh, xk, &gz2, &gp2);
  cx<T> gzp1 = cmulr(gp1, z1a);
  cx<T> gzp2 = cmulr(gp2, z2a);
  const T txk = T(xk);
  o->ezs = cmulc(k1, csub(cmulr(cmulr(csub(gz2, gz1), cs), txk), cmulr(cadd(gzp2, gzp1), ss)));
  o->ezc = cmulc(mk1, cadd(cmulr(cmulr(cadd(gz2, gz1), ss), txk), cmulr(csub(gzp2, gzp1), cs)));
  o->erk = cmulr(cmulc(k1, csub(gp2, gp1)), rh);
  T cint, sint;
  intx(st, -shk, shk, rhk, ij, &cint, &sint);
  o->ezk = cmulc(mk1, cadd(csub(gzp2, gzp1), cmulr(cmk<T>(cint, -sint), T(xk * xk))));
  if (val(rh) >= 1.0e-10) {
    gzp1 = cmulr(gzp1, z1a);
    gzp2 = cmulr(gzp2, z2a);
    const cx<T> t1 = cmulr(cadd(cadd(cadd(gzp2, gzp1), gz2), gz1), ss);
    const cx<T> t2 = cmulr(cmulr(csub(cmulr(gz2, z2a), cmulr(gz1, z1a)), cs), txk);
    o->ers = cdivr(cmulc(mk1, csub(t1, t2)), rh);
    const cx<T> u1 = cmulr(csub(cadd(csub(gzp2, gzp1), gz2), gz1), cs);
    const cx<T> u2 = cmulr(cmulr(cadd(cmulr(gz2, z2a), cmulr(gz1, z1a)), ss), txk);
    o->erc = cdivr(cmulc(mk1, cadd(u1, u2)), rh);
    return;
  }
  o->ers = cmk<T>(T(0.), T(0.));
  o->erc = cmk<T>(T(0.), T(0.));
}

// ekscx, src/nec_context.cpp:3108-3188
__device__ void ekscx(double bx, double s, double z, double rhx, double xk, int ij, int inx1, int inx2, NearOut* o) {
  const zc k1 = zmk(0.0, kNearConst), mk1 = zneg(k1);
  int ira;
  double b, rh;
  if (rhx >= bx) {
    rh = rhx;
    b = bx;
    ira = 0;
  } else {
    rh = bx;
    b = rhx;
    ira = 1;
  }
  const double sh = .5 * s;
  GfState st;
  st.ija = ij;
  st.zpk = xk * z;
  const double rhk = xk * rh;
  st.rkb2 = rhk * rhk;
  const double shk = xk * sh;
  double ss, cs;
  sincos(shk, &ss, &cs);
  const double z2a = sh - z;
  const double z1a = -(sh + z);
  const double a2 = b * b;
  zc gz1, gz2, gzp1, gzp2, gr1, gr2, grp1, grp2, grk1, grk2, gzz1, gzz2;
  if (inx1 != 2) {
    gxx(z1a, rh, b, a2, xk, ira, &gz1, &gzp1, &gr1, &grp1, &grk1, &gzz1);
  } else {
    gx(z1a, rhx, xk, &gz1, &grk1);
    gzp1 = zmulr(grk1, z1a);
    gr1 = zdivr(gz1, rhx);
    grp1 = zdivr(gzp1, rhx);
    grk1 = zmulr(grk1, rhx);
    gzz1 = zmk(0., 0.);
  }
  if (inx2 != 2) {
    gxx(z2a, rh, b, a2, xk, ira, &gz2, &gzp2, &gr2, &grp2, &grk2, &gzz2);
  } else {
    gx(z2a, rhx, xk, &gz2, &grk2);
    gzp2 = zmulr(grk2, z2a);
    gr2 = zdivr(gz2, rhx);
    grp2 = zdivr(gzp2, rhx);
    grk2 = zmulr(grk2, rhx);
    gzz2 = zmk(0., 0.);
  }
  o->ezs = zmul(k1, zsub(zmulr(zmulr(zsub(gz2, gz1), cs), xk), zmulr(zadd(gzp2, gzp1), ss)));
  o->ezc = zmul(mk1, zadd(zmulr(zmulr(zadd(gz2, gz1), ss), xk), zmulr(zsub(gzp2, gzp1), cs)));
  {
    const zc t1 = zmulr(zadd(zadd(zadd(zmulr(grp2, z2a), zmulr(grp1, z1a)), gr2), gr1), ss);
    const zc t2 = zmulr(zmulr(zsub(zmulr(gr2, z2a), zmulr(gr1, z1a)), cs), xk);
    o->ers = zmul(mk1, zsub(t1, t2));
    const zc u1 = zmulr(zsub(zadd(zsub(zmulr(grp2, z2a), zmulr(grp1, z1a)), gr2), gr1), cs);
    const zc u2 = zmulr(zmulr(zadd(zmulr(gr2, z2a), zmulr(gr1, z1a)), ss), xk);
    o->erc = zmul(mk1, zadd(u1, u2));
  }
  o->erk = zmul(k1, zsub(grk2, grk1));
  double cint, sint;
  intx(st, -shk, shk, rhk, ij, &cint, &sint);
  const double bk = b * xk;
  const double bk2 = bk * bk * .25;
  const zc term = zmulr(zmk(cint, -sint), (xk * xk) * (1. - bk2));
  o->ezk = zmul(mk1, zsub(zadd(zsub(gzp2, gzp1), term), zmulr(zsub(gzz2, gzz1), bk2)));
}

// Fields of the constant (k), sine (s) and cosine (c) currents of source
// segment j at observation segment i, projected on i's direction
// (efld_compute, src/nec_context.cpp:2777-2837, and the projection of cmww,
// :2617-2619).  Segment data by value, so that a Dn instantiation
// differentiates with respect to the centres, the observation radius, the
// source length and the source direction (the observation direction enters
// only the final projection; evec, when given, receives the nine unprojected
// field components e0..e8 for its derivative).  The extended kernel (EK
// card) exists for double only.
template <class T>
__device__ void pair_fields_v(const T& xi_, const T& yi_, const T& zi_, const T& ai, double cabi, double sabi,
                              double salpi, const T& xj_, const T& yj_, const T& zj_, const T& sj, const T& bj,
                              const T& cabj, const T& sabj, const T& salpj, bool self, double rkh, int use_ek,
                              int ind1, int ind2, cx<T>* etk, cx<T>* ets, cx<T>* etc, cx<double>* evec = nullptr,
                              double* rout = nullptr) {
  const T xij = xi_ - xj_;
  const T yij = yi_ - yj_;
  const T zij = zi_ - zj_;
  const T zp = xij * cabj + yij * sabj + zij * salpj;
  T rhox = xij - cabj * zp;
  T rhoy = yij - sabj * zp;
  T rhoz = zij - salpj * zp;
  const T rh = sqrt(rhox * rhox + rhoy * rhoy + rhoz * rhoz + ai * ai);
  if (val(rh) <= 1.e-10) {
    rhox = T(0.0);
    rhoy = T(0.0);
    rhoz = T(0.0);
  } else {
    // one reciprocal for the three components (a rounding apart from the
    // reference's three divisions; no decision depends on them)
    const T irh = T(1.0) / rh;
    rhox = rhox * irh;
    rhoy = rhoy * irh;
    rhoz = rhoz * irh;
  }
  const T r = sqrt(zp * zp + rh * rh);
  if (rout) *rout = val(r);
  cx<T> tezk, terk, tezc, terc;
  const cx<T> zero = cmk<T>(T(0), T(0));
  cx<T> e3 = zero, e4 = zero, e5 = zero;
  if (val(r) >= rkh) {
    // lumped current element, src/nec_context.cpp:2798-2816; the divisions
    // by r, by py and by 2 py and the two of 1 / rmag are multiplications by
    // three reciprocals here (no decision depends on these values)
    const double eta = em_impedance();
    T rmag = kTwoPi * r;
    const T ir = T(1.0) / r;
    const T cth = zp * ir;
    const T px = rh * ir;
    T sr, cr;
    sincos(rmag, &sr, &cr);
    const cx<T> e0 = cmk<T>(cr, -sr);
    const T py = kTwoPi * r * r;
    const T ipy = T(1.0) / py;
    const T irm = T(1.0) / rmag;
    const cx<T> e1 = cmulr(cmul(cmulr(e0, eta * cth), cmk<T>(T(1.0), -irm)), ipy);
    const cx<T> e2 = cmulr(cmul(cmulr(e0, eta * px), cmk<T>(T(1.0), rmag - irm)), 0.5 * ipy);
    tezk = csub(cmulr(e1, cth), cmulr(e2, px));
    terk = cadd(cmulr(e1, px), cmulr(e2, cth));
    rmag = sin(kPi * sj) / kPi;
    tezc = cmulr(tezk, rmag);
    terc = cmulr(terk, rmag);
    tezk = cmulr(tezk, sj);
    terk = cmulr(terk, sj);
  } else {
    NearOutT<T> o;
    if constexpr (std::is_same<T, double>::value) {
      if (use_ek) {
        NearOut ox;
        ekscx(val(bj), val(sj), val(zp), val(rh), kTwoPi, !self, ind1, ind2, &ox);
        o.ezs = cmk(ox.ezs.re, ox.ezs.im);
        o.ers = cmk(ox.ers.re, ox.ers.im);
        o.ezc = cmk(ox.ezc.re, ox.ezc.im);
        o.erc = cmk(ox.erc.re, ox.erc.im);
        o.ezk = cmk(ox.ezk.re, ox.ezk.im);
        o.erk = cmk(ox.erk.re, ox.erk.im);
      } else {
        eksc(sj, zp, rh, kTwoPi, !self, &o);
      }
    } else {
      eksc(sj, zp, rh, kTwoPi, !self, &o);
    }
    tezk = o.ezk;
    terk = o.erk;
    tezc = o.ezc;
    terc = o.erc;
    e3 = cadd(cmulr(o.ezs, cabj), cmulr(o.ers, rhox));
    e4 = cadd(cmulr(o.ezs, sabj), cmulr(o.ers, rhoy));
    e5 = cadd(cmulr(o.ezs, salpj), cmulr(o.ers, rhoz));
  }
  const cx<T> e0 = cadd(cmulr(tezk, cabj), cmulr(terk, rhox));
  const cx<T> e1 = cadd(cmulr(tezk, sabj), cmulr(terk, rhoy));
  const cx<T> e2 = cadd(cmulr(tezk, salpj), cmulr(terk, rhoz));
  const cx<T> e6 = cadd(cmulr(tezc, cabj), cmulr(terc, rhox));
  const cx<T> e7 = cadd(cmulr(tezc, sabj), cmulr(terc, rhoy));
  const cx<T> e8 = cadd(cmulr(tezc, salpj), cmulr(terc, rhoz));
  if (evec) {
    const cx<T>* ek[9] = {&e0, &e1, &e2, &e3, &e4, &e5, &e6, &e7, &e8};
    for (int q = 0; q < 9; ++q) evec[q] = cmk<double>(val(ek[q]->re), val(ek[q]->im));
  }
  *etk = cadd(cadd(cmulr(e0, T(cabi)), cmulr(e1, T(sabi))), cmulr(e2, T(salpi)));
  *ets = cadd(cadd(cmulr(e3, T(cabi)), cmulr(e4, T(sabi))), cmulr(e5, T(salpi)));
  *etc = cadd(cadd(cmulr(e6, T(cabi)), cmulr(e7, T(sabi))), cmulr(e8, T(salpi)));
}

__device__ void pair_fields(const DevSegments& g, int i, int j, double rkh, int use_ek, int ind1, int ind2, zc* etk,
                            zc* ets, zc* etc) {
  cx<double> a, b, c;
  pair_fields_v<double>(g.x[i], g.y[i], g.z[i], g.b[i], g.cab[i], g.sab[i], g.salp[i], g.x[j], g.y[j], g.z[j],
                        g.s[j], g.b[j], g.cab[j], g.sab[j], g.salp[j], i == j, rkh, use_ek, ind1, ind2, &a, &b, &c);
  *etk = zmk(a.re, a.im);
  *ets = zmk(b.re, b.im);
  *etc = zmk(c.re, c.im);
}

// The fields at (xi, yi, zi) (observation radius ai) of segment j's unit
// constant / sine / cosine currents over a ground, efld
// src/nec_context.cpp:2845-2956, unprojected (x, y, z of each): the direct
// fields e and the image fields t (efld_compute with zij = zi + zj, the
// source direction's z cosine negated, never the self term), t weighted for
// the finite ground by the reflection coefficients at the specular point,
// then d = e - t * frati.
__device__ void point_fields_ground(const DevSegments& g, int j, double xi, double yi, double zi, double ai,
                                    bool self, double rkh, int use_ek, int ind1, int ind2, const GroundSpec& gs,
                                    zc d[9]) {
  cx<double> a, b, c, e[9], t[9];
  pair_fields_v<double>(xi, yi, zi, ai, 1.0, 0.0, 0.0, g.x[j], g.y[j], g.z[j], g.s[j], g.b[j], g.cab[j], g.sab[j],
                        g.salp[j], self, rkh, use_ek, ind1, ind2, &a, &b, &c, e);
  double r;
  // zi - (-zj) == zi + zj exactly
  pair_fields_v<double>(xi, yi, zi, ai, 1.0, 0.0, 0.0, g.x[j], g.y[j], -g.z[j], g.s[j], g.b[j], g.cab[j], g.sab[j],
                        -g.salp[j], false, rkh, use_ek, ind1, ind2, &a, &b, &c, t, &r);
  if (gs.type == 2) {
    const zc zratx = zmk(gs.zr_re, gs.zr_im);
    const double xij = xi - g.x[j], yij = yi - g.y[j], zij = zi + g.z[j];
    const double rmag = r;
    const double xymag = sqrt(xij * xij + yij * yij);
    double px, py, cth;
    zc zrsin;
    if (xymag <= 1.0e-6) {
      px = 0.;
      py = 0.;
      cth = 1.;
      zrsin = zmk(1.0, 0.0);
    } else {
      px = -yij / xymag;
      py = xij / xymag;
      cth = zij / rmag;
      // 1 - zratx^2 (1 - cth^2)
      const zc q = zmulr(zmul(zratx, zratx), 1.0 - cth * cth);
      zrsin = zsqrt(zmk(1.0 - q.re, -q.im));
    }
    const zc zz = zmul(zratx, zrsin);
    const zc refs = zdiv(zmk(cth - zz.re, -zz.im), zmk(cth + zz.re, zz.im));
    const zc zc1 = zmulr(zratx, cth);
    zc refps = zdiv(zneg(zsub(zc1, zrsin)), zadd(zc1, zrsin));
    refps = zsub(refps, refs);
    for (int q = 0; q < 9; q += 3) {
      const zc tx = zmk(t[q].re, t[q].im), ty = zmk(t[q + 1].re, t[q + 1].im), tz = zmk(t[q + 2].re, t[q + 2].im);
      zc epy = zadd(zmulr(tx, px), zmulr(ty, py));
      const zc epx = zmulr(epy, px);
      epy = zmulr(epy, py);
      const zc nx = zadd(zmul(refs, tx), zmul(refps, epx));
      const zc ny = zadd(zmul(refs, ty), zmul(refps, epy));
      const zc nz = zmul(refs, tz);
      t[q] = cmk<double>(nx.re, nx.im);
      t[q + 1] = cmk<double>(ny.re, ny.im);
      t[q + 2] = cmk<double>(nz.re, nz.im);
    }
  }
  const zc fr = zmk(gs.fr_re, gs.fr_im);
  for (int q = 0; q < 9; ++q) d[q] = zsub(zmk(e[q].re, e[q].im), zmul(zmk(t[q].re, t[q].im), fr));
}

// Pair (i, j) over a ground: the point fields at segment i's centre with its
// radius, projected on i's direction.
__device__ void pair_fields_ground(const DevSegments& g, int i, int j, double rkh, int use_ek, int ind1, int ind2,
                                   const GroundSpec& gs, zc* etk, zc* ets, zc* etc) {
  zc d[9];
  point_fields_ground(g, j, g.x[i], g.y[i], g.z[i], g.b[i], i == j, rkh, use_ek, ind1, ind2, gs, d);
  const double ci = g.cab[i], si = g.sab[i], li = g.salp[i];
  *etk = zadd(zadd(zmulr(d[0], ci), zmulr(d[1], si)), zmulr(d[2], li));
  *ets = zadd(zadd(zmulr(d[3], ci), zmulr(d[4], si)), zmulr(d[5], li));
  *etc = zadd(zadd(zmulr(d[6], ci), zmulr(d[7], si)), zmulr(d[8], li));
}

// GROUND: the image term of every pair (a separate instantiation keeps the
// free-space kernel's register use as it is).
template <bool GROUND>
__global__ void __launch_bounds__(256) fill_fields_kernel(DevSegments g, int n, int i0, int rows, double rkh,
                                                          int use_ek, const int* ekind, GroundSpec gs,
                                                          const int* __restrict__ ord, double* __restrict__ E) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (t >= n || r >= rows) return;
  const int j = ord[t];  // the source at position t of the field order
  const int i = i0 + r;
  int ind1 = 2, ind2 = 2;
  if (use_ek) {
    ind1 = ekind[2 * j];
    ind2 = ekind[2 * j + 1];
  }
  zc etk, ets, etc;
  if (GROUND) pair_fields_ground(g, i, j, rkh, use_ek, ind1, ind2, gs, &etk, &ets, &etc);
  else pair_fields(g, i, j, rkh, use_ek, ind1, ind2, &etk, &ets, &etc);
  double* o = E + ((size_t)r * n + t) * 6;
  o[0] = etk.re;
  o[1] = etk.im;
  o[2] = ets.re;
  o[3] = ets.im;
  o[4] = etc.re;
  o[5] = etc.im;
}

// Z[i][k] for RB consecutive rows i of the pass and one basis k per thread:
// the sum over k's support (ascending source segment j) of etk*a + ets*b +
// etc*c, as cmset/cmww accumulate it (src/nec_context.cpp:2635-2638), then
// a load on segment i right after its own term.  The support entries are
// read once for the RB rows, whose field loads are independent.  Thread t
// takes basis ord[t] and finds source j's fields at position cjp[q] (the
// field order), so that neighbouring threads, whose supports meet at the
// same junctions, read neighbouring records.
template <int RB>
__global__ void __launch_bounds__(256) fill_assemble_kernel(const double* __restrict__ E, int n, int i0, int rows,
                                                            const int* __restrict__ colptr,
                                                            const int* __restrict__ colj,
                                                            const int* __restrict__ cjp,
                                                            const int* __restrict__ ord,
                                                            const double* __restrict__ colabc,
                                                            const double* __restrict__ zload,
                                                            double* __restrict__ Z, long long ldz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int r0 = blockIdx.y * RB;
  if (t >= n || r0 >= rows) return;
  const int k = ord[t];
  const int nr = min(RB, rows - r0);
  double accr[RB], acci[RB];
#pragma unroll
  for (int u = 0; u < RB; ++u) accr[u] = acci[u] = 0.0;
  const int q1 = colptr[k + 1];
  for (int q = colptr[k]; q < q1; ++q) {
    const int jp = cjp[q];
    const double a = colabc[3 * q], b = colabc[3 * q + 1], c = colabc[3 * q + 2];
    double e[RB][6];
#pragma unroll
    for (int u = 0; u < RB; ++u) {
      const double* p = E + ((size_t)(r0 + min(u, nr - 1)) * n + jp) * 6;
#pragma unroll
      for (int m = 0; m < 6; ++m) e[u][m] = p[m];
    }
#pragma unroll
    for (int u = 0; u < RB; ++u) {
      // etk*a + ets*b + etc*c, then Z += term
      const double tr = ((e[u][0] * a) + (e[u][2] * b)) + (e[u][4] * c);
      const double ti = ((e[u][1] * a) + (e[u][3] * b)) + (e[u][5] * c);
      accr[u] += tr;
      acci[u] += ti;
      if (zload && colj[q] == i0 + r0 + u) {
        // the load of segment i, right after its own term
        // (src/nec_context.cpp:2126-2142): -= (A + C) z
        const int j = colj[q];
        const double sl = a + c;
        accr[u] -= sl * zload[2 * j];
        acci[u] -= sl * zload[2 * j + 1];
      }
    }
  }
#pragma unroll
  for (int u = 0; u < RB; ++u)
    if (u < nr) {
      double* zp = Z + 2 * ((size_t)(i0 + r0 + u) * ldz + k);
      zp[0] = accr[u];
      zp[1] = acci[u];
    }
}

// The fill in one pass.  A workgroup takes RB consecutive observation rows
// and one cluster of bases: segments consecutive in the Morton order of
// their centres, listed in ascending index (cl_base), together with every
// source their supports draw on (cl_src, ascending; a basis' support entry
// q finds its source at position cjl[q] of that list).  The fields of the
// rows against the cluster's sources go to LDS (six planes of RB x smax
// doubles), then each entry Z[i][k] gathers them over k's support in
// ascending source order with the arithmetic of fill_assemble_kernel, and
// the load of segment i right after its own term.  Sources shared by two
// clusters are evaluated once for each; Z is written once and the fields
// never leave the CU.
template <bool GROUND, int RB>
__global__ void __launch_bounds__(256) fill_tile_kernel(DevSegments g, int n, double rkh, int use_ek,
                                                        const int* __restrict__ ekind, GroundSpec gs,
                                                        const int* __restrict__ cl_src_ptr,
                                                        const int* __restrict__ cl_src,
                                                        const int* __restrict__ cl_base_ptr,
                                                        const int* __restrict__ cl_base,
                                                        const int* __restrict__ colptr, const int* __restrict__ colj,
                                                        const int* __restrict__ cjl,
                                                        const double* __restrict__ colabc,
                                                        const double* __restrict__ zload, int smax,
                                                        double* __restrict__ Z, long long ldz) {
  extern __shared__ double fe[];  // [6][RB][smax]
  const int q = blockIdx.x;
  const int i0 = blockIdx.y * RB;
  const int nr = min(RB, n - i0);
  const int s0 = cl_src_ptr[q];
  const int ns = cl_src_ptr[q + 1] - s0;
  const size_t pl = (size_t)RB * smax;
  for (int e = threadIdx.x; e < nr * ns; e += blockDim.x) {
    const int r = e / ns, sidx = e - r * ns;
    const int i = i0 + r, j = cl_src[s0 + sidx];
    int ind1 = 2, ind2 = 2;
    if (use_ek) {
      ind1 = ekind[2 * j];
      ind2 = ekind[2 * j + 1];
    }
    zc etk, ets, etc;
    if (GROUND) pair_fields_ground(g, i, j, rkh, use_ek, ind1, ind2, gs, &etk, &ets, &etc);
    else pair_fields(g, i, j, rkh, use_ek, ind1, ind2, &etk, &ets, &etc);
    double* o = fe + (size_t)r * smax + sidx;
    o[0] = etk.re;
    o[pl] = etk.im;
    o[2 * pl] = ets.re;
    o[3 * pl] = ets.im;
    o[4 * pl] = etc.re;
    o[5 * pl] = etc.im;
  }
  __syncthreads();
  const int b0 = cl_base_ptr[q];
  const int nbq = cl_base_ptr[q + 1] - b0;
  for (int e = threadIdx.x; e < nbq; e += blockDim.x) {
    const int k = cl_base[b0 + e];
    double accr[RB], acci[RB];
#pragma unroll
    for (int u = 0; u < RB; ++u) accr[u] = acci[u] = 0.0;
    const int q1 = colptr[k + 1];
    for (int qq = colptr[k]; qq < q1; ++qq) {
      const int sidx = cjl[qq];
      const double a = colabc[3 * qq], b = colabc[3 * qq + 1], c = colabc[3 * qq + 2];
#pragma unroll
      for (int u = 0; u < RB; ++u) {
        if (u < nr) {
          const double* f = fe + (size_t)u * smax + sidx;
          // etk*a + ets*b + etc*c, then Z += term
          const double tr = ((f[0] * a) + (f[2 * pl] * b)) + (f[4 * pl] * c);
          const double ti = ((f[pl] * a) + (f[3 * pl] * b)) + (f[5 * pl] * c);
          accr[u] += tr;
          acci[u] += ti;
          if (zload && colj[qq] == i0 + u) {
            // the load of segment i, right after its own term
            // (src/nec_context.cpp:2126-2142): -= (A + C) z
            const int j = colj[qq];
            const double sl = a + c;
            accr[u] -= sl * zload[2 * j];
            acci[u] -= sl * zload[2 * j + 1];
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < RB; ++u)
      if (u < nr) {
        double* zp = Z + 2 * ((size_t)(i0 + u) * ldz + k);
        zp[0] = accr[u];
        zp[1] = acci[u];
      }
  }
}

// Rows per thread of the fill assembly: 4, or 8 above 8,192 segments
// (NECGPU_FILL_RB = 1, 2, 4 or 8 overrides, for measurements).  With the
// Morton field order, the 3,280-segment grid assembles in 510 us with 4
// rows against 580 with 8; the 16,020-segment grid fills in 29.1 ms with 8
// rows against 30.7 with 4.
int fill_assemble_rows(int n) {
  const char* v = std::getenv("NECGPU_FILL_RB");
  const int r = (v && *v) ? std::atoi(v) : (n > 8192 ? 8 : 4);
  return (r == 1 || r == 2 || r == 8) ? r : 4;
}

// Batched fill: one workgroup per (solve, block of R consecutive rows), the
// block's solve and first row from the table blk (2 ints per workgroup, made
// on the host).  Threads run over the block's R x n (row, source) pairs, so
// lanes stay busy when n is not a multiple of the workgroup size; the fields
// go to LDS (R n 6 doubles, dynamic), then each entry Z[i][k] gathers them
// over basis k's support in ascending j, as fill_assemble does.  The
// register budget is capped at three waves per SIMD (167 VGPRs, a few
// spilled): batches of Yagi-sized structures (n = 63 .. 252) fill faster
// than at two waves (173 VGPRs) or four
// (128 VGPRs, heavy spilling).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
batch_fill_kernel(const BatchSolve* __restrict__ sol, const int* __restrict__ blk, int R,
                  const double* __restrict__ seg, const int* __restrict__ colptr, const int* __restrict__ colj,
                  const double* __restrict__ colabc, double* __restrict__ Zall, double rkh) {
  extern __shared__ double sE[];
  const BatchSolve d = sol[blk[2 * blockIdx.x]];
  const int r0 = blk[2 * blockIdx.x + 1];
  const int n = d.n;
  const int rows = min(R, n - r0);
  const int np = rows * n;
  const double* p = seg + d.seg0;
  DevSegments g;
  g.x = p;
  g.y = p + n;
  g.z = p + 2 * n;
  g.s = p + 3 * n;
  g.b = p + 4 * n;
  g.cab = p + 5 * n;
  g.sab = p + 6 * n;
  g.salp = p + 7 * n;
  for (int q = threadIdx.x; q < np; q += blockDim.x) {
    const int r = q / n, j = q - r * n;
    zc etk, ets, etc;
    pair_fields(g, r0 + r, j, rkh, 0, 2, 2, &etk, &ets, &etc);
    double* o = sE + 6 * q;
    o[0] = etk.re;
    o[1] = etk.im;
    o[2] = ets.re;
    o[3] = ets.im;
    o[4] = etc.re;
    o[5] = etc.im;
  }
  __syncthreads();
  const int* cp = colptr + d.cptr0;
  const int* cj = colj + d.csr0;
  const double* cv = colabc + 3 * d.csr0;
  double* Z = Zall + 2 * (d.z0 + (long long)r0 * n);
  for (int q = threadIdx.x; q < np; q += blockDim.x) {
    const int r = q / n, k = q - r * n;
    const double* er = sE + 6 * (r * n);
    double accr = 0.0, acci = 0.0;
    for (int e = cp[k]; e < cp[k + 1]; ++e) {
      const double* f = er + 6 * cj[e];
      const double a = cv[3 * e], b = cv[3 * e + 1], c = cv[3 * e + 2];
      accr += ((f[0] * a) + (f[2] * b)) + (f[4] * c);
      acci += ((f[1] * a) + (f[3] * b)) + (f[5] * c);
    }
    Z[2 * q] = accr;
    Z[2 * q + 1] = acci;
  }
}

// Adjoint contraction of the fill for one structure.  With lam the adjoint
// vector (Z^T lam = dJ/dx) and U_j = (A_j, B_j, C_j) the current
// coefficients of segment j (abc), pair (i, j) adds
//   -2 Re( lam_i * d(E_ij . U_j) / dv )
// to the gradient of each geometry variable v, where E_ij = (etk, ets, etc)
// are the fields of pair_fields_v.  The fields are differentiated on the fly
// (Dn, no dZ/dv is stored): with respect to the centre difference c_i - c_j
// (the fields depend on the centres through it only), the observation
// radius, the source length and the source direction; the observation
// direction enters through the projection of the unprojected components.
// One workgroup per observation row i (rows i0 .. i0+rows-1), threads over
// source segments j.  Row-side sums go to gi (7 per row: d/dc_i x y z,
// d/db_i, d/d(cab, sab, salp)_i); source-side terms of each pair go to pj
// (13 per pair: d/dc_j x y z, d/ds_j, d/d(cab, sab, salp)_j, and
// lam_i E_ij as 3 complex) and are summed over rows by adjoint_reduce in
// row order, so the result does not depend on scheduling.
constexpr int ADJ_PJ = 13;
constexpr int ADJ_GI = 7;
constexpr int ADJ_THREADS = 256;

// Fields of pair (i, j) with forward derivatives in one group of four of the
// eight variables: G = 0: dx, dy, dz (c_i - c_j; a segment's own pair has no
// difference to vary) and b_i; G = 1: s_j, cab_j, sab_j, salp_j.  Each
// derivative component is computed from the values and its own seed only, so
// two passes of four give the same numbers as one of eight, with half the
// registers live.
template <int G>
__device__ __forceinline__ void adjoint_fields(const DevSegments& g, int i, int j, double rkh, cx<Dn<4>>* etk,
                                               cx<Dn<4>>* ets, cx<Dn<4>>* etc, cx<double>* ev) {
  typedef Dn<4> D;
  const bool self = (i == j);
  D dx(g.x[i] - g.x[j]), dy(g.y[i] - g.y[j]), dz(g.z[i] - g.z[j]);
  D ai(g.b[i]), sj(g.s[j]), bj(g.b[j]), cj(g.cab[j]), sbj(g.sab[j]), slj(g.salp[j]);
  if (G == 0) {
    if (!self) {
      dx.d[0] = 1.0;
      dy.d[1] = 1.0;
      dz.d[2] = 1.0;
    }
    ai.d[3] = 1.0;
  } else {
    sj.d[0] = 1.0;
    cj.d[1] = 1.0;
    sbj.d[2] = 1.0;
    slj.d[3] = 1.0;
  }
  const D zero(0.0);
  pair_fields_v<D>(dx, dy, dz, ai, g.cab[i], g.sab[i], g.salp[i], zero, zero, zero, sj, bj, cj, sbj, slj, self, rkh, 0,
                   2, 2, etk, ets, etc, ev);
}

// One pair's adjoint terms: oi (7, observation side: d/dc_i x y z, d/db_i,
// d/d(cab, sab, salp)_i) and oj (13, source side: d/dc_j x y z, d/ds_j,
// d/d(cab, sab, salp)_j, lam_i E_ij as 3 complex).  abc: 6 n doubles of the
// solve (A, B, C real and imaginary planes).
__device__ __forceinline__ void adjoint_pair(const DevSegments& g, int n, int i, int j, double rkh, double lr,
                                             double li, const double* __restrict__ abc, double* oi, double* oj) {
  const double Ar = abc[j], Ai = abc[n + j], Br = abc[2 * n + j], Bi = abc[3 * n + j], Cr = abc[4 * n + j],
               Ci = abc[5 * n + j];
  // -2 Re(lam * (dk A + ds B + dc C)) for complex derivative triples
  auto contract = [&](double kr, double ki, double sr, double si, double cr, double ci) {
    const double fr = ((kr * Ar - ki * Ai) + (sr * Br - si * Bi)) + (cr * Cr - ci * Ci);
    const double fi = ((kr * Ai + ki * Ar) + (sr * Bi + si * Br)) + (cr * Ci + ci * Cr);
    return -2.0 * (lr * fr - li * fi);
  };
  double cv[8];
  {
    cx<Dn<4>> etk, ets, etc;
    cx<double> ev[9];
    adjoint_fields<0>(g, i, j, rkh, &etk, &ets, &etc, ev);
    for (int v = 0; v < 4; ++v)
      cv[v] = contract(etk.re.d[v], etk.im.d[v], ets.re.d[v], ets.im.d[v], etc.re.d[v], etc.im.d[v]);
    // observation direction: etk = e0 cab + e1 sab + e2 salp (e3..e5: ets,
    // e6..e8: etc)
    for (int q = 0; q < 3; ++q)
      oi[4 + q] = contract(ev[q].re, ev[q].im, ev[3 + q].re, ev[3 + q].im, ev[6 + q].re, ev[6 + q].im);
    oj[7] = lr * etk.re.v - li * etk.im.v;
    oj[8] = lr * etk.im.v + li * etk.re.v;
    oj[9] = lr * ets.re.v - li * ets.im.v;
    oj[10] = lr * ets.im.v + li * ets.re.v;
    oj[11] = lr * etc.re.v - li * etc.im.v;
    oj[12] = lr * etc.im.v + li * etc.re.v;
  }
  {
    cx<Dn<4>> etk, ets, etc;
    adjoint_fields<1>(g, i, j, rkh, &etk, &ets, &etc, nullptr);
    for (int v = 0; v < 4; ++v)
      cv[4 + v] = contract(etk.re.d[v], etk.im.d[v], ets.re.d[v], ets.im.d[v], etc.re.d[v], etc.im.d[v]);
  }
  oi[0] = cv[0];
  oi[1] = cv[1];
  oi[2] = cv[2];
  oi[3] = cv[3];
  oj[0] = -cv[0];
  oj[1] = -cv[1];
  oj[2] = -cv[2];
  oj[3] = cv[4];
  oj[4] = cv[5];
  oj[5] = cv[6];
  oj[6] = cv[7];
}

// Sum of K per-thread values over the workgroup (ADJ_THREADS threads);
// thread c < K writes total c to out[c].
template <int K>
__device__ __forceinline__ void block_sum_out(const double* acc, double (*sred)[ADJ_THREADS / 64], double* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int v = 0; v < K; ++v) {
    double x = acc[v];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) sred[v][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double x = 0.0;
    for (int w = 0; w < ADJ_THREADS / 64; ++w) x += sred[threadIdx.x][w];
    out[threadIdx.x] = x;
  }
}

// Adjoint contraction of the fill for one structure.  With lam the adjoint
// vector (Z^T lam = dJ/dx) and U_j = (A_j, B_j, C_j) the current
// coefficients of segment j (abc), pair (i, j) adds
//   -2 Re( lam_i * d(E_ij . U_j) / dv )
// to the gradient of each geometry variable v, where E_ij = (etk, ets, etc)
// are the fields of pair_fields_v.  The fields are differentiated on the fly
// (Dn, no dZ/dv is stored): with respect to the centre difference c_i - c_j
// (the fields depend on the centres through it only), the observation
// radius, the source length and the source direction; the observation
// direction enters through the projection of the unprojected components.
// One workgroup per observation row i (rows i0 .. i0+rows-1), threads over
// source segments j.  Row-side sums go to gi (7 per row); source-side terms
// of each pair go to pj (13 per pair) and are summed over rows by
// adjoint_reduce in row order, so the result does not depend on scheduling.
__global__ void __launch_bounds__(ADJ_THREADS) adjoint_fill_kernel(DevSegments g, int n, int i0, double rkh,
                                                                   const double* __restrict__ lam,
                                                                   const double* __restrict__ abc,
                                                                   double* __restrict__ gi, double* __restrict__ pj) {
  __shared__ double sred[ADJ_PJ][ADJ_THREADS / 64];
  const int r = blockIdx.x;
  const int i = i0 + r;
  const double lr = lam[2 * i], li = lam[2 * i + 1];
  double acc[ADJ_GI] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double oi[ADJ_GI];
    double* o = pj + ((size_t)r * n + j) * ADJ_PJ;
    double oj[ADJ_PJ];
    adjoint_pair(g, n, i, j, rkh, lr, li, abc, oi, oj);
    for (int v = 0; v < ADJ_GI; ++v) acc[v] += oi[v];
    for (int v = 0; v < ADJ_PJ; ++v) o[v] = oj[v];
  }
  block_sum_out<ADJ_GI>(acc, sred, gi + (size_t)i * ADJ_GI);
}

// Batched contraction: one workgroup per block of ADJ_ROWS consecutive
// observation rows of one solve (table blk: solve, first row), threads over
// the sources j.  Each pair is evaluated once.  Its observation-side terms
// are summed over the workgroup per row (gi, 7 per row); its source-side
// terms are summed per source over the block's rows, in row order, in LDS and
// written to the block's slice of part (13 x maxn, [term][j]);
// batch_adjoint_reduce adds the slices of a solve in block order.  Every sum
// runs in a fixed order, so the result does not depend on scheduling.
constexpr int ADJ_ROWS = 16;
__global__ void __launch_bounds__(ADJ_THREADS) batch_adjoint_kernel(const BatchSolve* __restrict__ sol,
                                                                    const int* __restrict__ blk, int maxn,
                                                                    const double* __restrict__ seg,
                                                                    const double* __restrict__ lamall,
                                                                    const double* __restrict__ abcall, double rkh,
                                                                    double* __restrict__ gi,
                                                                    double* __restrict__ part) {
  __shared__ double sacc[ADJ_PJ * kBatchMaxN];  // [term][j]
  __shared__ double sred[2][ADJ_PJ][ADJ_THREADS / 64];
  const BatchSolve d = sol[blk[2 * blockIdx.x]];
  const int r0 = blk[2 * blockIdx.x + 1];
  const int n = d.n;
  const int rows = min(ADJ_ROWS, n - r0);
  const double* p = seg + d.seg0;
  DevSegments g;
  g.x = p;
  g.y = p + n;
  g.z = p + 2 * n;
  g.s = p + 3 * n;
  g.b = p + 4 * n;
  g.cab = p + 5 * n;
  g.sab = p + 6 * n;
  g.salp = p + 7 * n;
  const double* lam = lamall + 2 * d.row0;
  const double* abc = abcall + 6 * d.row0;
  // each thread owns the sacc entries of its own sources: no barrier needed
  for (int j = threadIdx.x; j < n; j += ADJ_THREADS)
    for (int v = 0; v < ADJ_PJ; ++v) sacc[v * kBatchMaxN + j] = 0.0;
  for (int r = 0; r < rows; ++r) {
    const int i = r0 + r;
    const double lr = lam[2 * i], li = lam[2 * i + 1];
    double acc[ADJ_GI] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < n; j += ADJ_THREADS) {
      double oi[ADJ_GI], oj[ADJ_PJ];
      adjoint_pair(g, n, i, j, rkh, lr, li, abc, oi, oj);
      for (int v = 0; v < ADJ_GI; ++v) acc[v] += oi[v];
      for (int v = 0; v < ADJ_PJ; ++v) sacc[v * kBatchMaxN + j] += oj[v];
    }
    // alternate reduction buffers: row r + 1 may write while row r is read
    block_sum_out<ADJ_GI>(acc, sred[r & 1], gi + (size_t)(d.row0 + i) * ADJ_GI);
  }
  double* o = part + (size_t)blockIdx.x * ADJ_PJ * maxn;
  for (int j = threadIdx.x; j < n; j += ADJ_THREADS)
    for (int v = 0; v < ADJ_PJ; ++v) o[(size_t)v * maxn + j] = sacc[v * kBatchMaxN + j];
}

// gj[row0 + j][v] = sum over the blocks fb[s] .. fb[s+1]-1 of solve s of
// their part slices, in block order.  Grid: (solve, chunk of 13 n).
__global__ void __launch_bounds__(256) batch_adjoint_reduce_kernel(const BatchSolve* __restrict__ sol,
                                                                   const int* __restrict__ fb, int maxn,
                                                                   const double* __restrict__ part,
                                                                   double* __restrict__ gj) {
  const int s = blockIdx.x;
  const int n = sol[s].n;
  const int e = blockIdx.y * blockDim.x + threadIdx.x;
  if (e >= ADJ_PJ * n) return;
  const int v = e / n, j = e - v * n;
  double x = 0.0;
  for (int b = fb[s]; b < fb[s + 1]; ++b) x += part[((size_t)b * ADJ_PJ + v) * maxn + j];
  gj[(size_t)(sol[s].row0 + j) * ADJ_PJ + v] = x;
}

// gj[j][c] += sum over the rows of a chunk, in row order.
__global__ void __launch_bounds__(256) adjoint_reduce_kernel(const double* __restrict__ pj, int n, int rows,
                                                             double* __restrict__ gj) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * ADJ_PJ) return;
  double x = gj[t];
  for (int r = 0; r < rows; ++r) x += pj[(size_t)r * n * ADJ_PJ + t];
  gj[t] = x;
}

}  // namespace

int adjoint_rows_per_chunk(int n, size_t budget_bytes) {
  const size_t per_row = (size_t)n * ADJ_PJ * sizeof(double);
  size_t rows = budget_bytes / (per_row ? per_row : 1);
  if (rows < 1) rows = 1;
  if (rows > (size_t)n) rows = n;
  return static_cast<int>(rows);
}

int launch_adjoint_fill(const DevSegments& g, int n, int i0, int rows, double rkh, const double* lam,
                        const double* abc, double* gi, double* pj, double* gj, hipStream_t stream) {
  if (n <= 0 || rows <= 0) return 0;
  launch_k(K_ADJOINT_FILL, adjoint_fill_kernel, dim3(rows), dim3(ADJ_THREADS), 0u, stream, g, n, i0, rkh, lam, abc,
           gi, pj);
  launch_k(K_ADJOINT_REDUCE, adjoint_reduce_kernel, dim3((n * ADJ_PJ + 255) / 256), dim3(256), 0u, stream, pj, n,
           rows, gj);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int batch_adjoint_rows() { return ADJ_ROWS; }

int launch_batch_adjoint_fill(const BatchSolve* sol, int S, const int* blk, const int* fb, int nblk, int maxn,
                              const double* seg, const double* lam, const double* abc, double rkh, double* gi,
                              double* part, double* gj, hipStream_t stream) {
  if (S <= 0 || nblk <= 0) return 0;
  if (maxn > kBatchMaxN) return 1;
  launch_k(K_ADJOINT_FILL, batch_adjoint_kernel, dim3((unsigned)nblk), dim3(ADJ_THREADS), 0u, stream, sol, blk, maxn,
           seg, lam, abc, rkh, gi, part);
  launch_k(K_ADJOINT_REDUCE, batch_adjoint_reduce_kernel, dim3((unsigned)S, (ADJ_PJ * maxn + 255) / 256), dim3(256),
           0u, stream, sol, fb, maxn, part, gj);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int batch_fill_rows(int maxn) { return std::max(1, std::min(8, 1024 / std::max(1, maxn))); }

int launch_batch_fill(const BatchSolve* sol, const int* blk, int nblk, int R, int maxn, const double* seg,
                      const int* colptr, const int* colj, const double* colabc, double* Z, double rkh,
                      hipStream_t stream) {
  if (nblk <= 0) return 0;
  const uint32_t lds = (uint32_t)R * (uint32_t)maxn * 6u * sizeof(double);
  launch_k(K_BATCH_FILL, batch_fill_kernel, dim3((unsigned)nblk), dim3(256), lds, stream, sol, blk, R, seg, colptr,
           colj, colabc, Z, rkh);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

namespace {

// ------------------------------------------------------------ near fields

// gh, src/nec_context.cpp:3582-3596: hr + j hi = (sin r / r^2 + cos r /
// r^3) + j (cos r / r^2 - sin r / r^3), r the distance from (rho^2 = rhks,
// z = zpka) to the axis point zk (radians).
__device__ __forceinline__ void gh(double zk, double zpka, double rhks, double* hr, double* hi) {
  const double dz = zk - zpka;
  const double rs = rhks + dz * dz;
  const double r = sqrt(rs);
  double skr, ckr;
  sincos(r, &skr, &ckr);
  const double rr2 = 1. / rs;
  const double rr3 = rr2 / r;
  *hr = skr * rr2 + ckr * rr3;
  *hi = ckr * rr2 - skr * rr3;
}

// hfk, src/nec_context.cpp:3681-3822: gh integrated over [el1, el2] by the
// adaptive Romberg scheme of intx (the same decisions), times rhk / 2.
__device__ void hfk(double el1, double el2, double rhk, double zpka, double* sgr, double* sgi) {
  const int nx = 1, nma = 65536, nts = 4;
  const double rx = 1.0e-4;
  const double rhks = rhk * rhk;
  double z = el1;
  const double ze = el2;
  const double s = ze - z;
  const double ep = s / (10. * nma);
  const double zend = ze - ep;
  double accr = 0.0, acci = 0.0;
  int ns = nx, nt = 0;
  bool fresh = true;
  double dz = 0., dzot = 0., g1r, g1i, g2r, g2i, g3r = 0., g3i = 0., g4r, g4i, g5r = 0., g5i = 0.;
  gh(z, zpka, rhks, &g1r, &g1i);
  // at most 65536 subdivisions of a finite interval; the guard only stops a
  // non-finite one from spinning
  for (long guard = 0; guard < (1L << 22); ++guard) {
    if (fresh) {
      const double fns = ns;
      dz = s / fns;
      if (z + dz > ze) {
        dz = ze - z;
        if (fabs(dz) <= ep) break;
      }
      dzot = dz * .5;
      gh(z + dzot, zpka, rhks, &g3r, &g3i);
      gh(z + dz, zpka, rhks, &g5r, &g5i);
    }
    const double t00r = (g1r + g5r) * dzot, t00i = (g1i + g5i) * dzot;
    const double t01r = (t00r + dz * g3r) * 0.5, t01i = (t00i + dz * g3i) * 0.5;
    const double t10r = (4.0 * t01r - t00r) / 3.0, t10i = (4.0 * t01i - t00i) / 3.0;
    double te1r, te1i;
    conv_test(t01r, t10r, &te1r, t01i, t10i, &te1i);
    double addr, addi;
    if ((te1i <= rx) && (te1r <= rx)) {
      addr = t10r;
      addi = t10i;
      nt += 2;
    } else {
      gh(z + dz * 0.25, zpka, rhks, &g2r, &g2i);
      gh(z + dz * 0.75, zpka, rhks, &g4r, &g4i);
      const double t02r = (t01r + dzot * (g2r + g4r)) * 0.5, t02i = (t01i + dzot * (g2i + g4i)) * 0.5;
      const double t11r = (4.0 * t02r - t01r) / 3.0, t11i = (4.0 * t02i - t01i) / 3.0;
      const double t20r = (16.0 * t11r - t10r) / 15.0, t20i = (16.0 * t11i - t10i) / 15.0;
      double te2r, te2i;
      conv_test(t11r, t20r, &te2r, t11i, t20i, &te2i);
      const bool bad = (te2i > rx) || (te2r > rx);
      if (bad && ns < nma) {
        nt = 0;
        ns *= 2;
        const double fns = ns;
        dz = s / fns;
        dzot = dz * 0.5;
        g5r = g3r;
        g5i = g3i;
        g3r = g2r;
        g3i = g2i;
        fresh = false;
        continue;
      }
      if (bad) nt = 0;  // step size limited: the 5-point value is accepted
      addr = t20r;
      addi = t20i;
      nt++;
    }
    accr += addr;
    acci += addi;
    z += dz;
    if (z >= zend) break;
    g1r = g5r;
    g1i = g5i;
    if (nt >= nts && ns > nx) {
      ns /= 2;
      nt = 1;
    }
    fresh = true;
  }
  *sgr = accr * rhk * .5;
  *sgi = acci * rhk * .5;
}

// hsflx, src/nec_context.cpp:4059-4131: the azimuthal H of a segment's
// constant (hpk), sine (hps) and cosine (hpc) currents at radial distance rh
// and axial offset zpx from its centre (wavelengths).
__device__ void hsflx(double s, double rh, double zpx, zc* hpk, zc* hps, zc* hpc) {
  if (rh < 1.0e-10) {
    *hpk = *hps = *hpc = zmk(0.0, 0.0);
    return;
  }
  const double hss = zpx >= 0.0 ? 1.0 : -1.0;
  const double zp = fabs(zpx);
  const double dh = 0.5 * s;
  const double z1 = zp + dh, z2a = zp - dh;
  const double rhz = z2a >= 1.0e-7 ? rh / z2a : 1.0;
  const double dk = kTwoPi * dh;
  double sdk, cdk;
  sincos(dk, &sdk, &cdk);
  double hkr, hki;
  hfk(-dk, dk, rh * kTwoPi, zp * kTwoPi, &hkr, &hki);
  *hpk = zmk(hkr, hki);
  if (rhz >= 1.0e-3) {
    const double rh2 = rh * rh;
    const double r1 = sqrt(rh2 + z1 * z1), r2 = sqrt(rh2 + z2a * z2a);
    double s1, c1, s2, c2;
    sincos(kTwoPi * r1, &s1, &c1);
    sincos(kTwoPi * r2, &s2, &c2);
    const zc ekr1 = zmk(c1, -s1), ekr2 = zmk(c2, -s2);  // exp(-j 2 pi r)
    const zc t1 = zdivr(zmulr(ekr1, z1), r1);
    const zc t2 = zdivr(zmulr(ekr2, z2a), r2);
    // hps = (cdk (ekr2 - ekr1) - j sdk (t2 + t1)) hss, hpc = -sdk (ekr2 +
    // ekr1) - j cdk (t2 - t1), both times -j / (4 pi rh)
    const zc d = zsub(ekr2, ekr1), sm = zadd(t2, t1);
    const zc ps = zmulr(zmk(cdk * d.re + sdk * sm.im, cdk * d.im - sdk * sm.re), hss);
    const zc a2 = zadd(ekr2, ekr1), df = zsub(t2, t1);
    const zc pc = zmk(-sdk * a2.re + cdk * df.im, -sdk * a2.im - cdk * df.re);
    const double cons = -1.0 / (2.0 * kTwoPi * rh);
    *hps = zmk(-cons * ps.im, cons * ps.re);
    *hpc = zmk(-cons * pc.im, cons * pc.re);
    return;
  }
  // near the axis beyond an end
  const zc ekr1 = zdivr(zmk(cdk, sdk), z2a * z2a);
  const zc ekr2 = zdivr(zmk(cdk, -sdk), z1 * z1);
  const double t1 = kTwoPi * (1.0 / z1 - 1.0 / z2a);
  double sz, cz;
  sincos(kTwoPi * zp, &sz, &cz);
  const zc t2 = zdivr(zmulr(zmk(cz, -sz), rh), kPi * 8.0);
  const zc sps = zadd(zmk(t1, 0.0), zmulr(zadd(ekr1, ekr2), sdk));
  *hps = zmulr(zmul(t2, sps), hss);
  const zc dpc = zmulr(zsub(ekr1, ekr2), cdk);
  *hpc = zmul(t2, zmk(dpc.re, dpc.im - t1));
}

// hsfld, src/nec_context.cpp:3927-4055: H at (xi, yi, zi) (observation
// radius ai) of segment j's constant / sine / cosine currents, h[3 m + axis]
// for m = k, s, c.  Over a ground the image segment (z and its z cosine
// negated) is subtracted, weighted for the finite ground by the reflection
// coefficients of the specular direction; a point on an axis (rho <= 1e-10)
// zeroes the nine components, the image's too.
template <bool GROUND>
__device__ void hsfld(const DevSegments& g, int j, double xi, double yi, double zi, double ai,
                      const GroundSpec& gs, zc h[9]) {
  const double xij = xi - g.x[j], yij = yi - g.y[j];
  const double cabj = g.cab[j], sabj = g.sab[j];
#pragma unroll
  for (int q = 0; q < 9; ++q) h[q] = zmk(0.0, 0.0);
  for (int loop = 0; loop < (GROUND ? 2 : 1); ++loop) {
    const double rfl = loop == 0 ? 1.0 : -1.0;
    const double salpr = g.salp[j] * rfl;
    const double zij = zi - rfl * g.z[j];
    const double zp = xij * cabj + yij * sabj + zij * salpr;
    double rhox = xij - cabj * zp, rhoy = yij - sabj * zp, rhoz = zij - salpr * zp;
    const double rh = sqrt(rhox * rhox + rhoy * rhoy + rhoz * rhoz + ai * ai);
    if (rh <= 1.0e-10) {
#pragma unroll
      for (int q = 0; q < 9; ++q) h[q] = zmk(0.0, 0.0);
      continue;
    }
    rhox = rhox / rh;
    rhoy = rhoy / rh;
    rhoz = rhoz / rh;
    const double phx = sabj * rhoz - salpr * rhoy;
    const double phy = salpr * rhox - cabj * rhoz;
    const double phz = cabj * rhoy - sabj * rhox;
    zc hp[3];
    hsflx(g.s[j], rh, zp, &hp[0], &hp[1], &hp[2]);
    if (loop == 0) {
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        h[3 * m] = zmulr(hp[m], phx);
        h[3 * m + 1] = zmulr(hp[m], phy);
        h[3 * m + 2] = zmulr(hp[m], phz);
      }
      continue;
    }
    if (gs.type == 1) {
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        h[3 * m] = zsub(h[3 * m], zmulr(hp[m], phx));
        h[3 * m + 1] = zsub(h[3 * m + 1], zmulr(hp[m], phy));
        h[3 * m + 2] = zsub(h[3 * m + 2], zmulr(hp[m], phz));
      }
      continue;
    }
    // finite ground: vertical (rrv) and horizontal (rrh) reflection coefficients
    const zc zratx = zmk(gs.zr_re, gs.zr_im);
    const double rmag = sqrt(zp * zp + rh * rh);
    const double xymag = sqrt(xij * xij + yij * yij);
    double px, py, cth;
    zc rrv;
    if (xymag <= 1.0e-6) {
      px = 0.;
      py = 0.;
      cth = 1.;
      rrv = zmk(1.0, 0.0);
    } else {
      px = -yij / xymag;
      py = xij / xymag;
      cth = zij / rmag;
      const zc q = zmulr(zmul(zratx, zratx), 1.0 - cth * cth);
      rrv = zsqrt(zmk(1.0 - q.re, -q.im));
    }
    zc rrh = zmulr(zratx, cth);
    rrh = zdiv(zneg(zsub(rrh, rrv)), zadd(rrh, rrv));
    rrv = zmul(zratx, rrv);
    rrv = zdiv(zmk(cth - rrv.re, -rrv.im), zmk(cth + rrv.re, rrv.im));
    zc qy = zmulr(zsub(rrv, rrh), phx * px + phy * py);
    const zc qx = zadd(zmulr(qy, px), zmulr(rrh, phx));
    qy = zadd(zmulr(qy, py), zmulr(rrh, phy));
    const zc qz = zmulr(rrh, phz);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      h[3 * m] = zsub(h[3 * m], zmul(hp[m], qx));
      h[3 * m + 1] = zsub(h[3 * m + 1], zmul(hp[m], qy));
      h[3 * m + 2] = zsub(h[3 * m + 2], zmul(hp[m], qz));
    }
  }
}

// nefld (src/nec_context.cpp:4304-4450, MAGNETIC false) or nhfld (:5189-5262,
// MAGNETIC true) for wires: one thread per (point, source segment) evaluates
// the nine field components of the segment's currents at the point (the
// observation radius ax of the point's own segment or 0; for E efld with
// ij = true, over a ground with its image) and contracts them with the
// segment's current coefficients A, B, C; the 256 segments of a block are
// summed in a fixed tree order, the blocks in order by near_finish_kernel.
template <bool MAGNETIC, bool GROUND>
__global__ void __launch_bounds__(256) near_field_kernel(DevSegments g, int n, const double* __restrict__ pts,
                                                         const double* __restrict__ ax,
                                                         const double* __restrict__ abc, int use_ek,
                                                         const int* __restrict__ ekind, double rkh, GroundSpec gs,
                                                         double* __restrict__ part) {
  __shared__ double sred[6][256];
  const int t = threadIdx.x;
  const int j = blockIdx.x * 256 + t;
  const int p = blockIdx.y;
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (j < n) {
    zc e[9];
    if constexpr (MAGNETIC) {
      hsfld<GROUND>(g, j, pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], ax[p], gs, e);
    } else {
      int ind1 = 2, ind2 = 2;
      if (use_ek) {
        ind1 = ekind[2 * j];
        ind2 = ekind[2 * j + 1];
      }
      if constexpr (GROUND) {
        point_fields_ground(g, j, pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], ax[p], false, rkh, use_ek, ind1, ind2,
                            gs, e);
      } else {
        cx<double> etk, ets, etc, ev[9];
        pair_fields_v<double>(pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], ax[p], 1.0, 0.0, 0.0, g.x[j], g.y[j],
                              g.z[j], g.s[j], g.b[j], g.cab[j], g.sab[j], g.salp[j], false, rkh, use_ek, ind1, ind2,
                              &etk, &ets, &etc, ev);
#pragma unroll
        for (int q = 0; q < 9; ++q) e[q] = zmk(ev[q].re, ev[q].im);
      }
    }
    const zc a = zmk(abc[j], abc[n + j]), b = zmk(abc[2 * n + j], abc[3 * n + j]),
             c = zmk(abc[4 * n + j], abc[5 * n + j]);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const zc f = zadd(zadd(zmul(e[q], a), zmul(e[q + 3], b)), zmul(e[q + 6], c));
      v[2 * q] = f.re;
      v[2 * q + 1] = f.im;
    }
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) sred[q][t] = v[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int q = 0; q < 6; ++q) sred[q][t] += sred[q][t + w];
    __syncthreads();
  }
  if (t < 6) part[((size_t)p * gridDim.x + blockIdx.x) * 6 + t] = sred[t][0];
}

__global__ void near_finish_kernel(const double* __restrict__ part, int npts, int nblk, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= npts * 6) return;
  const int p = e / 6, q = e % 6;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[((size_t)p * nblk + b) * 6 + q];
  out[e] = s;
}

}  // namespace

int launch_near_field(bool magnetic, const DevSegments& g, int n, const double* pts, const double* ax, int npts,
                      const double* abc, int use_ek, const int* ekind, double rkh, const GroundSpec& ground,
                      double* part, double* out, hipStream_t stream) {
  if (n <= 0 || npts <= 0) return 0;
  const int nblk = (n + 255) / 256;
  const dim3 grid(nblk, npts), block(256);
  const bool gr = ground.type != 0;
  if (magnetic && gr)
    launch_k(K_NEAR_FIELD, near_field_kernel<true, true>, grid, block, 0u, stream, g, n, pts, ax, abc, use_ek, ekind,
             rkh, ground, part);
  else if (magnetic)
    launch_k(K_NEAR_FIELD, near_field_kernel<true, false>, grid, block, 0u, stream, g, n, pts, ax, abc, use_ek,
             ekind, rkh, ground, part);
  else if (gr)
    launch_k(K_NEAR_FIELD, near_field_kernel<false, true>, grid, block, 0u, stream, g, n, pts, ax, abc, use_ek,
             ekind, rkh, ground, part);
  else
    launch_k(K_NEAR_FIELD, near_field_kernel<false, false>, grid, block, 0u, stream, g, n, pts, ax, abc, use_ek,
             ekind, rkh, ground, part);
  launch_k(K_NEAR_FIELD, near_finish_kernel, dim3((npts * 6 + 255) / 256), dim3(256), 0u, stream, part, npts, nblk,
           out);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_fill(const DevSegments& g, int n, int i0, int rows, double rkh, int use_ek, const int* ekind,
                const GroundSpec& ground, const double* zload, double* E, const int* colptr, const int* colj,
                const int* cjp, const int* ord, const double* colabc, double* Z, long long ldz, hipStream_t stream) {
  if (n <= 0 || rows <= 0) return 0;
  dim3 block(256);
  dim3 grid((n + 255) / 256, rows);
  if (ground.type != 0)
    launch_k(K_FILL_FIELDS, fill_fields_kernel<true>, grid, block, 0, stream, g, n, i0, rows, rkh, use_ek, ekind,
             ground, ord, E);
  else
    launch_k(K_FILL_FIELDS, fill_fields_kernel<false>, grid, block, 0, stream, g, n, i0, rows, rkh, use_ek, ekind,
             ground, ord, E);
  const int rb = fill_assemble_rows(n);
  const dim3 agrid((n + 255) / 256, (rows + rb - 1) / rb);
  if (rb == 8)
    launch_k(K_FILL_ASSEMBLE, fill_assemble_kernel<8>, agrid, block, 0, stream, E, n, i0, rows, colptr, colj, cjp,
             ord, colabc, zload, Z, ldz);
  else if (rb == 4)
    launch_k(K_FILL_ASSEMBLE, fill_assemble_kernel<4>, agrid, block, 0, stream, E, n, i0, rows, colptr, colj, cjp,
             ord, colabc, zload, Z, ldz);
  else if (rb == 2)
    launch_k(K_FILL_ASSEMBLE, fill_assemble_kernel<2>, agrid, block, 0, stream, E, n, i0, rows, colptr, colj, cjp,
             ord, colabc, zload, Z, ldz);
  else
    launch_k(K_FILL_ASSEMBLE, fill_assemble_kernel<1>, agrid, block, 0, stream, E, n, i0, rows, colptr, colj, cjp,
             ord, colabc, zload, Z, ldz);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The one-pass fill (fill_tile_kernel) over all rows; rb rows per
// workgroup (1, 2 or 4), smax the longest cluster source list.
int launch_fill_tile(const DevSegments& g, int n, double rkh, int use_ek, const int* ekind, const GroundSpec& ground,
                     const double* zload, int nclusters, const int* cl_src_ptr, const int* cl_src,
                     const int* cl_base_ptr, const int* cl_base, const int* colptr, const int* colj, const int* cjl,
                     const double* colabc, int smax, int rb, double* Z, long long ldz, hipStream_t stream) {
  if (n <= 0 || nclusters <= 0) return 0;
  const size_t lds = (size_t)6 * rb * smax * sizeof(double);
  const dim3 grid(nclusters, (n + rb - 1) / rb), block(256);
  auto go = [&](auto kern) {
    static bool attr = false;  // one attribute per instantiation
    if (!attr) {
      if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kFillTileLdsMax) != hipSuccess)
        return 1;
      attr = true;
    }
    launch_k(K_FILL_TILE, kern, grid, block, (uint32_t)lds, stream, g, n, rkh, use_ek, ekind, ground, cl_src_ptr,
             cl_src, cl_base_ptr, cl_base, colptr, colj, cjl, colabc, zload, smax, Z, ldz);
    return 0;
  };
  int rc;
  if (ground.type != 0)
    rc = rb == 4 ? go(fill_tile_kernel<true, 4>) : rb == 2 ? go(fill_tile_kernel<true, 2>) : go(fill_tile_kernel<true, 1>);
  else
    rc = rb == 4 ? go(fill_tile_kernel<false, 4>) : rb == 2 ? go(fill_tile_kernel<false, 2>)
                                                    : go(fill_tile_kernel<false, 1>);
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace necgpu
