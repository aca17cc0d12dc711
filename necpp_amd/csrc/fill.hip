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
#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

struct GfState {
  int ija;
  double zpk, rkb2;
};

// gf, src/nec_context.cpp:3553-3576
__device__ __forceinline__ void gf(const GfState& st, double zk, double* co, double* si) {
  const double zdk = zk - st.zpk;
  const double rk = sqrt(st.rkb2 + zdk * zdk);
  double s, c;
  sincos(rk, &s, &c);
  *si = s / rk;
  if (st.ija != 0) {
    *co = c / rk;
    return;
  }
  if (rk >= .2) {
    *co = (c - 1.0) / rk;
    return;
  }
  const double rks = rk * rk;
  *co = ((-1.38888889e-3 * rks + 4.16666667e-2) * rks - 0.5) * rk;
}

// test, src/c_evlcom.cpp:845-868 (dmin = 0)
__device__ __forceinline__ void conv_test(double f1r, double f2r, double* tr, double f1i, double f2i, double* ti) {
  double den = fabs(f2r);
  const double t = fabs(f2i);
  if (den < t) den = t;
  if (den < 1.0e-37) {
    *tr = 0.0;
    *ti = 0.0;
    return;
  }
  *tr = fabs((f1r - f2r) / den);
  *ti = fabs((f1i - f2i) / den);
}

// intx, src/nec_context.cpp:4135-4297: variable-step Romberg integration of
// the gf integrand; for the self term (ij == 0) half the segment plus the
// analytic log term.
__device__ void intx(const GfState& st, double el1, double el2, double b, int ij, double* sgr, double* sgi) {
  const int nx = 1, nma = 65536, nts = 4;
  const double rx = 1.0e-4;
  double z = el1;
  double ze = el2;
  if (ij == 0) ze = 0.;
  const double s = ze - z;
  const double fnm = nma;
  const double ep = s / (10. * fnm);
  const double zend = ze - ep;
  double accr = 0., acci = 0.;
  int ns = nx, nt = 0;
  bool fresh = true;
  double dz = 0., dzot = 0.;
  double g1r, g1i, g2r, g2i, g3r = 0, g3i = 0, g4r, g4i, g5r = 0, g5i = 0;
  gf(st, z, &g1r, &g1i);
  // The step count is bounded by 65536 subdivisions of a finite interval;
  // the guard only stops a non-finite interval from spinning forever.
  long guard = 0;
  if (!(fabs(s) < 1.0e300)) {
    *sgr = s;
    *sgi = s;
    return;
  }
  for (; guard < (1L << 22); ++guard) {
    if (fresh) {
      const double fns = ns;
      dz = s / fns;
      double zp = z + dz;
      if (zp > ze) {
        dz = ze - z;
        if (fabs(dz) <= ep) break;
      }
      dzot = dz * .5;
      zp = z + dzot;
      gf(st, zp, &g3r, &g3i);
      zp = z + dz;
      gf(st, zp, &g5r, &g5i);
    }
    const double t00r = (g1r + g5r) * dzot;
    const double t00i = (g1i + g5i) * dzot;
    const double t01r = (t00r + dz * g3r) * 0.5;
    const double t01i = (t00i + dz * g3i) * 0.5;
    const double t10r = (4.0 * t01r - t00r) / 3.0;
    const double t10i = (4.0 * t01i - t00i) / 3.0;
    double te1r, te1i;
    conv_test(t01r, t10r, &te1r, t01i, t10i, &te1i);
    if ((te1i <= rx) && (te1r <= rx)) {
      accr = accr + t10r;
      acci = acci + t10i;
      nt += 2;
      z += dz;
      if (z >= zend) break;
      g1r = g5r;
      g1i = g5i;
      if (nt >= nts && ns > nx) {
        ns = ns / 2;
        nt = 1;
      }
      fresh = true;
      continue;
    }
    double zp = z + dz * 0.25;
    gf(st, zp, &g2r, &g2i);
    zp = z + dz * 0.75;
    gf(st, zp, &g4r, &g4i);
    const double t02r = (t01r + dzot * (g2r + g4r)) * 0.5;
    const double t02i = (t01i + dzot * (g2i + g4i)) * 0.5;
    const double t11r = (4.0 * t02r - t01r) / 3.0;
    const double t11i = (4.0 * t02i - t01i) / 3.0;
    const double t20r = (16.0 * t11r - t10r) / 15.0;
    const double t20i = (16.0 * t11i - t10i) / 15.0;
    double te2r, te2i;
    conv_test(t11r, t20r, &te2r, t11i, t20i, &te2i);
    if ((te2i > rx) || (te2r > rx)) {
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
    z += dz;
    if (z >= zend) break;
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
__device__ __forceinline__ void gx(double zz, double rh, double xk, zc* gz, zc* gzp) {
  const double r2 = zz * zz + rh * rh;
  const double r = sqrt(r2);
  const double rkz = xk * r;
  double s, c;
  sincos(rkz, &s, &c);
  const zc temp = zdivr(zmk(c, -s), r);
  *gz = temp;
  *gzp = zdivr(zmul(zneg(zmk(1.0, rkz)), temp), r2);
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

// eksc, src/nec_context.cpp:3051-3100
__device__ void eksc(double s, double z, double rh, double xk, int ij, NearOut* o) {
  const zc k1 = zmk(0.0, kNearConst), mk1 = zneg(k1);
  GfState st;
  st.ija = ij;
  st.zpk = xk * z;
  const double rhk = xk * rh;
  st.rkb2 = rhk * rhk;
  const double sh = .5 * s;
  const double shk = xk * sh;
  double ss, cs;
  sincos(shk, &ss, &cs);
  const double z2a = sh - z;
  const double z1a = -(sh + z);
  zc gz1, gz2, gp1, gp2;
  gx(z1a, rh, xk, &gz1, &gp1);
  gx(z2a, rh, xk, &gz2, &gp2);
  zc gzp1 = zmulr(gp1, z1a);
  zc gzp2 = zmulr(gp2, z2a);
  o->ezs = zmul(k1, zsub(zmulr(zmulr(zsub(gz2, gz1), cs), xk), zmulr(zadd(gzp2, gzp1), ss)));
  o->ezc = zmul(mk1, zadd(zmulr(zmulr(zadd(gz2, gz1), ss), xk), zmulr(zsub(gzp2, gzp1), cs)));
  o->erk = zmulr(zmul(k1, zsub(gp2, gp1)), rh);
  double cint, sint;
  intx(st, -shk, shk, rhk, ij, &cint, &sint);
  o->ezk = zmul(mk1, zadd(zsub(gzp2, gzp1), zmulr(zmk(cint, -sint), xk * xk)));
  if (rh >= 1.0e-10) {
    gzp1 = zmulr(gzp1, z1a);
    gzp2 = zmulr(gzp2, z2a);
    const zc t1 = zmulr(zadd(zadd(zadd(gzp2, gzp1), gz2), gz1), ss);
    const zc t2 = zmulr(zmulr(zsub(zmulr(gz2, z2a), zmulr(gz1, z1a)), cs), xk);
    o->ers = zdivr(zmul(mk1, zsub(t1, t2)), rh);
    const zc u1 = zmulr(zsub(zadd(zsub(gzp2, gzp1), gz2), gz1), cs);
    const zc u2 = zmulr(zmulr(zadd(zmulr(gz2, z2a), zmulr(gz1, z1a)), ss), xk);
    o->erc = zdivr(zmul(mk1, zadd(u1, u2)), rh);
    return;
  }
  o->ers = zmk(0., 0.);
  o->erc = zmk(0., 0.);
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
// :2617-2619).
__device__ void pair_fields(const DevSegments& g, int i, int j, double rkh, int use_ek, int ind1, int ind2, zc* etk,
                            zc* ets, zc* etc) {
  const double cabj = g.cab[j], sabj = g.sab[j], salpj = g.salp[j];
  const double sj = g.s[j];
  const double xij = g.x[i] - g.x[j];
  const double yij = g.y[i] - g.y[j];
  const double zij = g.z[i] - g.z[j];
  const double ai = g.b[i];
  const double zp = xij * cabj + yij * sabj + zij * salpj;
  double rhox = xij - cabj * zp;
  double rhoy = yij - sabj * zp;
  double rhoz = zij - salpj * zp;
  const double rh = sqrt(rhox * rhox + rhoy * rhoy + rhoz * rhoz + ai * ai);
  if (rh <= 1.e-10) {
    rhox = 0.0;
    rhoy = 0.0;
    rhoz = 0.0;
  } else {
    rhox /= rh;
    rhoy /= rh;
    rhoz /= rh;
  }
  const double r = sqrt(zp * zp + rh * rh);
  zc tezk, terk, tezc, terc;
  zc e3 = zmk(0, 0), e4 = zmk(0, 0), e5 = zmk(0, 0);
  if (r >= rkh) {
    // lumped current element, src/nec_context.cpp:2798-2816
    const double eta = em_impedance();
    double rmag = kTwoPi * r;
    const double cth = zp / r;
    const double px = rh / r;
    double sr, cr;
    sincos(rmag, &sr, &cr);
    const zc e0 = zmk(cr, -sr);
    const double py = kTwoPi * r * r;
    const zc e1 = zdivr(zmul(zmulr(e0, eta * cth), zmk(1.0, -1.0 / rmag)), py);
    const zc e2 = zdivr(zmul(zmulr(e0, eta * px), zmk(1.0, rmag - 1.0 / rmag)), 2. * py);
    tezk = zsub(zmulr(e1, cth), zmulr(e2, px));
    terk = zadd(zmulr(e1, px), zmulr(e2, cth));
    rmag = sin(kPi * sj) / kPi;
    tezc = zmulr(tezk, rmag);
    terc = zmulr(terk, rmag);
    tezk = zmulr(tezk, sj);
    terk = zmulr(terk, sj);
  } else {
    NearOut o;
    if (!use_ek)
      eksc(sj, zp, rh, kTwoPi, i != j, &o);
    else
      ekscx(g.b[j], sj, zp, rh, kTwoPi, i != j, ind1, ind2, &o);
    tezk = o.ezk;
    terk = o.erk;
    tezc = o.ezc;
    terc = o.erc;
    e3 = zadd(zmulr(o.ezs, cabj), zmulr(o.ers, rhox));
    e4 = zadd(zmulr(o.ezs, sabj), zmulr(o.ers, rhoy));
    e5 = zadd(zmulr(o.ezs, salpj), zmulr(o.ers, rhoz));
  }
  const zc e0 = zadd(zmulr(tezk, cabj), zmulr(terk, rhox));
  const zc e1 = zadd(zmulr(tezk, sabj), zmulr(terk, rhoy));
  const zc e2 = zadd(zmulr(tezk, salpj), zmulr(terk, rhoz));
  const zc e6 = zadd(zmulr(tezc, cabj), zmulr(terc, rhox));
  const zc e7 = zadd(zmulr(tezc, sabj), zmulr(terc, rhoy));
  const zc e8 = zadd(zmulr(tezc, salpj), zmulr(terc, rhoz));
  const double cabi = g.cab[i], sabi = g.sab[i], salpi = g.salp[i];
  *etk = zadd(zadd(zmulr(e0, cabi), zmulr(e1, sabi)), zmulr(e2, salpi));
  *ets = zadd(zadd(zmulr(e3, cabi), zmulr(e4, sabi)), zmulr(e5, salpi));
  *etc = zadd(zadd(zmulr(e6, cabi), zmulr(e7, sabi)), zmulr(e8, salpi));
}

__global__ void __launch_bounds__(256) fill_fields_kernel(DevSegments g, int n, int i0, int rows, double rkh,
                                                          int use_ek, const int* ekind, double* __restrict__ E) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (j >= n || r >= rows) return;
  const int i = i0 + r;
  int ind1 = 2, ind2 = 2;
  if (use_ek) {
    ind1 = ekind[2 * j];
    ind2 = ekind[2 * j + 1];
  }
  zc etk, ets, etc;
  pair_fields(g, i, j, rkh, use_ek, ind1, ind2, &etk, &ets, &etc);
  double* o = E + ((size_t)r * n + j) * 6;
  o[0] = etk.re;
  o[1] = etk.im;
  o[2] = ets.re;
  o[3] = ets.im;
  o[4] = etc.re;
  o[5] = etc.im;
}

__global__ void __launch_bounds__(256) fill_assemble_kernel(const double* __restrict__ E, int n, int i0, int rows,
                                                            const int* __restrict__ colptr,
                                                            const int* __restrict__ colj,
                                                            const double* __restrict__ colabc,
                                                            double* __restrict__ Z, long long ldz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (k >= n || r >= rows) return;
  const double* Er = E + (size_t)r * n * 6;
  double accr = 0.0, acci = 0.0;
  const int q1 = colptr[k + 1];
  for (int q = colptr[k]; q < q1; ++q) {
    const double* e = Er + (size_t)colj[q] * 6;
    const double a = colabc[3 * q], b = colabc[3 * q + 1], c = colabc[3 * q + 2];
    // etk*a + ets*b + etc*c, then Z += term (src/nec_context.cpp:2635-2638)
    const double tr = ((e[0] * a) + (e[2] * b)) + (e[4] * c);
    const double ti = ((e[1] * a) + (e[3] * b)) + (e[5] * c);
    accr += tr;
    acci += ti;
  }
  double* zp = Z + 2 * ((size_t)(i0 + r) * ldz + k);
  zp[0] = accr;
  zp[1] = acci;
}

// Batched fill: one workgroup per (solve, row i).  The fields of every
// source j at i go to LDS, then each entry Z[i][k] gathers them over basis
// k's support in ascending j, as fill_assemble does.
__global__ void __launch_bounds__(256) batch_fill_kernel(const BatchSolve* __restrict__ sol, int S,
                                                         const double* __restrict__ seg,
                                                         const int* __restrict__ colptr,
                                                         const int* __restrict__ colj,
                                                         const double* __restrict__ colabc, double* __restrict__ Zall,
                                                         double rkh) {
  __shared__ double sE[kBatchMaxN * 6];
  const long long blk = blockIdx.x;
  int lo = 0, hi = S - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sol[mid].row0 <= blk) lo = mid;
    else hi = mid - 1;
  }
  const BatchSolve d = sol[lo];
  const int n = d.n;
  const int i = static_cast<int>(blk - d.row0);
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
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    zc etk, ets, etc;
    pair_fields(g, i, j, rkh, 0, 2, 2, &etk, &ets, &etc);
    double* o = sE + 6 * j;
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
  double* zrow = Zall + 2 * (d.z0 + (long long)i * n);
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    double accr = 0.0, acci = 0.0;
    for (int q = cp[k]; q < cp[k + 1]; ++q) {
      const double* e = sE + 6 * cj[q];
      const double a = cv[3 * q], b = cv[3 * q + 1], c = cv[3 * q + 2];
      accr += ((e[0] * a) + (e[2] * b)) + (e[4] * c);
      acci += ((e[1] * a) + (e[3] * b)) + (e[5] * c);
    }
    zrow[2 * k] = accr;
    zrow[2 * k + 1] = acci;
  }
}

}  // namespace

int launch_batch_fill(const BatchSolve* sol, int S, long long rows_total, const double* seg, const int* colptr,
                      const int* colj, const double* colabc, double* Z, double rkh, hipStream_t stream) {
  if (S <= 0 || rows_total <= 0) return 0;
  launch_k(K_BATCH_FILL, batch_fill_kernel, dim3((unsigned)rows_total), dim3(256), 0u, stream, sol, S, seg, colptr,
           colj, colabc, Z, rkh);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_fill(const DevSegments& g, int n, int i0, int rows, double rkh, int use_ek, const int* ekind, double* E,
                const int* colptr, const int* colj, const double* colabc, double* Z, long long ldz,
                hipStream_t stream) {
  if (n <= 0 || rows <= 0) return 0;
  dim3 block(256);
  dim3 grid((n + 255) / 256, rows);
  launch_k(K_FILL_FIELDS, fill_fields_kernel, grid, block, 0, stream, g, n, i0, rows, rkh, use_ek, ekind, E);
  launch_k(K_FILL_ASSEMBLE, fill_assemble_kernel, grid, block, 0, stream, E, n, i0, rows, colptr, colj, colabc, Z, ldz);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace necgpu
