// adjoint.cpp -- host side of the adjoint gradient.
//
// The solve is Z x = e (Z, x, e complex; the geometry in wavelengths).  The
// objective J depends on the segment current coefficients U_j = (A, B, C)_j
// = sum_k c_jk x_k (c_jk the real basis coefficients of basis k on segment
// j, functions of segment lengths and radii) and on the geometry through
// the far-field phase and shape factors.  With the Wirtinger derivatives
// dJ/dU_j, the adjoint right-hand side is g_k = sum_j c_jk . dJ/dU_j, the
// adjoint vector solves Z^T lam = g, and for every geometry variable v
//   dJ/dv = dJ/dv|_U + 2 Re( sum_j (dJ/dU_j - W_j) . dU_j/dv|_x )
//         + 2 Re( lam^T de/dv ) - 2 Re( sum_ij lam_i dE_ij/dv . U_j )
// where E_ij are the three field components of the fill (Z_ik = sum over
// the support of k of E_ij . c_jk) and W_j = sum_i lam_i E_ij.  The device
// does the transposed solve and the last sum together with W (the fields
// are differentiated inside the fill code, not stored); this file forms J,
// g, the far-field and basis-coefficient derivatives and the excitation
// term.  The objective follows the reference's gain and impedance
// definitions: nec_radiation_pattern::analyze (power gain, gcop,
// src/nec_radiation_pattern.cpp:221-262), nec_context::ffld
// (src/nec_context.cpp:6639-6887), the antenna input parameters of netwk
// (src/nec_context.cpp:4850-5047) and the applied-field voltage source of
// setup_excitation (src/nec_context.cpp:1561-1617).
#include "adjoint.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

#include "basis.h"
#include "dual.h"

namespace necgpu {

namespace {

constexpr int kCap = 64;

struct Direction {
  double rox, roy, roz, thx, thy, thz, phx, phy;
};

Direction direction(double theta_deg, double phi_deg) {
  const double th = deg_to_rad(theta_deg), ph = deg_to_rad(phi_deg);
  const double sth = std::sin(th), cth = std::cos(th), sph = std::sin(ph), cph = std::cos(ph);
  Direction d;
  d.phx = -sph;
  d.phy = cph;
  d.roz = cth;
  const double thz = -sth;
  d.rox = -thz * d.phy;
  d.roy = thz * d.phx;
  d.thx = d.roz * d.phy;
  d.thy = -d.roz * d.phx;
  d.thz = thz;
  return d;
}

// Far-field sum terms of one segment (ffld for free space,
// src/nec_context.cpp:6815-6870): the segment's contribution to the current
// moment (cx, cy, cz) in direction d, and its three shape factors.
template <class T>
void segment_far(const Direction& d, const T& x, const T& y, const T& z, const T& s, const T& cab, const T& sab,
                 const T& salp, zc A, zc B, zc C, cx<T>* ox, cx<T>* oy, cx<T>* oz, T* fa = nullptr, T* fb = nullptr,
                 T* fc = nullptr, T* arg_out = nullptr) {
  const T el = kPi * s;
  const T omega = -(d.rox * cab + d.roy * sab + d.roz * salp);
  const T sill = omega * el;
  const T top = el + sill;
  const T bot = el - sill;
  T a, too, boo;
  if (fabs(val(omega)) >= 1.0e-7) a = 2. * sin(sill) / omega;
  else a = (2. - omega * omega * el * el / 3.) * el;
  if (fabs(val(top)) >= 1.0e-7) too = sin(top) / top;
  else too = 1. - top * top / 6.;
  if (fabs(val(bot)) >= 1.0e-7) boo = sin(bot) / bot;
  else boo = 1. - bot * bot / 6.;
  const T b = el * (boo - too);
  const T c = el * (boo + too);
  const T rr = a * A.re + b * B.im + c * C.re;
  const T ri = a * A.im - b * B.re + c * C.im;
  const T arg = kTwoPi * (x * d.rox + y * d.roy + z * d.roz);
  T sa, ca;
  sincos(arg, &sa, &ca);
  const cx<T> exa = cmk<T>(ca * rr - sa * ri, ca * ri + sa * rr);
  *ox = cmulr(exa, cab);
  *oy = cmulr(exa, sab);
  *oz = cmulr(exa, salp);
  if (fa) {
    *fa = a;
    *fb = b;
    *fc = c;
    *arg_out = arg;
  }
}

zc cplx(double r, double i) { return zmk(r, i); }

// Segments whose end sets meet: union-find over the 2n segment ends with the
// links of icon1 / icon2 (icon1[i] = +k: end 1 of i meets end 2 of k,
// -k: end 1 of k; icon2[i] = +k: end 2 of i meets end 1 of k, -k: end 2).
struct EndGroups {
  std::vector<int> parent;
  std::vector<std::vector<int>> members;  // per root: segments with an end there
  int find(int a) {
    while (parent[a] != a) a = parent[a] = parent[parent[a]];
    return a;
  }
  void build(const std::vector<int32_t>& icon1, const std::vector<int32_t>& icon2, int n) {
    parent.resize(2 * (size_t)n);
    std::iota(parent.begin(), parent.end(), 0);
    auto unite = [&](int a, int b) {
      a = find(a);
      b = find(b);
      if (a != b) parent[a] = b;
    };
    for (int i = 0; i < n; ++i) {
      const int c1 = icon1[i], c2 = icon2[i];
      if (c1 != 0 && std::abs(c1) <= n && std::abs(c1) != i + 1) {
        const int k = std::abs(c1) - 1;
        unite(2 * i, c1 > 0 ? 2 * k + 1 : 2 * k);
      }
      if (c2 != 0 && std::abs(c2) <= n && std::abs(c2) != i + 1) {
        const int k = std::abs(c2) - 1;
        unite(2 * i + 1, c2 > 0 ? 2 * k : 2 * k + 1);
      }
    }
    members.assign(2 * (size_t)n, {});
    for (int e = 0; e < 2 * n; ++e) members[find(e)].push_back(e / 2);
  }
  const std::vector<int>& at(int seg, int end) { return members[find(2 * seg + end)]; }
};

}  // namespace

void adjoint_front(const ScaledSegments& seg, const BasisTables& t, int src, zc volt, double wavelength,
                   const double* abc, const AdjointSpec& spec, AdjointFront* f) {
  const int n = static_cast<int>(seg.x.size());
  if (src < 0 || src >= n) throw NecError("adjoint: the source segment is out of range");
  const double kdb = 10.0 / std::log(10.0);
  const Direction dr = direction(spec.theta_deg, spec.phi_deg);
  const zc const3 = zmk(0.0, -em_impedance() / kFourPi);
  auto U = [&](int j, int c) { return cplx(abc[(size_t)(2 * c) * n + j], abc[(size_t)(2 * c + 1) * n + j]); };

  // ---- objective: far field in the direction, input current and power
  zc Cx = zmk(0, 0), Cy = zmk(0, 0), Cz = zmk(0, 0);
  std::vector<double> fa(n), fb(n), fc(n), farg(n);
  for (int j = 0; j < n; ++j) {
    cx<double> ox, oy, oz;
    segment_far<double>(dr, seg.x[j], seg.y[j], seg.z[j], seg.s[j], seg.cab[j], seg.sab[j], seg.salp[j], U(j, 0),
                        U(j, 1), U(j, 2), &ox, &oy, &oz, &fa[j], &fb[j], &fc[j], &farg[j]);
    Cx = zadd(Cx, zmk(ox.re, ox.im));
    Cy = zadd(Cy, zmk(oy.re, oy.im));
    Cz = zadd(Cz, zmk(oz.re, oz.im));
  }
  f->eth = zmul(zadd(zadd(zmulr(Cx, dr.thx), zmulr(Cy, dr.thy)), zmulr(Cz, dr.thz)), const3);
  f->eph = zmul(zadd(zmulr(Cx, dr.phx), zmulr(Cy, dr.phy)), const3);
  const zc eth = f->eth, eph = f->eph;
  const double S = (eth.re * eth.re + eth.im * eth.im) + (eph.re * eph.re + eph.im * eph.im);
  f->usrc = zadd(U(src, 0), U(src, 2));  // current at the segment centre: A + C
  const zc I = zmulr(f->usrc, wavelength);
  const zc zin = zdiv(volt, I);
  const double P = 0.5 * zmul(volt, zconj(I)).re;
  if (!(S > 0.0) || !(P > 0.0)) throw NecError("adjoint: the gain is undefined (no radiated field or input power)");
  const double gcon = wavelength * wavelength * 2.0 * kPi / (em_impedance() * P);
  const double G = 10.0 * std::log10(gcon * S);
  const zc dz = zsub(zin, spec.z0);
  f->gain_db = G;
  f->zin = zin;
  f->J = -G + spec.weight * (dz.re * dz.re + dz.im * dz.im);

  // ---- Wirtinger derivatives dJ/dU_j (3 complex per segment)
  std::vector<zc>& dU = f->dU;
  dU.assign(3 * (size_t)n, zmk(0.0, 0.0));
  const zc ceth = zconj(eth), ceph = zconj(eph);
  f->dJdS = -kdb / S;
  for (int j = 0; j < n; ++j) {
    const zc ph = zmk(std::cos(farg[j]), std::sin(farg[j]));
    const zc pth = zmul(const3, zmulr(ph, (seg.cab[j] * dr.thx + seg.sab[j] * dr.thy) + seg.salp[j] * dr.thz));
    const zc pph = zmul(const3, zmulr(ph, seg.cab[j] * dr.phx + seg.sab[j] * dr.phy));
    const zc Q = zmulr(zadd(zmul(ceth, pth), zmul(ceph, pph)), f->dJdS);
    dU[3 * j + 0] = zmulr(Q, fa[j]);
    dU[3 * j + 1] = zmul(Q, zmk(0.0, -fb[j]));
    dU[3 * j + 2] = zmulr(Q, fc[j]);
  }
  // Z_in = V / I and P = Re(V conj(I)) / 2, I = lambda (A + C)_src
  f->dJdI = zadd(zmulr(zmul(zconj(dz), zneg(zdiv(volt, zmul(I, I)))), spec.weight),
                 zmulr(zconj(volt), 0.25 * kdb / P));
  dU[3 * src + 0] = zadd(dU[3 * src + 0], zmulr(f->dJdI, wavelength));
  dU[3 * src + 2] = zadd(dU[3 * src + 2], zmulr(f->dJdI, wavelength));

  // ---- adjoint right-hand side through the current coefficients
  f->g.assign(n, zmk(0.0, 0.0));
  for (int j = 0; j < n; ++j)
    for (int q = t.segptr[j]; q < t.segptr[j + 1]; ++q) {
      const int k = t.segbasis[q];
      const double* c = &t.segabc[3 * (size_t)q];
      f->g[k] = zadd(f->g[k],
                     zadd(zadd(zmulr(dU[3 * j], c[0]), zmulr(dU[3 * j + 1], c[1])), zmulr(dU[3 * j + 2], c[2])));
    }
}

void adjoint_back(const Geometry& geo, const ScaledSegments& seg, int src, zc volt, double wavelength,
                  const double* abc, const AdjointSpec& spec, const AdjointFront& f, const zc* lam, const zc* x,
                  const double* gi, const double* gj, AdjointResult* out) {
  const int n = static_cast<int>(seg.x.size());
  const double kdb = 10.0 / std::log(10.0);
  const Direction dr = direction(spec.theta_deg, spec.phi_deg);
  const zc const3 = zmk(0.0, -em_impedance() / kFourPi);
  auto U = [&](int j, int c) { return cplx(abc[(size_t)(2 * c) * n + j], abc[(size_t)(2 * c + 1) * n + j]); };
  const zc ceth = zconj(f.eth), ceph = zconj(f.eph);
  out->J = f.J;
  out->gain_db = f.gain_db;
  out->zin = f.zin;
  std::vector<double>& D = out->dseg;
  D.assign(8 * (size_t)n, 0.0);
  for (int m = 0; m < n; ++m) {
    double* d = &D[8 * (size_t)m];
    const double* a = &gi[7 * (size_t)m];
    const double* b = &gj[13 * (size_t)m];
    d[0] += a[0] + b[0];
    d[1] += a[1] + b[1];
    d[2] += a[2] + b[2];
    d[3] += b[3];
    d[4] += a[3];
    d[5] += a[4] + b[4];
    d[6] += a[5] + b[5];
    d[7] += a[6] + b[6];
  }

  // ---- far-field shape and phase factors at fixed currents
  typedef Dn<7> D7;
  for (int j = 0; j < n; ++j) {
    D7 v[7] = {D7(seg.x[j]), D7(seg.y[j]), D7(seg.z[j]), D7(seg.s[j]), D7(seg.cab[j]), D7(seg.sab[j]),
               D7(seg.salp[j])};
    for (int q = 0; q < 7; ++q) v[q].d[q] = 1.0;
    cx<D7> ox, oy, oz;
    segment_far<D7>(dr, v[0], v[1], v[2], v[3], v[4], v[5], v[6], U(j, 0), U(j, 1), U(j, 2), &ox, &oy, &oz);
    static const int slot[7] = {0, 1, 2, 3, 5, 6, 7};
    for (int q = 0; q < 7; ++q) {
      const zc dcx = zmk(ox.re.d[q], ox.im.d[q]), dcy = zmk(oy.re.d[q], oy.im.d[q]), dcz = zmk(oz.re.d[q], oz.im.d[q]);
      const zc deth = zmul(zadd(zadd(zmulr(dcx, dr.thx), zmulr(dcy, dr.thy)), zmulr(dcz, dr.thz)), const3);
      const zc deph = zmul(zadd(zmulr(dcx, dr.phx), zmulr(dcy, dr.phy)), const3);
      const double dS = 2.0 * (zmul(ceth, deth).re + zmul(ceph, deph).re);
      D[8 * (size_t)j + slot[q]] += f.dJdS * dS;
    }
  }

  // ---- excitation: e_src = -V / (s_src lambda)
  const zc esrc = zdivr(zneg(volt), seg.s[src] * wavelength);
  D[8 * (size_t)src + 3] += 2.0 * zmul(lam[src], zdivr(zneg(esrc), seg.s[src])).re;

  // ---- basis coefficients (functions of lengths and radii) at fixed x:
  // 2 Re(sum_j H_j . dU_j/dv), H_j = dJ/dU_j - W_j
  {
    std::vector<int> icon1(geo.icon1.begin(), geo.icon1.end()), icon2(geo.icon2.begin(), geo.icon2.end());
    EndGroups groups;
    groups.build(geo.icon1, geo.icon2, n);
    typedef Dn<2> D2;
    std::vector<D2> sD(n), bD(n);
    for (int m = 0; m < n; ++m) {
      sD[m] = D2(seg.s[m]);
      bD[m] = D2(seg.b[m]);
    }
    const SegView sv{icon1.data(), icon2.data(), seg.s.data(), seg.b.data()};
    const SegViewT<D2> svd{icon1.data(), icon2.data(), sD.data(), bD.data()};
    int jco[kCap];
    double ax[kCap], bx[kCap], cxx[kCap];
    D2 axd[kCap], bxd[kCap], cxd[kCap];
    std::vector<int> cand;
    for (int j = 0; j < n; ++j) {
      const int cnt = trio(sv, j + 1, jco, ax, bx, cxx, kCap);
      if (cnt < 0) throw NecError("adjoint: segment connection error");
      zc H[3];
      for (int c = 0; c < 3; ++c) {
        const double* w = &gj[13 * (size_t)j + 7 + 2 * c];
        H[c] = zsub(f.dU[3 * j + c], zmk(w[0], w[1]));
      }
      // segments read by trio(j): each basis on j, and every segment with an
      // end at either end of that basis' centre segment
      cand.clear();
      for (int q = 0; q < cnt; ++q) {
        const int k = jco[q] - 1;
        cand.push_back(k);
        for (int e = 0; e < 2; ++e)
          for (int m : groups.at(k, e)) cand.push_back(m);
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (int m : cand) {
        sD[m].d[0] = 1.0;
        bD[m].d[1] = 1.0;
        const int c2 = trio(svd, j + 1, jco, axd, bxd, cxd, kCap);
        sD[m].d[0] = 0.0;
        bD[m].d[1] = 0.0;
        if (c2 != cnt) throw NecError("adjoint: basis table mismatch");
        for (int v = 0; v < 2; ++v) {
          zc du[3] = {zmk(0, 0), zmk(0, 0), zmk(0, 0)};
          for (int q = 0; q < cnt; ++q) {
            const zc xk = x[jco[q] - 1];
            du[0] = zadd(du[0], zmulr(xk, axd[q].d[v]));
            du[1] = zadd(du[1], zmulr(xk, bxd[q].d[v]));
            du[2] = zadd(du[2], zmulr(xk, cxd[q].d[v]));
          }
          const double contrib =
              2.0 * ((zmul(H[0], du[0]).re + zmul(H[1], du[1]).re) + zmul(H[2], du[2]).re);
          D[8 * (size_t)m + (v == 0 ? 3 : 4)] += contrib;
        }
      }
    }
  }

  // ---- explicit wavelength dependence (for the frequency derivative):
  // I = lambda (A + C)_src, gcon ~ lambda^2, e_src ~ 1 / lambda
  out->dlambda = 2.0 * zmul(f.dJdI, f.usrc).re - kdb * 2.0 / wavelength +
                 2.0 * zmul(lam[src], zdivr(zneg(esrc), wavelength)).re;
}

void adjoint_gradient(const Geometry& geo, const ScaledSegments& seg, int src, zc volt, double wavelength,
                      double rkh, const std::vector<double>& abc, DeviceSolver& dev, const AdjointSpec& spec,
                      AdjointResult* out) {
  const auto t0 = std::chrono::steady_clock::now();
  const int n = static_cast<int>(geo.size());
  if ((int)abc.size() != 6 * n) throw NecError("adjoint: no current coefficients of the last solve");
  BasisTables t;
  build_basis_tables(geo, seg, &t);
  AdjointFront f;
  adjoint_front(seg, t, src, volt, wavelength, abc.data(), spec, &f);
  const auto t1 = std::chrono::steady_clock::now();
  std::vector<zc> lam, x;
  std::vector<double> gi, gj;
  dev.adjoint(rkh, f.g, &lam, &x, &gi, &gj);
  out->device_ms = dev.adjoint_ms();
  const auto t2 = std::chrono::steady_clock::now();
  adjoint_back(geo, seg, src, volt, wavelength, abc.data(), spec, f, lam.data(), x.data(), gi.data(), gj.data(), out);
  const auto t3 = std::chrono::steady_clock::now();
  out->host_ms = std::chrono::duration<double, std::milli>((t1 - t0) + (t3 - t2)).count();
}

void adjoint_endpoints(const ScaledSegments& seg, double wavelength, double freq_mhz, const AdjointResult& r,
                       std::vector<double>* grad7, double* dfreq_mhz) {
  const size_t n = seg.x.size();
  grad7->assign(7 * n, 0.0);
  double scale_sum = 0.0;
  for (size_t m = 0; m < n; ++m) {
    const double* d = &r.dseg[8 * m];
    const double dir[3] = {seg.cab[m], seg.sab[m], seg.salp[m]};
    const double gd[3] = {d[5], d[6], d[7]};
    const double dot = dir[0] * gd[0] + dir[1] * gd[1] + dir[2] * gd[2];
    double* o = &(*grad7)[7 * m];
    for (int c = 0; c < 3; ++c) {
      // centre = (p1 + p2) / 2, length = |p2 - p1|, direction = (p2 - p1) / length
      const double perp = (gd[c] - dir[c] * dot) / seg.s[m];
      o[c] = (0.5 * d[c] - d[3] * dir[c] - perp) / wavelength;
      o[3 + c] = (0.5 * d[c] + d[3] * dir[c] + perp) / wavelength;
    }
    o[6] = d[4] / wavelength;
    scale_sum += (seg.x[m] * d[0] + seg.y[m] * d[1] + seg.z[m] * d[2]) + seg.s[m] * d[3] + seg.b[m] * d[4];
  }
  // geometry in wavelengths scales with f; lambda = c / f
  *dfreq_mhz = (scale_sum - wavelength * r.dlambda) / freq_mhz;
}

}  // namespace necgpu
