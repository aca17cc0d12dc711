// context.cpp -- card state machine of the nec_context behind the C API.
#include "context.h"

#include <cmath>
#include <complex>

namespace necgpu {

Context::Context() { calc_prepare(); }
Context::~Context() = default;

DeviceSolver& Context::device() {
  if (!dev_) dev_ = std::make_unique<DeviceSolver>();
  return *dev_;
}

void Context::calc_prepare() {
  // src/nec_context.cpp:123-164
  iflow_ = 1;
  state_ = MEMORY_ALLOC;
  rkh_ = 1.;
  use_ek_ = false;
  gn_iperf_ = -1;  // ground.default_values(): free space
  loads_.clear();
  networks_.clear();
  ifar_ = -1;
  freq_mhz_ = em_speed_of_light() / 1.0e6;
  nfrq_ = 1;
}

void Context::reset_to_structure_loading() {
  if (state_ != MEMORY_ALLOC) state_ = STRUCTURE_LOADING;
}

void Context::reset_to_excitation_setup() {
  if (state_ != MEMORY_ALLOC && state_ != STRUCTURE_LOADING) state_ = EXCITATION_SETUP;
}

void Context::wire(int tag, int ns, double x1, double y1, double z1, double x2, double y2, double z2, double rad,
                   double rdel, double rrad) {
  geo_.add_wire(tag, ns, x1, y1, z1, x2, y2, z2, rad, rdel, rrad);
}

void Context::move(double rox, double roy, double roz, double xs, double ys, double zs, int its, int nrpt,
                   int itgi) {
  geo_.move(rox, roy, roz, xs, ys, zs, its, nrpt, itgi);
}

void Context::scale(double f) { geo_.scale(f); }

void Context::geometry_complete(int gpflag) {
  geo_.complete(gpflag);
  calc_prepare();
}

void Context::fr_card(int ifrq, int nfrq, double freq_mhz, double del_freq) {
  // src/nec_context.cpp:406-420
  ifrq_ = ifrq;
  nfrq_ = (nfrq == 0) ? 1 : nfrq;
  freq_mhz_ = freq_mhz;
  delfrq_ = del_freq;
  state_ = MEMORY_ALLOC;
  iflow_ = 1;
}

void Context::ex_card(int type, int i2, int i3, int i4, double f1, double f2, double f3, double f4, double f5,
                      double f6) {
  // src/nec_context.cpp:519-612 (voltage sources only)
  (void)i4;
  (void)f3;
  (void)f4;
  (void)f5;
  (void)f6;
  if (type != 0) throw NecError("EX: excitation type " + std::to_string(type) + " not supported by the GPU path");
  if (!geo_.is_complete()) throw NecError("EX: geometry is not complete");
  if (iflow_ != 5) {
    src_seg_.clear();
    src_v_.clear();
    iflow_ = 5;
    reset_to_excitation_setup();
  }
  const int64_t seg = geo_.segment_number(i2, i3);
  if (seg > geo_.size())
    throw NecError("CHECK DATA, PARAMETER SPECIFYING EXCITATION SOURCE SEGMENT [" + std::to_string(seg) +
                   "] IS TOO LARGE");
  zc v = zmk(f1, f2);
  if (std::hypot(v.re, v.im) < 1.e-20) v = zmk(1.0, 0.0);
  src_seg_.push_back(seg);
  src_v_.push_back(v);
}

void Context::network_card(int type, int i1, int i2, int i3, int i4, const double* x) {
  // tl_card / nt_card, src/nec_context.cpp:630-726: consecutive TL / NT
  // cards accumulate; the first after any other card starts a new list, and
  // with a port-1 segment of -1 only clears it
  if (!geo_.is_complete()) throw NecError(std::string(type == 1 ? "NT" : "TL") + ": geometry is not complete");
  if (iflow_ != 6) {
    networks_.clear();
    iflow_ = 6;
    reset_to_excitation_setup();
    if (i2 == -1) return;
  }
  Network net;
  net.type = type;
  net.seg1 = geo_.segment_number(i1, i2);
  net.seg2 = geo_.segment_number(i3, i4);
  if (net.seg1 < 1 || net.seg1 > geo_.size() || net.seg2 < 1 || net.seg2 > geo_.size())
    throw NecError("NETWORK DATA ERROR: PORT SEGMENT OUT OF RANGE");
  for (int q = 0; q < 6; ++q) net.x[q] = x[q];
  if (type == 2 && x[0] <= 0.0) {
    // a negative characteristic impedance: the crossed line
    net.type = 3;
    net.x[0] = -x[0];
  }
  networks_.push_back(net);
}

void Context::tl_card(int i1, int i2, int i3, int i4, double z0, double len, double y1r, double y1i, double y2r,
                      double y2i) {
  const double x[6] = {z0, len, y1r, y1i, y2r, y2i};
  network_card(2, i1, i2, i3, i4, x);
}

void Context::nt_card(int i1, int i2, int i3, int i4, double y11r, double y11i, double y12r, double y12i,
                      double y22r, double y22i) {
  const double x[6] = {y11r, y11i, y12r, y12i, y22r, y22i};
  network_card(1, i1, i2, i3, i4, x);
}

void Context::kh_card(double rkh) {
  rkh_ = rkh;
  reset_to_structure_loading();
  iflow_ = 1;
}

void Context::gn_card(int iperf, int nradl, double epsr, double sig, double t3, double t4, double t5, double t6) {
  // nec_ground::parse_gn, src/nec_ground.cpp:75-121
  (void)t3;
  (void)t4;
  (void)t5;
  (void)t6;  // cliff parameters: used by RP modes 2-5 only, which are not supported
  if (iperf == -1) {
    gn_iperf_ = -1;
  } else {
    if (iperf == 2) throw NecError("GN: the Sommerfeld ground (type 2) is not supported by the GPU path");
    if (iperf != 0 && iperf != 1) throw NecError("GN: ground type " + std::to_string(iperf) + " is not supported");
    if (nradl != 0) throw NecError("GN: the radial-wire ground screen is not supported by the GPU path");
    if (iperf == 0) {
      if (epsr == 0.0) throw NecError("GROUND DIELECTRIC CONSTANT IS ZERO");
      if (sig == 0.0) throw NecError("GROUND CONDUCTIVITY IS ZERO");
    }
    gn_iperf_ = iperf;
    gn_epsr_ = epsr;
    gn_sig_ = sig;
  }
  iflow_ = 4;
  reset_to_structure_loading();
}

void Context::ld_card(int type, int tag, int tagf, int tagt, double zlr, double zli, double zlc) {
  // src/nec_context.cpp:424-474
  if (iflow_ != 3) {
    iflow_ = 3;
    loads_.clear();
    reset_to_structure_loading();
    if (type == -1) return;
  }
  if (tagt == 0) tagt = tagf;
  if (tagt < tagf)
    throw NecError("DATA FAULT ON LOADING CARD No: " + std::to_string(loads_.size() + 1) + ": ITAG STEP1: " +
                   std::to_string(tagf) + " IS GREATER THAN ITAG STEP2: " + std::to_string(tagt));
  loads_.push_back(LoadCard{type, tag, tagf, tagt, zlr, zli, zlc});
}

namespace {

// nec_context::zint, src/nec_context.cpp:6321-6406: lambda * internal
// impedance per unit length of a round wire (sigl = conductivity x
// wavelength, rolam = radius in wavelengths), from the published polynomial
// approximations of the Kelvin functions (x <= 8), their asymptotic
// expansions (x > 8) and the skin-effect limit (x > 110).  The coefficient
// tables are the reference's, so that the loads agree to the last bit.
std::complex<double> zint(double sigl, double rolam) {
  typedef std::complex<double> C;
  const C cc1(6.0e-7, 1.9e-6), cc2(-3.4e-6, 5.1e-6), cc3(-2.52e-5, 0.0), cc4(-9.06e-5, -9.01e-5);
  const C cc5(0., -9.765e-4), cc6(.0110486, -.0110485), cc7(0., -.3926991), cc8(1.6e-6, -3.2e-6);
  const C cc9(1.17e-5, -2.4e-6), cc10(3.46e-5, 3.38e-5), cc11(5.0e-7, 2.452e-4), cc12(-1.3813e-3, 1.3811e-3);
  const C cc13(-6.25001e-2, -1.0e-7), cc14(.7071068, .7071068), cn = cc14;
  const double tpcmu = 2.368705e+3, cmotp = 60.00;
  const C j1(0.0, 1.0);
  const double x = std::sqrt(tpcmu * sigl) * rolam;
  auto th = [&](double d) { return (((((cc1 * d + cc2) * d + cc3) * d + cc4) * d + cc5) * d + cc6) * d + cc7; };
  auto ph = [&](double d) { return (((((cc8 * d + cc9) * d + cc10) * d + cc11) * d + cc12) * d + cc13) * d + cc14; };
  if (x > 110.0) {
    const C br1(0.70710678, -0.70710678);
    return j1 * std::sqrt(cmotp / sigl) * br1 / rolam;
  }
  if (x > 8.0) {
    const C fx = std::sqrt((kPi / 2.0) / x) * std::exp(-cn * x + th(-8. / x));
    const C gx = std::exp(cn * x + th(8. / x)) / std::sqrt((2.0 * kPi) * x);
    C br2 = j1 * fx / kPi;
    C br1 = gx + br2;
    br2 = gx * ph(8.0 / x) - br2 * ph(-8.0 / x);
    br1 = br1 / br2;
    return j1 * std::sqrt(cmotp / sigl) * br1 / rolam;
  }
  const double x8 = x / 8.0, y = x8 * x8, s = y * y;
  double ber =
      ((((((-9.01e-6 * s + 1.22552e-3) * s - .08349609) * s + 2.6419140) * s - 32.363456) * s + 113.77778) * s - 64.) *
          s +
      1.;
  double bei =
      ((((((1.1346e-4 * s - .01103667) * s + .52185615) * s - 10.567658) * s + 72.817777) * s - 113.77778) * s + 16.) *
      y;
  C br1(ber, bei);
  ber = (((((((-3.94e-6 * s + 4.5957e-4) * s - 0.02609253) * s + 0.66047849) * s - 6.0681481) * s + 14.222222) * s -
          4.0) *
         y) *
        x;
  bei = ((((((4.609e-5 * s - 3.79386e-3) * s + .14677204) * s - 2.3116751) * s + 11.377778) * s - 10.666667) * s +
         .5) *
        x;
  const C br2(ber, bei);
  br1 = br1 / br2;
  return j1 * std::sqrt(cmotp / sigl) * br1 / rolam;
}

}  // namespace

std::vector<zc> Context::load_impedances(const ScaledSegments& seg) const {
  // nec_context::load, src/nec_context.cpp:1861-2055 (no symmetry)
  std::vector<zc> out;
  if (loads_.empty()) return out;
  const int64_t n = geo_.size();
  std::vector<std::complex<double>> za(n, std::complex<double>(0.0, 0.0));
  const std::complex<double> tpcj(0.0, 1.883698955e+9);
  const double wl = wavelength_;
  for (size_t k = 0; k < loads_.size(); ++k) {
    const LoadCard& L = loads_[k];
    if (L.type > 5 || L.type < 0) throw NecError("IMPROPER LOAD TYPE CHOSEN, REQUESTED TYPE IS " + std::to_string(L.type));
    int ichk = 0;
    int64_t l1 = 1, l2 = n;
    if (L.tag == 0 && (L.tagf != 0 || L.tagt != 0)) {
      l1 = L.tagf;
      l2 = L.tagt;
    }
    for (int64_t i = l1 - 1; i < l2; ++i) {
      if (i < 0 || i >= n) continue;
      if (L.tag != 0) {
        if (L.tag != geo_.tags[i]) continue;
        if (L.tagf != 0) {
          ++ichk;
          if (ichk < L.tagf || ichk > L.tagt) continue;
        } else {
          ichk = 1;
        }
      } else {
        ichk = 1;
      }
      const double sl = seg.s[i];
      std::complex<double> zt;
      switch (L.type + 1) {
        case 1:
          zt = L.zlr / sl + tpcj * L.zli / (sl * wl);
          if (std::fabs(L.zlc) > 1.0e-20) zt += wl / (tpcj * sl * L.zlc);
          break;
        case 2:
          zt = tpcj * sl * L.zlc / wl;
          if (std::fabs(L.zli) > 1.0e-20) zt += sl * wl / (tpcj * L.zli);
          if (std::fabs(L.zlr) > 1.0e-20) zt += sl / L.zlr;
          zt = 1. / zt;
          break;
        case 3:
          zt = L.zlr * wl + tpcj * L.zli;
          if (std::fabs(L.zlc) > 1.0e-20) zt += 1. / (tpcj * sl * sl * L.zlc);
          break;
        case 4:
          zt = tpcj * sl * sl * L.zlc;
          if (std::fabs(L.zli) > 1.0e-20) zt += 1. / (tpcj * L.zli);
          if (std::fabs(L.zlr) > 1.0e-20) zt += 1. / (L.zlr * wl);
          zt = 1. / zt;
          break;
        case 5:
          zt = std::complex<double>(L.zlr, L.zli) / sl;
          break;
        default:
          zt = zint(L.zlr * wl, seg.b[i]);
          break;
      }
      za[i] += zt;
    }
    if (ichk == 0) throw NecError("LOADING DATA CARD ERROR, NO SEGMENT HAS AN ITAG = " + std::to_string(L.tag));
  }
  out.resize(n);
  for (int64_t i = 0; i < n; ++i) out[i] = zmk(za[i].real(), za[i].imag());
  return out;
}

GroundSpec Context::ground_spec() {
  // nec_ground::calculate_antenna_environment, src/nec_ground.cpp:176-197:
  // frati = 1; the finite ground's zrati = 1 / sqrt(epsc), epsc = epsr -
  // j sig lambda eta / (2 pi); a negative sig is converted once and kept
  GroundSpec g;
  if (gn_iperf_ == -1) return g;
  g.type = (gn_iperf_ == 1) ? 1 : 2;
  if (g.type == 2) {
    const double eta2pi = em_impedance() / (2.0 * kPi);
    if (gn_sig_ < 0.) gn_sig_ = -gn_sig_ / (eta2pi * wavelength_);
    const std::complex<double> epsc(gn_epsr_, -gn_sig_ * wavelength_ * eta2pi);
    const std::complex<double> zr = 1.0 / std::sqrt(epsc);
    g.zr_re = zr.real();
    g.zr_im = zr.imag();
  }
  return g;
}

void Context::ek_card(bool on) {
  use_ek_ = on;
  reset_to_structure_loading();
  iflow_ = 1;
}

void Context::xq_card(int itmp1) {
  // src/nec_context.cpp:741-779
  if ((iflow_ == 10 && itmp1 == 0) || (nfrq_ == 1 && itmp1 == 0 && iflow_ > 7)) return;
  if (itmp1 == 0) {
    iflow_ = (iflow_ > 7) ? 11 : 7;
    last_was_pattern_ = false;
  } else {
    ifar_ = 0;
    rfld_ = 0.;
    ipd_ = 0;
    iavp_ = 0;
    rp_norm_ = 0;
    rp_format_ = 0;
    nth_ = 91;
    nph_ = 1;
    thets_ = 0.0;
    phis_ = 0.0;
    dth_ = 1.0;
    dph_ = 0.0;
    if (itmp1 == 2) phis_ = 90.0;
    if (itmp1 == 3) {
      nph_ = 2;
      dph_ = 90.0;
    }
  }
  simulate(true);
}

void Context::rp_card(int calc_mode, int n_theta, int n_phi, int output_format, int normalization, int D, int A,
                      double theta0, double phi0, double delta_theta, double delta_phi, double radial_distance,
                      double gain_norm) {
  // src/nec_context.cpp:828-868
  if (calc_mode != 0)
    throw NecError("RP: calculation mode " + std::to_string(calc_mode) + " not supported by the GPU path");
  ifar_ = calc_mode;
  nth_ = (n_theta == 0) ? 1 : n_theta;
  nph_ = (n_phi == 0) ? 1 : n_phi;
  rp_format_ = (output_format != 0) ? 1 : 0;
  rp_norm_ = normalization;
  ipd_ = (D != 0) ? 1 : 0;
  iavp_ = A;
  if (nth_ < 2 || nph_ < 2 || ifar_ == 1) iavp_ = 0;
  if (rp_norm_ < 0 || rp_norm_ > 5) throw NecError("Unknown Gain Normalization Encountered.");
  thets_ = theta0;
  phis_ = phi0;
  dth_ = delta_theta;
  dph_ = delta_phi;
  rfld_ = radial_distance;
  gnor_ = gain_norm;
  iflow_ = 10;
  simulate(true);
}

void Context::adjoint(const AdjointSpec& spec, AdjointResult* r, std::vector<double>* grad7, double* dfreq_mhz) {
  if (inputs_.empty() || state_ < EXCITATION_LOOP) throw NecError("adjoint: no solve to differentiate (XQ or RP first)");
  if (src_seg_.size() != 1) throw NecError("adjoint: the objective needs exactly one voltage source");
  if (precision_ != 0) throw NecError("adjoint: needs the FP64 path (precision mode 0)");
  if (ground_now_.type != 0) throw NecError("adjoint: free space only (no GN card)");
  if (!zload_.empty()) throw NecError("adjoint: unloaded structures only (no LD card)");
  const ScaledSegments seg = geo_.scaled(freq_mhz_);
  adjoint_gradient(geo_, seg, static_cast<int>(src_seg_[0] - 1), src_v_[0], wavelength_, rkh_, abc_, device(), spec,
                   r);
  adjoint_endpoints(seg, wavelength_, freq_mhz_, *r, grad7, dfreq_mhz);
}

void Context::ne_card(int near, int nrx, int nry, int nrz, double xnr, double ynr, double znr, double dxnr,
                      double dynr, double dznr) {
  near_card(0, near, nrx, nry, nrz, xnr, ynr, znr, dxnr, dynr, dznr);
}

void Context::nh_card(int near, int nrx, int nry, int nrz, double xnr, double ynr, double znr, double dxnr,
                      double dynr, double dznr) {
  near_card(1, near, nrx, nry, nrz, xnr, ynr, znr, dxnr, dynr, dznr);
}

void Context::near_card(int nfeh, int near, int nrx, int nry, int nrz, double xnr, double ynr, double znr,
                        double dxnr, double dynr, double dznr) {
  // ne_nh_card, src/nec_context.cpp:930-955: the last NE or NH card is the
  // pending request; blank counts default to 1; with one frequency the
  // field is computed at once
  near_spec_.nfeh = nfeh;
  near_spec_.near = near;
  near_spec_.nrx = nrx == 0 ? 1 : nrx;
  near_spec_.nry = nry == 0 ? 1 : nry;
  near_spec_.nrz = nrz == 0 ? 1 : nrz;
  near_spec_.xnr = xnr;
  near_spec_.ynr = ynr;
  near_spec_.znr = znr;
  near_spec_.dxnr = dxnr;
  near_spec_.dynr = dynr;
  near_spec_.dznr = dznr;
  near_pending_ = true;
  iflow_ = 8;
  if (nfrq_ == 1) simulate(false);
}

void Context::compute_near_field() {
  // nfpat, src/nec_context.cpp:5068-5188: the points accumulate as the
  // reference steps them (z or theta outermost, x or r innermost)
  const NearSpec& ns = near_spec_;
  NearFieldResult r;
  r.freq_hz = freq_mhz_ / 1.e-6;
  double cth = 0., sth = 0., cph = 0., sph = 0.;
  double znrt = ns.znr - ns.dznr;
  for (int i = 0; i < ns.nrz; ++i) {
    znrt += ns.dznr;
    if (ns.near != 0) {
      cth = std::cos(deg_to_rad(znrt));
      sth = std::sin(deg_to_rad(znrt));
    }
    double ynrt = ns.ynr - ns.dynr;
    for (int j = 0; j < ns.nry; ++j) {
      ynrt += ns.dynr;
      if (ns.near != 0) {
        cph = std::cos(deg_to_rad(ynrt));
        sph = std::sin(deg_to_rad(ynrt));
      }
      double xnrt = ns.xnr - ns.dxnr;
      for (int k = 0; k < ns.nrx; ++k) {
        xnrt += ns.dxnr;
        if (ns.near != 0) {
          r.xyz.push_back(xnrt * sth * cph);
          r.xyz.push_back(xnrt * sth * sph);
          r.xyz.push_back(xnrt * cth);
        } else {
          r.xyz.push_back(xnrt);
          r.xyz.push_back(ynrt);
          r.xyz.push_back(znrt);
        }
      }
    }
  }
  // points in wavelengths, and the radius of the segment each lies on
  // (nefld, src/nec_context.cpp:4315-4331)
  const ScaledSegments seg = geo_.scaled(freq_mhz_);
  const size_t np = r.xyz.size() / 3;
  std::vector<double> pts(r.xyz.size()), ax(np, 0.0);
  for (size_t p = 0; p < np; ++p) {
    for (int c = 0; c < 3; ++c) pts[3 * p + c] = r.xyz[3 * p + c] / wavelength_;
    for (size_t i = 0; i < seg.x.size(); ++i) {
      const double xj = pts[3 * p] - seg.x[i], yj = pts[3 * p + 1] - seg.y[i], zj = pts[3 * p + 2] - seg.z[i];
      double zp = seg.cab[i] * xj + seg.sab[i] * yj + seg.salp[i] * zj;
      if (std::fabs(zp) > 0.5001 * seg.s[i]) continue;
      zp = xj * xj + yj * yj + zj * zj - zp * zp;
      if (zp > 0.9 * seg.b[i] * seg.b[i]) continue;
      ax[p] = seg.b[i];
      break;
    }
  }
  r.magnetic = ns.nfeh == 1;
  r.e = device().near_field(r.magnetic, pts, ax, rkh_);
  near_fields_.push_back(std::move(r));
}

void Context::simulate(bool far_field_flag) {
  // src/nec_context.cpp:1031-1159 for voltage excitation.
  if (!geo_.is_complete()) throw NecError("geometry is not complete");
  State igox = (far_field_flag && state_ == NEAR_FIELD) ? FAR_FIELD : state_;
  int mhz = 1;
  do {
    if (igox <= MEMORY_ALLOC) {
      if (mhz != 1) {
        if (ifrq_ == 1) freq_mhz_ *= delfrq_;
        else freq_mhz_ += delfrq_;
      }
      wavelength_ = em_speed_of_light() / (1.0e6 * freq_mhz_);
      device().load(geo_, geo_.scaled(freq_mhz_), use_ek_);
      state_ = STRUCTURE_LOADING;
    }
    if (igox <= STRUCTURE_LOADING) {
      device().set_precision(precision_);
      ground_now_ = ground_spec();
      device().set_ground(ground_now_);
      zload_ = load_impedances(geo_.scaled(freq_mhz_));
      device().set_loads(zload_);
      device().fill(rkh_);
      device().factor();
      times_.fill_ms = device().times().fill_ms;
      times_.factor_ms = device().times().factor_ms;
      state_ = EXCITATION_SETUP;
    }
    // excitation_loop, src/nec_context.cpp:1793-1855
    if (igox < EXCITATION_LOOP) {
      excitation(mhz);
      state_ = EXCITATION_LOOP;
      if (iflow_ == 7) {
        if (nfrq_ != 1) continue;
        return;
      }
    }
    // excitation_compute_near_field, src/nec_context.cpp:1741-1757: a
    // pending NE card is evaluated at every frequency of the run (before
    // the far field) and cleared after the last; with one frequency the
    // run ends there, otherwise the far field of the frequency follows
    if (igox < NEAR_FIELD) state_ = NEAR_FIELD;
    if (igox != FAR_FIELD && near_pending_) {
      compute_near_field();
      if (mhz == nfrq_) near_pending_ = false;
      if (nfrq_ == 1) return;
    }
    far_field();
    if (mhz == nfrq_) ifar_ = -1;
    if (nfrq_ == 1) return;
  } while (++mhz <= nfrq_);
}

void Context::excitation(int mhz) {
  // setup_excitation + netwk (src/nec_context.cpp:1561-1617, 3193-3210,
  // 4850-5047): RHS of the voltage sources, solve, current coefficients,
  // antenna input parameters.
  (void)mhz;
  const int n = static_cast<int>(geo_.size());
  const ScaledSegments seg = geo_.scaled(freq_mhz_);
  std::vector<zc> e(n, zmk(0.0, 0.0));
  for (size_t s = 0; s < src_seg_.size(); ++s) {
    const int64_t k = src_seg_[s] - 1;
    e[k] = zdivr(zneg(src_v_[s]), seg.s[k] * wavelength_);
  }
  std::vector<zc> curx;
  std::vector<zc> src_cur(src_seg_.size());
  network_loss_ = 0.0;
  if (networks_.empty()) {
    device().solve(e, &curx, &abc_);
    for (size_t s = 0; s < src_seg_.size(); ++s) src_cur[s] = curx[src_seg_[s] - 1];
  } else {
    solve_with_networks(seg, e, &curx, &src_cur);
  }
  times_.solve_ms = device().times().solve_ms;
  if (src_seg_.empty()) return;  // the reference records no antenna input without sources
  AntennaInput in;
  in.freq_hz = freq_mhz_ / 1.e-6;
  input_power_ = 0.0;
  for (size_t s = 0; s < src_seg_.size(); ++s) {
    const int64_t k = src_seg_[s] - 1;
    const zc volt = src_v_[s];
    const zc cur = zmulr(src_cur[s], wavelength_);
    const zc adm = zdiv(cur, volt);
    const zc imp = zdiv(volt, cur);
    const double power = 0.5 * zmul(volt, zconj(cur)).re;
    input_power_ += power;
    in.tag.push_back(geo_.tags[k]);
    in.segment.push_back(static_cast<int>(k + 1));
    in.voltage.push_back(volt);
    in.current.push_back(cur);
    in.impedance.push_back(imp);
    in.admittance.push_back(adm);
    in.power.push_back(power);
  }
  in.segment_currents.resize(n);
  for (int i = 0; i < n; ++i) in.segment_currents[i] = zmulr(curx[i], wavelength_);
  // ohmic loss in the loads, nec_context::compute_structure_power_loss
  // (src/nec_context.cpp:1423-1441)
  structure_loss_ = 0.0;
  if (!zload_.empty())
    for (int i = 0; i < n; ++i)
      if (std::fabs(zload_[i].re) >= 1.e-20) {
        const zc ci = zmulr(curx[i], wavelength_);
        const double cm = std::hypot(ci.re, ci.im);
        structure_loss_ += 0.5 * cm * cm * zload_[i].re * seg.s[i];
      }
  inputs_.push_back(std::move(in));
}

namespace {

// In-place partial-pivot LU of the row-major complex m x m matrix A (pivot:
// the first largest |z|, Eigen's PartialPivLU rule) and the solve with it:
// the network equations are a few dozen at most, so they stay on the host,
// as the reference keeps them (lu_decompose / solve, src/matrix_algebra.cpp
// :217-307).
void small_lu(std::vector<zc>& A, int m, std::vector<int>& piv) {
  piv.assign(m, 0);
  for (int k = 0; k < m; ++k) {
    int p = k;
    double best = std::hypot(A[(size_t)k * m + k].re, A[(size_t)k * m + k].im);
    for (int i = k + 1; i < m; ++i) {
      const double a = std::hypot(A[(size_t)i * m + k].re, A[(size_t)i * m + k].im);
      if (a > best) {
        best = a;
        p = i;
      }
    }
    piv[k] = p;
    if (p != k)
      for (int c = 0; c < m; ++c) std::swap(A[(size_t)k * m + c], A[(size_t)p * m + c]);
    const zc d = A[(size_t)k * m + k];
    if (d.re == 0.0 && d.im == 0.0) throw NecError("NETWORK DATA ERROR: SINGULAR NETWORK EQUATIONS");
    for (int i = k + 1; i < m; ++i) {
      const zc l = zdiv(A[(size_t)i * m + k], d);
      A[(size_t)i * m + k] = l;
      for (int c = k + 1; c < m; ++c) A[(size_t)i * m + c] = zsub(A[(size_t)i * m + c], zmul(l, A[(size_t)k * m + c]));
    }
  }
}

void small_solve(const std::vector<zc>& LU, int m, const std::vector<int>& piv, std::vector<zc>& b) {
  for (int k = 0; k < m; ++k)
    if (piv[k] != k) std::swap(b[k], b[piv[k]]);
  for (int i = 1; i < m; ++i)
    for (int k = 0; k < i; ++k) b[i] = zsub(b[i], zmul(LU[(size_t)i * m + k], b[k]));
  for (int i = m - 1; i >= 0; --i) {
    for (int k = i + 1; k < m; ++k) b[i] = zsub(b[i], zmul(LU[(size_t)i * m + k], b[k]));
    b[i] = zdiv(b[i], LU[(size_t)i * m + i]);
  }
}

}  // namespace

void Context::solve_with_networks(const ScaledSegments& seg, std::vector<zc>& e, std::vector<zc>* curx,
                                  std::vector<zc>* src_cur) {
  // netwk, src/nec_context.cpp:4496-5045, with the structure solves on the
  // GPU.  Port segments that are not sources get network equations, in
  // order of appearance; ports on sources become source rows (numbered from
  // the end, their voltage known).  M(j, i) = current at equation segment j
  // of a unit field at equation segment i, less y s of the networks; M v =
  // rhnx + the incident field's currents, then the port fields v are taken
  // from the incident field and the structure solved again.
  const int n = static_cast<int>(geo_.size());
  const double wl = wavelength_;
  // lines of length <= 0: the distance of the port centres, set once
  // (calculate_network_data, :1210-1241)
  for (Network& net : networks_)
    if (net.type >= 2 && net.x[1] <= 0.0) {
      const int64_t a = net.seg1 - 1, b = net.seg2 - 1;
      const double xx = seg.x[b] - seg.x[a], yy = seg.y[b] - seg.y[a], zz = seg.z[b] - seg.z[a];
      net.x[1] = wl * std::sqrt(xx * xx + yy * yy + zz * zz);
    }
  const int nsrc = static_cast<int>(src_seg_.size());
  const int ndimn = 2 * static_cast<int>(networks_.size()) + nsrc;
  std::vector<int64_t> nteqa, ntsca;
  std::vector<zc> vsrc;
  std::vector<zc> cmn((size_t)ndimn * ndimn, zmk(0.0, 0.0));  // fill(i, j) at i + j ndimn
  std::vector<zc> rhnx(ndimn, zmk(0.0, 0.0));
  for (const Network& net : networks_) {
    // short-circuit admittance parameters (:4660-4688)
    zc y11, y12, y22;
    if (net.type <= 1) {
      y11 = zmk(net.x[0], net.x[1]);
      y12 = zmk(net.x[2], net.x[3]);
      y22 = zmk(net.x[4], net.x[5]);
    } else {
      const double bl = (2.0 * kPi) * net.x[1] / wl;
      const double y12i = 1. / (net.x[0] * std::sin(bl));
      const double yc = -y12i * std::cos(bl);
      y11 = zmk(net.x[2], yc + net.x[3]);
      y12 = net.type == 2 ? zmk(0., y12i) : zmk(-0., -y12i);
      y22 = zmk(net.x[4], yc + net.x[5]);
    }
    const int64_t sp[2] = {net.seg1, net.seg2};
    int isc[2], row[2];
    for (int p = 0; p < 2; ++p) {
      isc[p] = -1;
      for (int i = 0; i < nsrc; ++i)
        if (sp[p] == src_seg_[i]) {
          isc[p] = i;
          break;
        }
      int found = -1;
      if (isc[p] < 0) {
        for (size_t i = 0; i < nteqa.size(); ++i)
          if (nteqa[i] == sp[p]) {
            found = static_cast<int>(i);
            break;
          }
        if (found < 0) {
          found = static_cast<int>(nteqa.size());
          nteqa.push_back(sp[p]);
        }
      } else {
        for (size_t i = 0; i < ntsca.size(); ++i)
          if (ntsca[i] == sp[p]) {
            found = ndimn - static_cast<int>(i + 1);
            break;
          }
        if (found < 0) {
          found = ndimn - static_cast<int>(ntsca.size() + 1);
          ntsca.push_back(sp[p]);
          vsrc.push_back(src_v_[isc[p]]);
        }
      }
      row[p] = found;
    }
    const double s1 = seg.s[sp[0] - 1], s2 = seg.s[sp[1] - 1];
    auto at = [&](int i, int j) -> zc& { return cmn[(size_t)i + (size_t)j * ndimn]; };
    if (isc[0] < 0) {
      at(row[0], row[0]) = zsub(at(row[0], row[0]), zmulr(y11, s1));
      at(row[0], row[1]) = zsub(at(row[0], row[1]), zmulr(y12, s1));
    } else {
      const zc v = src_v_[isc[0]];
      rhnx[row[0]] = zadd(rhnx[row[0]], zdivr(zmul(y11, v), wl));
      rhnx[row[1]] = zadd(rhnx[row[1]], zdivr(zmul(y12, v), wl));
    }
    if (isc[1] < 0) {
      at(row[1], row[1]) = zsub(at(row[1], row[1]), zmulr(y22, s2));
      at(row[1], row[0]) = zsub(at(row[1], row[0]), zmulr(y12, s2));
    } else {
      const zc v = src_v_[isc[1]];
      rhnx[row[0]] = zadd(rhnx[row[0]], zdivr(zmul(y12, v), wl));
      rhnx[row[1]] = zadd(rhnx[row[1]], zdivr(zmul(y22, v), wl));
    }
  }
  const int nteq = static_cast<int>(nteqa.size());
  const int ntsc = static_cast<int>(ntsca.size());
  DeviceSolver& d = device();
  std::vector<double> abc_tmp;
  // the structure's admittances between the equation segments: one GPU
  // solve per equation segment
  for (int i = 0; i < nteq; ++i) {
    std::vector<zc> unit(n, zmk(0.0, 0.0)), c;
    unit[nteqa[i] - 1] = zmk(1.0, 0.0);
    d.solve(unit, &c, &abc_tmp);
    for (int j = 0; j < nteq; ++j)
      cmn[(size_t)i + (size_t)j * ndimn] = zadd(cmn[(size_t)i + (size_t)j * ndimn], c[nteqa[j] - 1]);
  }
  // M(r, c) = fill(c, r)
  std::vector<zc> M((size_t)nteq * nteq);
  for (int r = 0; r < nteq; ++r)
    for (int c = 0; c < nteq; ++c) M[(size_t)r * nteq + c] = cmn[(size_t)c + (size_t)r * ndimn];
  std::vector<int> mpiv;
  if (nteq > 0) small_lu(M, nteq, mpiv);
  // the incident field's currents at the equation segments, the network
  // equations, the port fields off the incident field, the final solve
  std::vector<zc> c0;
  d.solve(e, &c0, &abc_tmp);
  std::vector<zc> rhnt(nteq);
  for (int i = 0; i < nteq; ++i) rhnt[i] = zadd(rhnx[i], c0[nteqa[i] - 1]);
  if (nteq > 0) small_solve(M, nteq, mpiv, rhnt);
  for (int i = 0; i < nteq; ++i) e[nteqa[i] - 1] = zsub(e[nteqa[i] - 1], rhnt[i]);
  d.solve(e, curx, &abc_);
  const std::vector<zc>& cur = *curx;
  // the power the networks take (netwk_compute_currents, :4891-4918)
  double loss = 0.0;
  for (int i = 0; i < nteq; ++i) {
    const int64_t r = nteqa[i] - 1;
    const zc volt = zmulr(zmulr(rhnt[i], seg.s[r]), wl);
    const zc c = zmulr(cur[r], wl);
    loss -= 0.5 * zmul(volt, zconj(c)).re;
  }
  for (int i = 0; i < ntsc; ++i) {
    const zc c = zmulr(cur[ntsca[i] - 1], wl);
    loss -= 0.5 * zmul(vsrc[i], zconj(c)).re;
  }
  // a source on a port: the network's current joins the structure's
  // (netwk_compute_inputs, :4954-4985)
  for (int s = 0; s < nsrc; ++s) {
    zc c = cur[src_seg_[s] - 1];
    if (ntsc != 0) {
      int j = -1;
      for (int q = 0; q < ntsc; ++q)
        if (ntsca[q] == src_seg_[s]) {
          j = q;
          break;
        }
      if (j < 0)
        throw NecError("NETWORK DATA ERROR: SOURCE SEGMENT [" + std::to_string(src_seg_[s]) +
                       "] NOT FOUND IN NETWORK TABLE");
      const int rowi = ndimn - (j + 1);
      zc t = rhnx[rowi];
      for (int k = 0; k < nteq; ++k) t = zsub(t, zmul(cmn[(size_t)k + (size_t)rowi * ndimn], rhnt[k]));
      c = zadd(c, t);
      const zc ci = zmulr(c, wl);
      loss += 0.5 * zmul(src_v_[s], zconj(ci)).re;
    }
    (*src_cur)[s] = c;
  }
  network_loss_ = loss;
}

PatternSpec Context::pattern_spec(std::vector<double>* trig_out) const {
  // gain constants and angle tables of nec_radiation_pattern::analyze
  // (src/nec_radiation_pattern.cpp:221-262)
  const double eta = em_impedance();
  const double pinr = input_power_, pnlr = network_loss_;
  PatternSpec sp;
  sp.n_theta = nth_;
  sp.n_phi = nph_;
  sp.ground = ground_now_.type;
  sp.zr_re = ground_now_.zr_re;
  sp.zr_im = ground_now_.zr_im;
  sp.output_format = rp_format_;
  sp.normalization = rp_norm_;
  sp.ipd = ipd_;
  sp.iavp = iavp_;
  sp.wavelength = wavelength_;
  sp.gcop = wavelength_ * wavelength_ * 2.0 * kPi / (eta * pinr);
  const double prad = pinr - structure_loss_ - pnlr;
  sp.gcon = sp.gcop;
  if (ipd_ != 0) sp.gcon = sp.gcon * pinr / prad;
  sp.rfld = rfld_;
  sp.exrm = 0.;
  sp.exra = 0.;
  if (rfld_ >= 1.0e-20) {
    sp.exrm = 1. / rfld_;
    sp.exra = rfld_ / wavelength_;
    sp.exra = -360. * (sp.exra - std::floor(sp.exra));
  }
  sp.dphi_rad = deg_to_rad(dph_);
  sp.half_dtheta_rad = 0.5 * deg_to_rad(dth_);
  // the reference accumulates the angles by repeated addition
  std::vector<double> trig(3 * (size_t)nth_ + 2 * (size_t)nph_ + (sp.ground ? (size_t)nth_ : 0));
  double thet = thets_ - dth_;
  for (int k = 0; k < nth_; ++k) {
    thet += dth_;
    const double tha = deg_to_rad(thet);
    trig[k] = std::sin(tha);
    trig[nth_ + k] = std::cos(tha);
    trig[2 * nth_ + k] = tha;
    // over a ground, directions below the horizon are skipped
    // (src/nec_radiation_pattern.cpp:263-264)
    if (sp.ground) trig[3 * (size_t)nth_ + 2 * (size_t)nph_ + k] = (thet > 90.01) ? 1.0 : 0.0;
  }
  double phi = phis_ - dph_;
  for (int k = 0; k < nph_; ++k) {
    phi += dph_;
    const double pha = deg_to_rad(phi);
    trig[3 * nth_ + k] = std::sin(pha);
    trig[3 * nth_ + nph_ + k] = std::cos(pha);
  }
  *trig_out = std::move(trig);
  return sp;
}

void Context::far_field() {
  // excitation_compute_far_field + nec_radiation_pattern::analyze
  // (src/nec_context.cpp:1761-1791, src/nec_radiation_pattern.cpp:209-450)
  if (ifar_ == -1) return;
  last_was_pattern_ = true;
  PatternResult pr;
  pr.freq_hz = freq_mhz_ / 1.e-6;
  pr.n_theta = nth_;
  pr.n_phi = nph_;
  pr.theta0 = thets_;
  pr.phi0 = phis_;
  pr.dtheta = dth_;
  pr.dphi = dph_;
  pr.normalization = rp_norm_;
  pr.output_format = rp_format_;
  pr.iavp = iavp_;
  std::vector<double> trig;
  const PatternSpec sp = pattern_spec(&trig);
  device().pattern(sp, trig, &pr.a);
  times_.pattern_ms = device().times().pattern_ms;
  const int64_t nd = (int64_t)nth_ * nph_;
  if (iavp_ != 0) {
    double pint = 0.0;
    for (int64_t d = 0; d < nd; ++d) pint += pr.a.pint[d];
    const double ts = deg_to_rad(thets_);
    const double tr = deg_to_rad(dth_) * static_cast<double>(nth_ - 1);
    const double prr = deg_to_rad(dph_) * static_cast<double>(nph_ - 1);
    const double sa = std::fabs(prr * (std::cos(ts) - std::cos(ts + tr)));
    pr.average_power_gain = pint / sa;
    pr.average_power_solid_angle = sa / kPi;
  }
  double mx = pr.a.gain[0];
  for (int64_t d = 1; d < nd; ++d)
    if (pr.a.gain[d] > mx) mx = pr.a.gain[d];
  pr.max_gain = mx;
  // statistics, nec_radiation_pattern.h:252-333 (rows = theta, cols = phi)
  const std::vector<double>* arrs[3] = {&pr.a.pg_tot, &pr.a.pg_rhcp, &pr.a.pg_lhcp};
  for (int a = 0; a < 3; ++a) {
    const std::vector<double>& p = *arrs[a];
    double mxv = p[0], mnv = p[0];
    for (int64_t d = 1; d < nd; ++d) {
      if (p[d] > mxv) mxv = p[d];
      if (p[d] < mnv) mnv = p[d];
    }
    double sum = 0.0;
    for (int i = 0; i < nth_; ++i)
      for (int j = 0; j < nph_; ++j) sum += p[(int64_t)j * nth_ + i] * pr.a.avg_scale[(int64_t)j * nth_ + i];
    const double norm = static_cast<double>(nd * 2) / kPi;
    const double mean = sum / norm;
    double s2 = 0.0;
    for (int i = 0; i < nth_; ++i)
      for (int j = 0; j < nph_; ++j) {
        const double diff = p[(int64_t)j * nth_ + i] - mean;
        s2 += diff * diff * pr.a.avg_scale[(int64_t)j * nth_ + i];
      }
    pr.stats[4 * a + 0] = mxv;
    pr.stats[4 * a + 1] = mnv;
    pr.stats[4 * a + 2] = mean;
    pr.stats[4 * a + 3] = std::sqrt(s2 / norm);
  }
  patterns_.push_back(std::move(pr));
}

const AntennaInput* Context::input(int k) const {
  if (k < 0 || k >= static_cast<int>(inputs_.size())) return nullptr;
  return &inputs_[k];
}

const PatternResult* Context::pattern(int k) const {
  if (k < 0 || k >= static_cast<int>(patterns_.size())) return nullptr;
  return &patterns_[k];
}

void Context::benchmark_solve(int warmup, int steps, double* out) {
  if (!dev_ || inputs_.empty()) throw NecError("benchmark_solve needs a completed XQ or RP first");
  const int n = static_cast<int>(geo_.size());
  const ScaledSegments seg = geo_.scaled(freq_mhz_);
  std::vector<zc> e(n, zmk(0.0, 0.0));
  std::vector<int> segs;
  for (size_t s = 0; s < src_seg_.size(); ++s) {
    const int64_t k = src_seg_[s] - 1;
    e[k] = zdivr(zneg(src_v_[s]), seg.s[k] * wavelength_);
    segs.push_back(static_cast<int>(k));
  }
  std::vector<double> trig;
  PatternSpec sp{};
  const PatternSpec* psp = nullptr;
  if (last_was_pattern_ && nth_ > 0 && nph_ > 0) {
    sp = pattern_spec(&trig);
    psp = &sp;
  }
  double ms = 0;
  StageTimes st;
  std::vector<double> zin;
  device().benchmark_pipeline(rkh_, e, segs, src_v_, wavelength_, psp, trig, warmup, steps, &ms, &st, &zin);
  out[0] = ms;
  out[1] = st.fill_ms;
  out[2] = st.factor_ms;
  out[3] = st.solve_ms;
  out[4] = st.pattern_ms;
  out[5] = zin.empty() ? 0.0 : zin[0];
  out[6] = zin.empty() ? 0.0 : zin[1];
}

void Context::fill_matrix(double freq_mhz, std::vector<double>* Z) {
  if (!geo_.is_complete()) throw NecError("geometry is not complete");
  DeviceSolver& d = device();
  d.load(geo_, geo_.scaled(freq_mhz), use_ek_);
  const double wl = wavelength_;
  wavelength_ = em_speed_of_light() / (1.0e6 * freq_mhz);
  d.set_ground(ground_spec());
  d.set_loads(load_impedances(geo_.scaled(freq_mhz)));
  wavelength_ = wl;
  d.fill(rkh_);
  d.download_matrix(Z);
  state_ = MEMORY_ALLOC;  // the device matrix no longer holds factors
}

}  // namespace necgpu
