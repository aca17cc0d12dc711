// context.h -- the nec_context behind the C interface.
//
// Card state and the simulate() state machine of nec_context
// (src/nec_context.cpp:84-164, 406-420, 519-612, 741-868, 1031-1159,
// 1523-1855) for the wire / free-space / voltage-source subset, with the
// numerical stages delegated to the DeviceSolver.  Results accumulate in
// order and are fetched by ordinal like nec_results (src/nec_results.h:711-722).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "adjoint.h"
#include "engine.h"
#include "geometry.h"

namespace necgpu {

struct AntennaInput {
  double freq_hz = 0;
  std::vector<int> tag, segment;
  std::vector<zc> voltage, current, impedance, admittance;
  std::vector<double> power;
  std::vector<zc> segment_currents;  // all segments, amperes (curx * wavelength)
};

struct PatternResult {
  double freq_hz = 0;
  int n_theta = 0, n_phi = 0;
  double theta0 = 0, phi0 = 0, dtheta = 0, dphi = 0;
  int normalization = 0, output_format = 0, iavp = 0;
  HostPattern a;
  double max_gain = -999.0;
  double average_power_gain = 0, average_power_solid_angle = 0;
  double stats[12] = {0};  // tot/rhcp/lhcp x (max, min, mean, sd)
};

class Context {
 public:
  Context();
  ~Context();

  // cards
  void wire(int tag, int ns, double x1, double y1, double z1, double x2, double y2, double z2, double rad,
            double rdel, double rrad);
  // GM card (nec_context::move, src/nec_context.cpp:345-350): angles in radians
  void move(double rox, double roy, double roz, double xs, double ys, double zs, int its, int nrpt, int itgi);
  // GS card of a deck: scale the segments defined so far
  void scale(double f);
  void geometry_complete(int gpflag);
  void fr_card(int ifrq, int nfrq, double freq_mhz, double del_freq);
  void ex_card(int type, int i2, int i3, int i4, double f1, double f2, double f3, double f4, double f5, double f6);
  void xq_card(int itmp1);
  void rp_card(int calc_mode, int n_theta, int n_phi, int output_format, int normalization, int D, int A,
               double theta0, double phi0, double delta_theta, double delta_phi, double radial_distance,
               double gain_norm);
  void kh_card(double rkh);
  // GN: iperf -1 free space, 0 finite ground (reflection coefficients), 1
  // perfect ground; no radial-wire screen, no Sommerfeld ground
  void gn_card(int iperf, int nradl, double epsr, double sig, double t3, double t4, double t5, double t6);
  // LD: types 0-5 (series / parallel RLC lumped or per metre, fixed
  // impedance, wire conductivity); consecutive LD cards accumulate, -1 clears
  void ld_card(int type, int tag, int tagf, int tagt, double zlr, double zli, double zlc);
  // NE card (src/nec_context.cpp:896-955): near electric field on a
  // rectangular (near = 0) or spherical (near = 1) grid of points
  void tl_card(int i1, int i2, int i3, int i4, double z0, double len, double y1r, double y1i, double y2r,
               double y2i);
  void nt_card(int i1, int i2, int i3, int i4, double y11r, double y11i, double y12r, double y12i, double y22r,
               double y22i);
  void ne_card(int near, int nrx, int nry, int nrz, double xnr, double ynr, double znr, double dxnr, double dynr,
               double dznr);
  void nh_card(int near, int nrx, int nry, int nrz, double xnr, double ynr, double znr, double dxnr, double dynr,
               double dznr);
  // near-field results: points (metres, 3 each) and E or H (3 complex each)
  struct NearFieldResult {
    double freq_hz = 0.0;
    bool magnetic = false;  // H (NH card) rather than E (NE card)
    std::vector<double> xyz;
    std::vector<zc> e;      // the field, 3 complex per point
  };
  const NearFieldResult* near_field(int k) const {
    return (k >= 0 && k < (int)near_fields_.size()) ? &near_fields_[k] : nullptr;
  }
  void ek_card(bool on);

  // results
  const AntennaInput* input(int k) const;
  const PatternResult* pattern(int k) const;
  int input_count() const { return static_cast<int>(inputs_.size()); }
  int pattern_count() const { return static_cast<int>(patterns_.size()); }

  // GPU extras
  // takes effect at the next XQ / RP, which refills and refactors
  void set_precision(int mode) {
    if (mode != precision_) state_ = MEMORY_ALLOC;
    precision_ = mode;
  }
  int precision() const { return precision_; }
  const StageTimes& last_times() const { return times_; }
  const Geometry& geometry() const { return geo_; }
  double wavelength() const { return wavelength_; }
  double frequency_mhz() const { return freq_mhz_; }
  bool use_ek() const { return use_ek_; }
  double rkh() const { return rkh_; }
  // Fill only (no factorisation) at the current frequency; Z row-major 2*n*n.
  void fill_matrix(double freq_mhz, std::vector<double>* Z);
  // Repeat the last frequency's whole solve (fill, LU, solve, Z_in and, if
  // the last card was RP, the pattern) on the device.  out: ms per solve,
  // fill / factor / solve / pattern ms, Z_in of feed 0 (re, im).
  void benchmark_solve(int warmup, int steps, double* out);
  // Adjoint gradient of the last solve (one voltage source, FP64 path) for
  // the objective of `spec`: the segment endpoint and radius gradient (7
  // per segment, per metre) and dJ/d(frequency in MHz).
  void adjoint(const AdjointSpec& spec, AdjointResult* r, std::vector<double>* grad7, double* dfreq_mhz);
  DeviceSolver& device();

 private:
  enum State { MEMORY_ALLOC = 1, STRUCTURE_LOADING = 2, EXCITATION_SETUP = 3, EXCITATION_LOOP = 4,
               NEAR_FIELD = 5, FAR_FIELD = 6 };
  void calc_prepare();
  void simulate(bool far_field_flag);
  void excitation(int mhz);
  void far_field();
  PatternSpec pattern_spec(std::vector<double>* trig) const;
  void reset_to_structure_loading();
  void reset_to_excitation_setup();

  Geometry geo_;
  std::unique_ptr<DeviceSolver> dev_;
  int iflow_ = 1;
  State state_ = MEMORY_ALLOC;
  int ifrq_ = 0, nfrq_ = 1;
  double freq_mhz_ = 0, delfrq_ = 0, wavelength_ = 0;
  double rkh_ = 1.0;
  bool use_ek_ = false;
  int gn_iperf_ = -1;  // -1: free space
  double gn_epsr_ = 0.0, gn_sig_ = 0.0;
  // ground of the current frequency (nec_ground::calculate_antenna_environment)
  GroundSpec ground_spec();
  GroundSpec ground_now_;  // the ground of the last fill
  struct LoadCard {
    int type, tag, tagf, tagt;
    double zlr, zli, zlc;
  };
  std::vector<LoadCard> loads_;
  struct NearSpec {
    int nfeh = 0;  // 0 electric (NE), 1 magnetic (NH)
    int near = 0, nrx = 1, nry = 1, nrz = 1;
    double xnr = 0, ynr = 0, znr = 0, dxnr = 0, dynr = 0, dznr = 0;
  } near_spec_;
  // an NE card waits for evaluation (the reference's m_near != -1): in every
  // frequency of the next run, before the far field, until the last one
  bool near_pending_ = false;
  std::vector<NearFieldResult> near_fields_;
  void compute_near_field();
  void near_card(int nfeh, int near, int nrx, int nry, int nrz, double xnr, double ynr, double znr, double dxnr,
                 double dynr, double dznr);
  // lambda * impedance per unit length of every segment at the current
  // frequency (nec_context::load); empty without loads
  std::vector<zc> load_impedances(const ScaledSegments& seg) const;
  std::vector<zc> zload_;          // of the last fill
  double structure_loss_ = 0.0;    // ohmic loss in the loads (W), last solve
  int precision_ = 0;
  // voltage sources (EX 0)
  std::vector<int64_t> src_seg_;
  std::vector<zc> src_v_;
  // networks and transmission lines (NT / TL cards, src/nec_context.cpp
  // :630-726): type 1 network, 2 straight line, 3 crossed line; 1-based
  // port segments; x as the card gives it (a line's length <= 0 is replaced
  // by the distance of the port centres at the first solve)
  struct Network {
    int type;
    int64_t seg1, seg2;
    double x[6];
  };
  std::vector<Network> networks_;
  double network_loss_ = 0.0;  // network_power_loss of the last solve (W)
  void network_card(int type, int i1, int i2, int i3, int i4, const double* x);
  // the solve of the structure with its networks (netwk), see excitation
  void solve_with_networks(const ScaledSegments& seg, std::vector<zc>& e, std::vector<zc>* curx,
                           std::vector<zc>* src_current);
  // RP card
  int ifar_ = -1;
  int nth_ = 0, nph_ = 0, rp_format_ = 0, rp_norm_ = 0, ipd_ = 0, iavp_ = 0;
  double thets_ = 0, phis_ = 0, dth_ = 0, dph_ = 0, rfld_ = 0, gnor_ = 0;
  double input_power_ = 0;
  bool last_was_pattern_ = false;
  std::vector<double> abc_;  // current coefficients of the last solve
  StageTimes times_;
  std::vector<AntennaInput> inputs_;
  std::vector<PatternResult> patterns_;
};

// The card-deck front end (deck.cpp): runs NEC input text on ctx through
// the libnecpp calls; throws NecError naming the first failing card.
void run_deck(Context* ctx, const char* text);

}  // namespace necgpu
