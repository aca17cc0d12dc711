// engine.h -- device-resident MoM solver for one structure (one GPU).
//
// Owns the HBM buffers of the path and runs the stages that nec_context's
// simulate() runs for a voltage-excited structure at one frequency:
//   fill      cmset                 -> Z (row-major, "cm" layout)
//   factor    factrs / lu_decompose -> L\U in place, pivots
//   solve     etmns -> solves -> get_current_coefficients -> Z_in
//   pattern   nec_radiation_pattern::analyze / ffld
// Host-side tables (segments in wavelengths, basis coefficients) are built
// from the Geometry and uploaded once per frequency.  Every stage has an
// enqueue_* form that only queues work on the solver's stream, so a whole
// solve can run without a host round trip (benchmark_pipeline).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "geometry.h"
#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

void hip_check(hipError_t e, const char* what);

// Device allocation that grows on demand and frees on destruction.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  void* ensure(size_t bytes);
  void release();
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr_);
  }
  size_t size() const { return bytes_; }

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

struct StageTimes {
  double fill_ms = 0, factor_ms = 0, solve_ms = 0, pattern_ms = 0;
};

// Per-direction pattern arrays on the host (n_theta * n_phi each).
struct HostPattern {
  std::vector<double> gain, pg_vert, pg_horiz, pg_tot, pg_rhcp, pg_lhcp, axial, tilt, avg_scale, pint;
  std::vector<int> sense;
  std::vector<double> e_theta, e_phi;  // interleaved complex
};

// Basis tables of one structure at one frequency.  Column CSR (from trio):
// for basis k, the source segments j with support, ascending j, and their
// (A, B, C) coefficients -- the fill's gather order.  Segment CSR (from tbf):
// for segment i, the bases with support on it, ascending, and coefficients --
// the current coefficients' gather order.  Indices are 0-based.
struct BasisTables {
  std::vector<int> colptr, colj, segptr, segbasis;
  std::vector<double> colabc, segabc;
};
void build_basis_tables(const Geometry& geo, const ScaledSegments& seg, BasisTables* out);
// The index arrays of build_basis_tables (which depend on the connections
// only) and, for the q-th entry of trio(j) (at tptr[j] + q) and of tbf(i)
// (at bptr[i] + q), the position cpos / spos of its coefficients in the
// column / segment CSR: what the device needs to place the coefficients
// it computes.
struct BasisShape {
  std::vector<int> colptr, colj, tptr, cpos, segptr, segbasis, bptr, spos;
};
void build_basis_shape(const Geometry& geo, const ScaledSegments& seg, BasisShape* out);

class DeviceSolver {
 public:
  DeviceSolver();
  ~DeviceSolver();
  DeviceSolver(const DeviceSolver&) = delete;
  DeviceSolver& operator=(const DeviceSolver&) = delete;

  // Upload the segment table and basis tables of one frequency.
  void load(const Geometry& geo, const ScaledSegments& seg, bool use_ek);

  // synchronous stages (timed with HIP events)
  void fill(double rkh);
  void factor();
  // e: right-hand side (n complex).  On return curx = segment-centre currents
  // (get_current_coefficients, not yet scaled by the wavelength) and abc (6n)
  // the current coefficients.
  void solve(const std::vector<zc>& e, std::vector<zc>* curx, std::vector<double>* abc);
  // Pattern of the last solve.  trig: sin/cos/angle tables, see kernels.h.
  void pattern(const PatternSpec& spec, const std::vector<double>& trig, HostPattern* out);

  // Whole solve (fill, LU, solve, currents, Z_in, pattern) repeated on the
  // device with inputs already in HBM; returns the mean time per solve and
  // the mean per-stage times.  sources: 0-based segments and voltages.
  void benchmark_pipeline(double rkh, const std::vector<zc>& e, const std::vector<int>& src_seg,
                          const std::vector<zc>& src_v, double wavelength, const PatternSpec* spec,
                          const std::vector<double>& trig, int warmup, int steps, double* ms_per_solve,
                          StageTimes* stage_ms, std::vector<double>* zin_out);

  // Adjoint pass of the last solve, with the FP64 factors: lam = Z^-T g
  // (g: n complex), then the fill's adjoint contraction (launch_adjoint_fill):
  // gi (7 per segment) and gj (13 per segment).  x receives the solution of
  // the last solve (n complex).
  void adjoint(double rkh, const std::vector<zc>& g, std::vector<zc>* lam, std::vector<zc>* x,
               std::vector<double>* gi, std::vector<double>* gj);
  // time of the last adjoint pass (transposed solve + contraction), ms
  double adjoint_ms() const { return adjoint_ms_; }

  void download_matrix(std::vector<double>* Z);  // current content of Z (2*n*n)
  int n() const { return n_; }
  int lu_info() const { return lu_info_; }
  const StageTimes& times() const { return times_; }
  // per-kernel launch times of the timed steps of the last benchmark_pipeline
  const KernelClock& kernel_clock() const { return kclock_; }
  // per-launch start/stop events in benchmark_pipeline (default on); their
  // marker packets cost about 5 % of an N = 3,280 solve, so headline
  // timings run with them off
  void set_kernel_timing(bool on) { kernel_timing_ = on; }
  // 0: FP64 LU.  1: mixed (FP32 LU of a copy of Z + FP64 iterative
  // refinement against Z until |b - Z x| <= refine_tol |b|).
  // ground of the next fills (the image term of efld)
  void set_ground(const GroundSpec& g) { ground_ = g; }
  // loads of the next fills: lambda * impedance per unit length of every
  // segment (n values), or none (empty)
  void set_loads(const std::vector<zc>& z);
  const GroundSpec& ground() const { return ground_; }
  void set_precision(int mode) { precision_ = mode; }
  int precision() const { return precision_; }
  void set_refine_tol(double tol) { refine_tol_ = tol; }
  // refinement steps and final relative residual of the last mixed solve
  int refine_steps() const { return refine_steps_; }
  // pivot mode of the last FP32 factors (1 block-local, 0 partial) and how
  // often the refinement fell back to partial pivoting
  int mixed_pivot() const { return mixed_pivot_; }
  long refine_fallbacks() const { return refine_fallbacks_; }
  // factorisations repeated on one stream after a panel exchange timed out
  // (the panel's workgroups could not all be resident beside the look-ahead
  // update, e.g. on a device shared with other work)
  long lu_fallbacks() const { return lu_fallbacks_; }
  // Near electric (magnetic false) or magnetic field of the last solve's
  // currents, over the ground of the last fill, at points given in
  // wavelengths (3 per point), ax as for launch_near_field; 3 complex per
  // point.
  std::vector<zc> near_field(bool magnetic, const std::vector<double>& pts, const std::vector<double>& ax,
                             double rkh);
  double refine_residual() const { return refine_residual_; }
  hipStream_t stream() const { return stream_; }

 private:
  void ensure_device();
  DevSegments dev_segments() const;
  void enqueue_fill(double rkh);
  void enqueue_factor_f32(int pivot_mode);
  void enqueue_factor();
  void ensure_solve_work();
  void enqueue_solve(const void* rhs_dev, hipMemcpyKind kind, const void* rhs_host);
  PatternOut pattern_buffers(int nd);
  void enqueue_pattern(const PatternSpec& spec, const PatternOut& po);
  float elapsed(hipEvent_t a, hipEvent_t b);

  int n_ = 0;
  long long ld_ = 0;
  bool use_ek_ = false;
  GroundSpec ground_;
  DevBuf zload_dev_;
  bool has_loads_ = false;
  bool factored_ = false;
  // waits for the enqueued factorisation: returns the zero-pivot word and
  // whether a panel exchange timed out
  int factor_wait(bool* timed_out);
  // the same, throwing if a panel exchange timed out
  int factor_outcome();
  // false: the next factorisations keep every launch on stream_ (set for
  // the refactorisation after a timed-out panel exchange)
  bool lookahead_ = true;
  double last_rkh_ = 1.0;  // rkh of the last fill (a refactorisation refills)
  int lu_info_ = 0;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_[6] = {};
  StageTimes times_;
  KernelClock kclock_;
  int precision_ = 0;
  bool kernel_timing_ = true;
  double refine_tol_ = 1e-12;
  int refine_steps_ = 0;
  double refine_residual_ = 0.0;
  static constexpr int kMaxRefine = 30;
  // refinement steps allowed on block-local-pivot factors before falling
  // back to partial pivoting
  static constexpr int kMaxLocalRefine = 10;
  int mixed_pivot_ = 1;        // pivot mode of the current FP32 factors
  long refine_fallbacks_ = 0;
  long lu_fallbacks_ = 0;  // refactorisations with partial pivoting
  void refine_solve();

  DevBuf seg_, ekind_, colptr_, colj_, colabc_, segptr_, segbasis_, segabc_;
  DevBuf ford_, fcjp_;  // the fill's field order and the column CSR's sources in it
  // the one-pass fill's clusters of bases (fill_tile_kernel); fill_rb_ = 0
  // when the longest cluster source list does not fit the LDS
  DevBuf cl_src_ptr_, cl_src_, cl_base_ptr_, cl_base_, cjl_;
  int fill_clusters_ = 0, fill_smax_ = 0, fill_rb_ = 0;
  DevBuf Z_, E_, piv_, info_, lu_work_, x_, abc_, curx_, solve_work_, rhs_, src_, zin_;
  DevBuf trig_, pat_part_, pat_out_, pat_sense_;
  DevBuf Z32_, b_, r_, acc_;  // mixed path: FP32 factors, right-hand side, residual, norms
  DevBuf adj_g_, adj_lam_, adj_acc_, adj_gi_, adj_gj_, adj_pj_;
  double adjoint_ms_ = 0.0;
};

}  // namespace necgpu
