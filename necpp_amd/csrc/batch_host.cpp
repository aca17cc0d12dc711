// batch_host.cpp -- host side of the batched path: many small structures, each
// at many frequencies, solved with one launch per stage for the whole batch.
//
// For every (design, frequency) the host builds what nec_context would build
// for one run (geometry in wavelengths, basis tables; c_geometry.cpp
// geometry_complete / frequency_scale / trio / tbf) and the voltage-source
// right-hand side (etmns_voltage_source, nec_context.cpp:3193-3210), all
// concatenated.  The device pipeline is
//   batch fill -> batch LU -> batch solve (+ currents, Z_in) -> pattern cut,
// captured once in a hipGraph and replayed.  The batched LU's panel width
// is picked by timing the candidates on the batch itself (autotune).
#include "batch.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "engine.h"
#include "nec_math.h"

namespace necgpu {

namespace {

template <class T>
void upload(DevBuf& buf, const std::vector<T>& v, hipStream_t s) {
  void* p = buf.ensure(std::max<size_t>(1, v.size() * sizeof(T)));
  if (!v.empty()) hip_check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}

struct StreamGuard {
  hipStream_t s = nullptr;
  ~StreamGuard() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  EventPair() {
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&b), "event");
  }
  ~EventPair() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  float ms() const {
    float v = 0.0f;
    hip_check(hipEventElapsedTime(&v, a, b), "event time");
    return v;
  }
};

}  // namespace

void run_batch(const BatchInput& in, BatchOutput* out) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    throw NecError("no HIP device available: the GPU path needs an MI355X (gfx950)");
  if (in.n_designs <= 0 || in.n_freq <= 0) throw NecError("batch: no designs or no frequencies");
  const auto t_host0 = std::chrono::steady_clock::now();
  const int S = in.n_designs * in.n_freq;
  std::vector<BatchSolve> sol(S);
  std::vector<double> seg;
  std::vector<int> colptr, colj, segptr, segbasis;
  std::vector<double> colabc, segabc;
  long long rows = 0, zents = 0;
  for (int d = 0; d < in.n_designs; ++d) {
    Geometry geo;
    for (int w = in.wire_ptr[d]; w < in.wire_ptr[d + 1]; ++w) {
      const double* gm = in.wire_geom + 7 * (size_t)w;
      geo.add_wire(in.wire_tag[w], in.wire_ns[w], gm[0], gm[1], gm[2], gm[3], gm[4], gm[5], gm[6], 1.0, 1.0);
    }
    geo.complete(0);
    const int n = static_cast<int>(geo.size());
    if (n < 1 || n > kBatchMaxN)
      throw NecError("batch: design " + std::to_string(d) + " has " + std::to_string(n) +
                     " segments; the batched path takes 1 to " + std::to_string(kBatchMaxN));
    const int64_t srcno = geo.segment_number(in.src_tag[d], in.src_seg[d]);
    const int src = static_cast<int>(srcno - 1);
    const zc v = zmk(in.src_volt[2 * d], in.src_volt[2 * d + 1]);
    for (int f = 0; f < in.n_freq; ++f) {
      const double fmhz = in.freq_mhz[(size_t)d * in.n_freq + f];
      const ScaledSegments sc = geo.scaled(fmhz);
      BasisTables t;
      build_basis_tables(geo, sc, &t);
      BatchSolve& b = sol[(size_t)d * in.n_freq + f];
      b.n = n;
      b.src = src;
      b.row0 = rows;
      b.seg0 = (long long)seg.size();
      b.cptr0 = (long long)colptr.size();
      b.csr0 = (long long)colj.size();
      b.sptr0 = (long long)segptr.size();
      b.ssr0 = (long long)segbasis.size();
      b.z0 = zents;
      b.wavelength = em_speed_of_light() / (1.0e6 * fmhz);
      // e[src] = -V / (s lambda), s the scaled segment length
      const zc e = zdivr(zneg(v), sc.s[src] * b.wavelength);
      b.er = e.re;
      b.ei = e.im;
      b.vr = v.re;
      b.vi = v.im;
      for (const std::vector<double>* a : {&sc.x, &sc.y, &sc.z, &sc.s, &sc.b, &sc.cab, &sc.sab, &sc.salp})
        seg.insert(seg.end(), a->begin(), a->end());
      colptr.insert(colptr.end(), t.colptr.begin(), t.colptr.end());
      colj.insert(colj.end(), t.colj.begin(), t.colj.end());
      colabc.insert(colabc.end(), t.colabc.begin(), t.colabc.end());
      segptr.insert(segptr.end(), t.segptr.begin(), t.segptr.end());
      segbasis.insert(segbasis.end(), t.segbasis.begin(), t.segbasis.end());
      segabc.insert(segabc.end(), t.segabc.begin(), t.segabc.end());
      rows += n;
      zents += (long long)n * n;
    }
  }
  // pattern directions (degrees -> sin / cos, as the RP card's tables)
  std::vector<double> dir(4 * (size_t)std::max(0, in.n_dir));
  for (int q = 0; q < in.n_dir; ++q) {
    const double th = deg_to_rad(in.theta_deg[q]), ph = deg_to_rad(in.phi_deg[q]);
    dir[4 * q] = std::sin(th);
    dir[4 * q + 1] = std::cos(th);
    dir[4 * q + 2] = std::sin(ph);
    dir[4 * q + 3] = std::cos(ph);
  }
  out->host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host0).count();

  StreamGuard sg;
  hip_check(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking), "stream");
  const hipStream_t st = sg.s;
  DevBuf d_sol, d_seg, d_colptr, d_colj, d_colabc, d_segptr, d_segbasis, d_segabc, d_dir;
  DevBuf d_Z, d_Zbak, d_rp, d_info, d_x, d_abc, d_zin, d_gain;
  upload(d_sol, sol, st);
  upload(d_seg, seg, st);
  upload(d_colptr, colptr, st);
  upload(d_colj, colj, st);
  upload(d_colabc, colabc, st);
  upload(d_segptr, segptr, st);
  upload(d_segbasis, segbasis, st);
  upload(d_segabc, segabc, st);
  upload(d_dir, dir, st);
  double* Z = static_cast<double*>(d_Z.ensure((size_t)zents * 16));
  int* rp = static_cast<int*>(d_rp.ensure((size_t)rows * sizeof(int)));
  int* info = static_cast<int*>(d_info.ensure((size_t)S * sizeof(int)));
  double* x = static_cast<double*>(d_x.ensure((size_t)rows * 16));
  double* abc = static_cast<double*>(d_abc.ensure((size_t)rows * 48));
  double* zin = static_cast<double*>(d_zin.ensure((size_t)S * 32));
  double* gain = static_cast<double*>(d_gain.ensure(std::max<size_t>(8, (size_t)S * std::max(0, in.n_dir) * 8)));
  const BatchSolve* dsol = d_sol.as<BatchSolve>();
  const double rkh = in.rkh > 0.0 ? in.rkh : 1.0;

  auto fill = [&]() {
    if (launch_batch_fill(dsol, S, rows, d_seg.as<double>(), d_colptr.as<int>(), d_colj.as<int>(),
                          d_colabc.as<double>(), Z, rkh, st))
      hip_check(hipGetLastError(), "batch fill");
  };
  auto lu = [&](int P) {
    if (launch_batch_lu(dsol, S, Z, rp, info, P, st)) hip_check(hipGetLastError(), "batch lu");
  };
  auto rest = [&]() {
    if (launch_batch_solve(dsol, S, Z, rp, d_seg.as<double>(), d_segptr.as<int>(), d_segbasis.as<int>(),
                           d_segabc.as<double>(), x, abc, zin, st))
      hip_check(hipGetLastError(), "batch solve");
    if (in.n_dir > 0 &&
        launch_batch_pattern(dsol, S, d_seg.as<double>(), abc, zin, in.n_dir, d_dir.as<double>(), gain, st))
      hip_check(hipGetLastError(), "batch pattern");
  };

  // autotune: the LU panel width, timed on this batch's own matrices
  EventPair e1, e2, e3, e4;
  hip_check(hipEventRecord(e1.a, st), "event");
  fill();
  hip_check(hipEventRecord(e1.b, st), "event");
  int bestP = in.panel_width;
  if (bestP != 8 && bestP != 16 && bestP != 32) {
    void* bak = d_Zbak.ensure((size_t)zents * 16);
    hip_check(hipMemcpyAsync(bak, Z, (size_t)zents * 16, hipMemcpyDeviceToDevice, st), "copy");
    float best = 1e30f;
    for (int P : {8, 16, 32}) {
      hip_check(hipMemcpyAsync(Z, bak, (size_t)zents * 16, hipMemcpyDeviceToDevice, st), "copy");
      hip_check(hipEventRecord(e2.a, st), "event");
      lu(P);
      hip_check(hipEventRecord(e2.b, st), "event");
      hip_check(hipEventSynchronize(e2.b), "autotune");
      const float ms = e2.ms();
      out->autotune_ms[P == 8 ? 0 : (P == 16 ? 1 : 2)] = ms;
      if (ms < best) {
        best = ms;
        bestP = P;
      }
    }
    hip_check(hipMemcpyAsync(Z, bak, (size_t)zents * 16, hipMemcpyDeviceToDevice, st), "copy");
    d_Zbak.release();
  }
  out->panel_width = bestP;
  // one pass with per-stage events
  hip_check(hipEventRecord(e2.a, st), "event");
  lu(bestP);
  hip_check(hipEventRecord(e2.b, st), "event");
  hip_check(hipEventRecord(e3.a, st), "event");
  rest();
  hip_check(hipEventRecord(e3.b, st), "event");
  hip_check(hipEventSynchronize(e3.b), "batch");
  out->fill_ms = e1.ms();
  out->lu_ms = e2.ms();
  out->solve_pattern_ms = e3.ms();
  // eager results, to check the graph replays against
  std::vector<double> eager_zin(4 * (size_t)S), eager_gain((size_t)S * std::max(0, in.n_dir));
  hip_check(hipMemcpyAsync(eager_zin.data(), zin, (size_t)S * 32, hipMemcpyDeviceToHost, st), "zin");
  if (in.n_dir > 0)
    hip_check(hipMemcpyAsync(eager_gain.data(), gain, eager_gain.size() * 8, hipMemcpyDeviceToHost, st), "gain");
  hip_check(hipStreamSynchronize(st), "eager results");
  // the whole pipeline as one graph, replayed
  out->graph_ms = 0.0;
  if (in.repeats > 0) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "capture");
    KernelClock* saved = g_kernel_clock;
    g_kernel_clock = nullptr;  // plain launches inside the graph
    fill();
    lu(bestP);
    rest();
    g_kernel_clock = saved;
    hip_check(hipStreamEndCapture(st, &graph), "capture");
    hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    hip_check(hipGraphLaunch(exec, st), "graph launch");  // warm
    hip_check(hipEventRecord(e4.a, st), "event");
    for (int r = 0; r < in.repeats; ++r) hip_check(hipGraphLaunch(exec, st), "graph launch");
    hip_check(hipEventRecord(e4.b, st), "event");
    hip_check(hipEventSynchronize(e4.b), "graph");
    out->graph_ms = e4.ms() / in.repeats;
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
  }
  // results
  std::vector<int> hinfo(S);
  std::vector<double> hzin(4 * (size_t)S);
  hip_check(hipMemcpyAsync(hinfo.data(), info, (size_t)S * sizeof(int), hipMemcpyDeviceToHost, st), "info");
  hip_check(hipMemcpyAsync(hzin.data(), zin, (size_t)S * 32, hipMemcpyDeviceToHost, st), "zin");
  out->gain.assign((size_t)S * std::max(0, in.n_dir), 0.0);
  if (in.n_dir > 0)
    hip_check(hipMemcpyAsync(out->gain.data(), gain, out->gain.size() * 8, hipMemcpyDeviceToHost, st), "gain");
  hip_check(hipStreamSynchronize(st), "batch results");
  out->zin.resize(2 * (size_t)S);
  out->pin.resize(S);
  for (int s = 0; s < S; ++s) {
    if (hinfo[s] != 0)
      throw NecError("batch: solve " + std::to_string(s) + " has a zero pivot at column " +
                     std::to_string(hinfo[s]));
    out->zin[2 * s] = hzin[4 * s];
    out->zin[2 * s + 1] = hzin[4 * s + 1];
    out->pin[s] = hzin[4 * s + 2];
  }
  if (in.repeats > 0)
    out->graph_matches_eager =
        (std::memcmp(eager_zin.data(), hzin.data(), hzin.size() * 8) == 0 &&
         (in.n_dir <= 0 || std::memcmp(eager_gain.data(), out->gain.data(), out->gain.size() * 8) == 0))
            ? 1
            : 0;
  out->solves = S;
  out->segments = rows;
}

}  // namespace necgpu
