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
#include <map>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <tuple>
#include <cmath>
#include <cstring>

#include "adjoint.h"
#include "engine.h"
#include "nec_math.h"

namespace necgpu {

namespace {
// batched-LU panel widths chosen by timing, per (largest order, solves, threads)
std::mutex g_tune_mu;
std::map<std::tuple<int, long long, int>, int> g_tuned_width;
}  // namespace

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


// Host tables of a batch: per (design, frequency) the segment table in
// wavelengths, the basis CSRs and the BatchSolve offsets, concatenated; the
// designs' geometries are kept for the adjoint.
struct BatchPrep {
  int S = 0, maxn = 0;
  std::vector<BatchSolve> sol;
  std::vector<Geometry> geo;  // per design
  std::vector<double> freq;   // per solve, MHz
  // per design, concatenated: what the device needs to build every solve's
  // tables (BatchDesign offsets)
  std::vector<BatchDesign> des;
  std::vector<double> dgeo;
  std::vector<int> dicon, dcolptr, dcolj, dtptr, dcpos, dsegptr, dsegbasis, dbptr, dspos;
  // sizes of the per-solve arrays the device fills
  long long rows = 0, zents = 0, nseg = 0, ncptr = 0, ncsr = 0, nsptr = 0, nssr = 0;
  int fill_rows = 1;          // rows per fill workgroup
  std::vector<int> fill_blk;  // (solve, first row) per fill workgroup
};

// body(s) for s = 0 .. count-1 on the host's cores; the first exception is
// rethrown after every worker has stopped.
template <class F>
void parallel_for(int count, F body) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nthreads = static_cast<int>(std::min<unsigned>(std::min(hw, 16u), (unsigned)std::max(1, count)));
  std::atomic<int> next(0);
  std::mutex mu;
  std::string err;
  auto work = [&]() {
    for (;;) {
      const int s = next.fetch_add(1);
      if (s >= count) return;
      try {
        body(s);
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu);
        if (err.empty()) err = e.what();
        next.store(count);
      }
    }
  };
  std::vector<std::thread> pool;
  for (int k = 1; k < nthreads; ++k) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (!err.empty()) throw NecError(err);
}

void prepare_batch(const BatchInput& in, BatchPrep* p) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    throw NecError("no HIP device available: the GPU path needs an MI355X (gfx950)");
  if (in.n_designs <= 0 || in.n_freq <= 0) throw NecError("batch: no designs or no frequencies");
  const int S = in.n_designs * in.n_freq;
  p->S = S;
  p->sol.assign(S, BatchSolve{});
  p->geo.assign(in.n_designs, Geometry{});
  p->freq.assign(S, 0.0);
  // 1. every design's geometry and the shape of its basis tables (which
  //    depends on the connections only, not on the frequency), on the
  //    host's cores
  std::vector<BasisShape> shape(in.n_designs);
  std::vector<int> src_of(in.n_designs);
  parallel_for(in.n_designs, [&](int d) {
    Geometry& geo = p->geo[d];
    for (int w = in.wire_ptr[d]; w < in.wire_ptr[d + 1]; ++w) {
      const double* gm = in.wire_geom + 7 * (size_t)w;
      geo.add_wire(in.wire_tag[w], in.wire_ns[w], gm[0], gm[1], gm[2], gm[3], gm[4], gm[5], gm[6], 1.0, 1.0);
    }
    geo.complete(0);
    const int n = static_cast<int>(geo.size());
    if (n < 1 || n > kBatchMaxN)
      throw NecError("batch: design " + std::to_string(d) + " has " + std::to_string(n) +
                     " segments; the batched path takes 1 to " + std::to_string(kBatchMaxN));
    src_of[d] = static_cast<int>(geo.segment_number(in.src_tag[d], in.src_seg[d]) - 1);
    build_basis_shape(geo, geo.scaled(in.freq_mhz[(size_t)d * in.n_freq]), &shape[d]);
  });
  // 2. the design arrays, concatenated
  p->des.resize(in.n_designs);
  for (int d = 0; d < in.n_designs; ++d) {
    const Geometry& geo = p->geo[d];
    const BasisShape& sh = shape[d];
    BatchDesign& D = p->des[d];
    D.n = static_cast<int>(geo.size());
    D.g0 = (long long)p->dgeo.size();
    for (const std::vector<double>* v8 : {&geo.x, &geo.y, &geo.z, &geo.len, &geo.rad, &geo.cab, &geo.sab, &geo.salp})
      p->dgeo.insert(p->dgeo.end(), v8->begin(), v8->end());
    D.ic0 = (long long)p->dicon.size();
    p->dicon.insert(p->dicon.end(), geo.icon1.begin(), geo.icon1.end());
    p->dicon.insert(p->dicon.end(), geo.icon2.begin(), geo.icon2.end());
    D.cptr0 = (long long)p->dcolptr.size();
    p->dcolptr.insert(p->dcolptr.end(), sh.colptr.begin(), sh.colptr.end());
    D.csr0 = (long long)p->dcolj.size();
    p->dcolj.insert(p->dcolj.end(), sh.colj.begin(), sh.colj.end());
    p->dcpos.insert(p->dcpos.end(), sh.cpos.begin(), sh.cpos.end());
    D.tptr0 = (long long)p->dtptr.size();
    p->dtptr.insert(p->dtptr.end(), sh.tptr.begin(), sh.tptr.end());
    D.sptr0 = (long long)p->dsegptr.size();
    p->dsegptr.insert(p->dsegptr.end(), sh.segptr.begin(), sh.segptr.end());
    D.ssr0 = (long long)p->dsegbasis.size();
    p->dsegbasis.insert(p->dsegbasis.end(), sh.segbasis.begin(), sh.segbasis.end());
    p->dspos.insert(p->dspos.end(), sh.spos.begin(), sh.spos.end());
    D.bptr0 = (long long)p->dbptr.size();
    p->dbptr.insert(p->dbptr.end(), sh.bptr.begin(), sh.bptr.end());
  }
  // 3. every solve: offsets, frequency scale and right-hand side
  for (int s = 0; s < S; ++s) {
    const int d = s / in.n_freq;
    const Geometry& geo = p->geo[d];
    BatchSolve& b = p->sol[s];
    const int n = static_cast<int>(geo.size());
    const double fmhz = in.freq_mhz[s];
    b.n = n;
    b.src = src_of[d];
    b.design = d;
    p->maxn = std::max(p->maxn, n);
    b.row0 = p->rows;
    b.seg0 = p->nseg;
    b.cptr0 = p->ncptr;
    b.csr0 = p->ncsr;
    b.sptr0 = p->nsptr;
    b.ssr0 = p->nssr;
    b.z0 = p->zents;
    p->rows += n;
    p->zents += (long long)n * n;
    p->nseg += 8LL * n;
    p->ncptr += n + 1;
    p->nsptr += n + 1;
    p->ncsr += (long long)shape[d].colj.size();
    p->nssr += (long long)shape[d].segbasis.size();
    p->freq[s] = fmhz;
    // as Geometry::scaled: wavelengths per metre, the source segment's length
    b.fr = (1.0e6 * fmhz) / em_speed_of_light();
    b.wavelength = em_speed_of_light() / (1.0e6 * fmhz);
    const zc v = zmk(in.src_volt[2 * d], in.src_volt[2 * d + 1]);
    // e[src] = -V / (s lambda), s the scaled segment length
    const zc e = zdivr(zneg(v), (geo.len[b.src] * b.fr) * b.wavelength);
    b.er = e.re;
    b.ei = e.im;
    b.vr = v.re;
    b.vi = v.im;
  }
  p->fill_rows = batch_fill_rows(p->maxn);
  for (int s = 0; s < S; ++s)
    for (int r0 = 0; r0 < p->sol[s].n; r0 += p->fill_rows) {
      p->fill_blk.push_back(s);
      p->fill_blk.push_back(r0);
    }
}

// Device copies of a batch's design tables, the per-solve tables the
// device builds from them, and the solve buffers.
struct BatchDev {
  DevBuf sol, des, dgeo, dicon, dcolptr, dcolj, dtptr, dcpos, dsegptr, dsegbasis, dbptr, dspos, fill_blk;
  DevBuf seg, colptr, colj, colabc, segptr, segbasis, segabc, status;
  DevBuf Z, rp, info, x, abc, zin;
  int fill_rows = 1, fill_nblk = 0, maxn = 0;
  void upload_all(const BatchPrep& p, hipStream_t st) {
    upload(sol, p.sol, st);
    upload(fill_blk, p.fill_blk, st);
    fill_rows = p.fill_rows;
    fill_nblk = static_cast<int>(p.fill_blk.size() / 2);
    maxn = p.maxn;
    upload(des, p.des, st);
    upload(dgeo, p.dgeo, st);
    upload(dicon, p.dicon, st);
    upload(dcolptr, p.dcolptr, st);
    upload(dcolj, p.dcolj, st);
    upload(dtptr, p.dtptr, st);
    upload(dcpos, p.dcpos, st);
    upload(dsegptr, p.dsegptr, st);
    upload(dsegbasis, p.dsegbasis, st);
    upload(dbptr, p.dbptr, st);
    upload(dspos, p.dspos, st);
    seg.ensure((size_t)p.nseg * 8);
    colptr.ensure((size_t)p.ncptr * 4);
    segptr.ensure((size_t)p.nsptr * 4);
    colj.ensure((size_t)std::max(1LL, p.ncsr) * 4);
    colabc.ensure((size_t)std::max(1LL, p.ncsr) * 24);
    segbasis.ensure((size_t)std::max(1LL, p.nssr) * 4);
    segabc.ensure((size_t)std::max(1LL, p.nssr) * 24);
    status.ensure(sizeof(int));
    Z.ensure((size_t)p.zents * 16);
    rp.ensure((size_t)p.rows * sizeof(int));
    info.ensure((size_t)p.S * sizeof(int));
    x.ensure((size_t)p.rows * 16);
    abc.ensure((size_t)p.rows * 48);
    zin.ensure((size_t)p.S * 32);
  }
  // every solve's segment table and basis tables, on the device
  void build_tables(int S, hipStream_t st) {
    hip_check(hipMemsetAsync(status.as<void>(), 0, sizeof(int), st), "memset");
    if (launch_batch_tables(sol.as<BatchSolve>(), S, des.as<BatchDesign>(), dgeo.as<double>(), dicon.as<int>(),
                            dcolptr.as<int>(), dcolj.as<int>(), dtptr.as<int>(), dcpos.as<int>(), dsegptr.as<int>(),
                            dsegbasis.as<int>(), dbptr.as<int>(), dspos.as<int>(), seg.as<double>(), colptr.as<int>(),
                            colj.as<int>(), colabc.as<double>(), segptr.as<int>(), segbasis.as<int>(),
                            segabc.as<double>(), status.as<int>(), st))
      hip_check(hipGetLastError(), "batch tables");
    int h = 0;
    hip_check(hipMemcpyAsync(&h, status.as<void>(), sizeof(int), hipMemcpyDeviceToHost, st), "status");
    hip_check(hipStreamSynchronize(st), "batch tables");
    if (h) throw NecError("batch: a basis table's shape changed with the frequency");
  }
};

// threads per batched workgroup for structures of up to maxn segments
int batch_threads(int maxn) { return maxn <= 256 ? 256 : 512; }

void check_info(const std::vector<int>& hinfo) {
  for (size_t s = 0; s < hinfo.size(); ++s)
    if (hinfo[s] != 0)
      throw NecError("batch: solve " + std::to_string(s) + " has a zero pivot at column " +
                     std::to_string(hinfo[s]));
}

}  // namespace

void run_batch(const BatchInput& in, BatchOutput* out) {
  const auto t_host0 = std::chrono::steady_clock::now();
  BatchPrep prep;
  prepare_batch(in, &prep);
  const int S = prep.S;
  const long long rows = prep.rows;
  const long long zents = prep.zents;
  const int bt = batch_threads(prep.maxn);
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
  BatchDev dv;
  dv.upload_all(prep, st);
  dv.build_tables(S, st);
  DevBuf d_dir, d_Zbak, d_gain;
  upload(d_dir, dir, st);
  double* Z = dv.Z.as<double>();
  int* rp = dv.rp.as<int>();
  int* info = dv.info.as<int>();
  double* x = dv.x.as<double>();
  double* abc = dv.abc.as<double>();
  double* zin = dv.zin.as<double>();
  double* gain = static_cast<double*>(d_gain.ensure(std::max<size_t>(8, (size_t)S * std::max(0, in.n_dir) * 8)));
  const BatchSolve* dsol = dv.sol.as<BatchSolve>();
  const double rkh = in.rkh > 0.0 ? in.rkh : 1.0;
  DevBuf& d_seg = dv.seg;
  DevBuf& d_colptr = dv.colptr;
  DevBuf& d_colj = dv.colj;
  DevBuf& d_colabc = dv.colabc;
  DevBuf& d_segptr = dv.segptr;
  DevBuf& d_segbasis = dv.segbasis;
  DevBuf& d_segabc = dv.segabc;

  auto fill = [&]() {
    if (launch_batch_fill(dsol, dv.fill_blk.as<int>(), dv.fill_nblk, dv.fill_rows, dv.maxn, d_seg.as<double>(),
                          d_colptr.as<int>(), d_colj.as<int>(), d_colabc.as<double>(), Z, rkh, st))
      hip_check(hipGetLastError(), "batch fill");
  };
  auto lu = [&](int P) {
    if (launch_batch_lu(dsol, S, Z, rp, info, P, bt, st)) hip_check(hipGetLastError(), "batch lu");
  };
  auto rest = [&]() {
    if (launch_batch_solve(dsol, S, Z, rp, d_seg.as<double>(), d_segptr.as<int>(), d_segbasis.as<int>(),
                           d_segabc.as<double>(), x, abc, zin, bt, st))
      hip_check(hipGetLastError(), "batch solve");
    if (in.n_dir > 0 &&
        launch_batch_pattern(dsol, S, d_seg.as<double>(), abc, zin, in.n_dir, d_dir.as<double>(), gain, bt, st))
      hip_check(hipGetLastError(), "batch pattern");
  };

  // autotune: the LU panel width, timed on this batch's own matrices
  EventPair e1, e2, e3, e4;
  hip_check(hipEventRecord(e1.a, st), "event");
  fill();
  hip_check(hipEventRecord(e1.b, st), "event");
  int bestP = in.panel_width;
  if (bt == 512 && bestP == 32) bestP = 16;
  // a width found for the same largest order and batch size is reused
  const std::tuple<int, long long, int> tune_key(dv.maxn, (long long)S, bt);
  if (bestP != 8 && bestP != 16 && bestP != 32) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    const auto it = g_tuned_width.find(tune_key);
    if (it != g_tuned_width.end()) bestP = it->second;
  }
  if (bestP != 8 && bestP != 16 && bestP != 32) {
    void* bak = d_Zbak.ensure((size_t)zents * 16);
    hip_check(hipMemcpyAsync(bak, Z, (size_t)zents * 16, hipMemcpyDeviceToDevice, st), "copy");
    float best = 1e30f;
    for (int P : {8, 16, 32}) {
      if (bt == 512 && P == 32) continue;  // the 512-row panel fits LDS at P <= 16
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
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned_width[tune_key] = bestP;
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
  check_info(hinfo);
  for (int s = 0; s < S; ++s) {
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
  (void)zents;
}

void run_batch_adjoint(const BatchInput& in, const AdjointSpec& spec, BatchAdjointOutput* out) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  BatchPrep prep;
  prepare_batch(in, &prep);
  const int S = prep.S;
  const long long rows = prep.rows;
  const int bt = batch_threads(prep.maxn);
  const double rkh = in.rkh > 0.0 ? in.rkh : 1.0;
  int P = in.panel_width;
  if (P != 8 && P != 16 && P != 32) P = 16;
  if (bt == 512 && P == 32) P = 16;
  out->host_prep_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();

  StreamGuard sg;
  hip_check(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking), "stream");
  const hipStream_t st = sg.s;
  BatchDev dv;
  dv.upload_all(prep, st);
  dv.build_tables(S, st);
  const BatchSolve* dsol = dv.sol.as<BatchSolve>();
  EventPair e1, e2, e3, e4;
  hip_check(hipEventRecord(e1.a, st), "event");
  if (launch_batch_fill(dsol, dv.fill_blk.as<int>(), dv.fill_nblk, dv.fill_rows, dv.maxn, dv.seg.as<double>(),
                        dv.colptr.as<int>(), dv.colj.as<int>(), dv.colabc.as<double>(), dv.Z.as<double>(), rkh, st))
    hip_check(hipGetLastError(), "batch fill");
  hip_check(hipEventRecord(e1.b, st), "event");
  hip_check(hipEventRecord(e2.a, st), "event");
  if (launch_batch_lu(dsol, S, dv.Z.as<double>(), dv.rp.as<int>(), dv.info.as<int>(), P, bt, st))
    hip_check(hipGetLastError(), "batch lu");
  hip_check(hipEventRecord(e2.b, st), "event");
  hip_check(hipEventRecord(e3.a, st), "event");
  if (launch_batch_solve(dsol, S, dv.Z.as<double>(), dv.rp.as<int>(), dv.seg.as<double>(), dv.segptr.as<int>(),
                         dv.segbasis.as<int>(), dv.segabc.as<double>(), dv.x.as<double>(), dv.abc.as<double>(),
                         dv.zin.as<double>(), bt, st))
    hip_check(hipGetLastError(), "batch solve");
  hip_check(hipEventRecord(e3.b, st), "event");
  std::vector<int> hinfo(S);
  std::vector<double> habc(6 * (size_t)rows);
  std::vector<zc> hx(rows);
  hip_check(hipMemcpyAsync(hinfo.data(), dv.info.as<void>(), (size_t)S * sizeof(int), hipMemcpyDeviceToHost, st),
            "info");
  hip_check(hipMemcpyAsync(habc.data(), dv.abc.as<void>(), habc.size() * 8, hipMemcpyDeviceToHost, st), "abc");
  hip_check(hipMemcpyAsync(hx.data(), dv.x.as<void>(), (size_t)rows * 16, hipMemcpyDeviceToHost, st), "x");
  hip_check(hipStreamSynchronize(st), "batch forward");
  check_info(hinfo);
  out->fill_ms = e1.ms();
  out->lu_ms = e2.ms();
  out->solve_ms = e3.ms();

  // host: objective and adjoint right-hand side of every solve
  const auto t1 = clk::now();
  std::vector<AdjointFront> front(S);
  std::vector<zc> g(rows);
  auto solve_design = [&](int s) { return s / in.n_freq; };
  parallel_for(S, [&](int s) {
    const BatchSolve& b = prep.sol[s];
    const Geometry& geo = prep.geo[solve_design(s)];
    const ScaledSegments sc = geo.scaled(prep.freq[s]);
    BasisTables t;
    build_basis_tables(geo, sc, &t);
    adjoint_front(sc, t, b.src, zmk(b.vr, b.vi), b.wavelength, habc.data() + 6 * b.row0, spec, &front[s]);
    std::copy(front[s].g.begin(), front[s].g.end(), g.begin() + b.row0);
  });
  out->front_ms = std::chrono::duration<double, std::milli>(clk::now() - t1).count();

  // device: transposed solves and the fill contraction
  DevBuf d_g, d_lam, d_gi, d_gj, d_ablk, d_afb, d_part;
  upload(d_g, g, st);
  d_lam.ensure((size_t)rows * 16);
  d_gi.ensure((size_t)rows * 7 * 8);
  d_gj.ensure((size_t)rows * 13 * 8);
  // contraction blocks: (solve, first row) and the first block of each solve
  std::vector<int> ablk, afb(S + 1, 0);
  const int arows = batch_adjoint_rows();
  for (int s = 0; s < S; ++s) {
    afb[s] = static_cast<int>(ablk.size() / 2);
    for (int r0 = 0; r0 < prep.sol[s].n; r0 += arows) {
      ablk.push_back(s);
      ablk.push_back(r0);
    }
  }
  const int anblk = static_cast<int>(ablk.size() / 2);
  afb[S] = anblk;
  upload(d_ablk, ablk, st);
  upload(d_afb, afb, st);
  d_part.ensure((size_t)anblk * 13 * prep.maxn * 8);
  hip_check(hipEventRecord(e4.a, st), "event");
  if (launch_batch_solve_transposed(dsol, S, dv.Z.as<double>(), dv.rp.as<int>(), d_g.as<double>(), d_lam.as<double>(),
                                    bt, st))
    hip_check(hipGetLastError(), "batch transposed solve");
  if (launch_batch_adjoint_fill(dsol, S, d_ablk.as<int>(), d_afb.as<int>(), anblk, prep.maxn, dv.seg.as<double>(),
                                d_lam.as<double>(), dv.abc.as<double>(), rkh, d_gi.as<double>(), d_part.as<double>(),
                                d_gj.as<double>(), st))
    hip_check(hipGetLastError(), "batch adjoint fill");
  hip_check(hipEventRecord(e4.b, st), "event");
  std::vector<zc> lam(rows);
  std::vector<double> gi(7 * (size_t)rows), gj(13 * (size_t)rows);
  hip_check(hipMemcpyAsync(lam.data(), d_lam.as<void>(), (size_t)rows * 16, hipMemcpyDeviceToHost, st), "lam");
  hip_check(hipMemcpyAsync(gi.data(), d_gi.as<void>(), gi.size() * 8, hipMemcpyDeviceToHost, st), "gi");
  hip_check(hipMemcpyAsync(gj.data(), d_gj.as<void>(), gj.size() * 8, hipMemcpyDeviceToHost, st), "gj");
  hip_check(hipStreamSynchronize(st), "batch adjoint");
  out->adjoint_device_ms = e4.ms();

  // host: gradients
  const auto t2 = clk::now();
  out->J.assign(S, 0.0);
  out->gain_db.assign(S, 0.0);
  out->zin.assign(2 * (size_t)S, 0.0);
  out->dfreq.assign(S, 0.0);
  out->grad.assign(7 * (size_t)rows, 0.0);
  parallel_for(S, [&](int s) {
    const BatchSolve& b = prep.sol[s];
    const Geometry& geo = prep.geo[solve_design(s)];
    const ScaledSegments sc = geo.scaled(prep.freq[s]);
    AdjointResult r;
    adjoint_back(geo, sc, b.src, zmk(b.vr, b.vi), b.wavelength, habc.data() + 6 * b.row0, spec, front[s],
                 lam.data() + b.row0, hx.data() + b.row0, gi.data() + 7 * b.row0, gj.data() + 13 * b.row0, &r);
    std::vector<double> g7;
    double df = 0.0;
    adjoint_endpoints(sc, b.wavelength, prep.freq[s], r, &g7, &df);
    std::copy(g7.begin(), g7.end(), out->grad.begin() + 7 * b.row0);
    out->J[s] = r.J;
    out->gain_db[s] = r.gain_db;
    out->zin[2 * s] = r.zin.re;
    out->zin[2 * s + 1] = r.zin.im;
    out->dfreq[s] = df;
  });
  out->back_ms = std::chrono::duration<double, std::milli>(clk::now() - t2).count();
  out->total_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  out->solves = S;
  out->segments = rows;
}

}  // namespace necgpu
