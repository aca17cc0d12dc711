// engine.cpp -- device-resident MoM solver for one structure.
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "basis.h"

namespace necgpu {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw NecError(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

namespace {

// NECGPU_FAIL_ALLOC=<bytes>: every device allocation of at least that size
// fails, so tests can drive the allocation-failure path from the host.
size_t forced_alloc_failure_threshold() {
  const char* v = std::getenv("NECGPU_FAIL_ALLOC");
  if (!v || !*v) return 0;
  return static_cast<size_t>(std::strtoull(v, nullptr, 10));
}

constexpr int kTableCap = 64;
constexpr size_t kFieldBudget = size_t(1) << 30;  // bytes of E staged per fill pass

// test_ek_approximation, src/c_geometry.cpp:3449-3458
int ek_test(const ScaledSegments& s, int s1, int s2) {
  const double ratio = s.b[s2] / s.b[s1];
  const double xi = std::fabs(s.cab[s1] * s.cab[s2] + s.sab[s1] * s.sab[s2] + s.salp[s1] * s.salp[s2]);
  return (xi < 0.999999 || std::fabs(ratio - 1.0) > 1.e-6) ? 2 : 0;
}

// End conditions of the extended kernel for source j, cmww
// src/nec_context.cpp:2535-2601.
void ek_ends(const Geometry& geo, const ScaledSegments& s, int j, int* ind1, int* ind2) {
  const int jx = j + 1;
  const double cabj = s.cab[j], sabj = s.sab[j];
  int ipr = geo.icon1[j];
  const int iprx = std::abs(ipr) - 1;
  *ind1 = 2;
  if (ipr > kPchcon) *ind1 = 0;
  else if (ipr == 0) *ind1 = 1;
  else if (ipr == jx && cabj * cabj + sabj * sabj <= 1.e-8) *ind1 = 0;
  else if (ipr > 0 && geo.icon2[iprx] == jx) *ind1 = ek_test(s, j, ipr - 1);
  else if (ipr < 0 && geo.icon1[iprx] == -jx) *ind1 = ek_test(s, j, -ipr - 1);
  ipr = geo.icon2[j];
  if (ipr > kPchcon) *ind2 = 2;
  else if (ipr == 0) *ind2 = 1;
  else if (ipr < 0) {
    const int k = -ipr - 1;
    *ind2 = (-geo.icon2[k] != jx) ? 2 : ek_test(s, j, k);
  } else if (ipr != jx) {
    const int k = ipr - 1;
    *ind2 = (geo.icon1[k] != jx) ? 2 : ek_test(s, j, k);
  } else {
    *ind2 = (cabj * cabj + sabj * sabj > 1.e-8) ? 2 : 0;
  }
}

template <class T>
void upload(DevBuf& buf, const std::vector<T>& v, hipStream_t s) {
  void* p = buf.ensure(std::max<size_t>(v.size() * sizeof(T), 1));
  if (!v.empty()) hip_check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}

}  // namespace

void build_basis_tables(const Geometry& geo, const ScaledSegments& seg, BasisTables* out) {
  const int n = static_cast<int>(geo.size());
  // Basis tables.  trio(j) gives the bases with support on source j; the
  // fill gathers per column k in ascending j.  tbf(i) gives the support of
  // basis i; the current coefficients gather per segment in ascending i.
  std::vector<int> icon1(geo.icon1.begin(), geo.icon1.end()), icon2(geo.icon2.begin(), geo.icon2.end());
  SegView sv{icon1.data(), icon2.data(), seg.s.data(), seg.b.data()};
  std::vector<int> ccount(n + 1, 0), scount(n + 1, 0);
  std::vector<int> tj, tk, sj, si;
  std::vector<double> tv, sv3;
  tj.reserve(8 * (size_t)n);
  int jco[kTableCap];
  double ax[kTableCap], bx[kTableCap], cx[kTableCap];
  for (int j = 0; j < n; ++j) {
    const int cnt = trio(sv, j + 1, jco, ax, bx, cx, kTableCap);
    if (cnt < 0) throw NecError("TRIO - SEGMENT CONNECTION ERROR FOR SEGMENT " + std::to_string(j + 1));
    for (int q = 0; q < cnt; ++q) {
      tj.push_back(j);
      tk.push_back(jco[q] - 1);
      tv.push_back(ax[q]);
      tv.push_back(bx[q]);
      tv.push_back(cx[q]);
      ++ccount[jco[q]];
    }
  }
  for (int i = 0; i < n; ++i) {
    const int cnt = tbf(sv, i + 1, 1, jco, ax, bx, cx, kTableCap);
    if (cnt < 0) throw NecError("TBF - SEGMENT CONNECTION ERROR FOR SEGMENT " + std::to_string(i + 1));
    for (int q = 0; q < cnt; ++q) {
      si.push_back(i);
      sj.push_back(jco[q] - 1);
      sv3.push_back(ax[q]);
      sv3.push_back(bx[q]);
      sv3.push_back(cx[q]);
      ++scount[jco[q]];
    }
  }
  // counting sort keeps the source order (ascending j, resp. ascending i)
  std::vector<int>& cptr = out->colptr;
  std::vector<int>& sptr = out->segptr;
  cptr.assign(n + 1, 0);
  sptr.assign(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    cptr[k + 1] = cptr[k] + ccount[k + 1];
    sptr[k + 1] = sptr[k] + scount[k + 1];
  }
  std::vector<int>& cj = out->colj;
  std::vector<int>& sb = out->segbasis;
  std::vector<double>& cv = out->colabc;
  std::vector<double>& sval = out->segabc;
  cj.assign(tj.size(), 0);
  sb.assign(si.size(), 0);
  cv.assign(tv.size(), 0.0);
  sval.assign(sv3.size(), 0.0);
  std::vector<int> cfill(cptr.begin(), cptr.end() - 1), sfill(sptr.begin(), sptr.end() - 1);
  for (size_t e = 0; e < tj.size(); ++e) {
    const int p = cfill[tk[e]]++;
    cj[p] = tj[e];
    cv[3 * p] = tv[3 * e];
    cv[3 * p + 1] = tv[3 * e + 1];
    cv[3 * p + 2] = tv[3 * e + 2];
  }
  for (size_t e = 0; e < si.size(); ++e) {
    const int p = sfill[sj[e]]++;
    sb[p] = si[e];
    sval[3 * p] = sv3[3 * e];
    sval[3 * p + 1] = sv3[3 * e + 1];
    sval[3 * p + 2] = sv3[3 * e + 2];
  }
}

void build_basis_shape(const Geometry& geo, const ScaledSegments& seg, BasisShape* out) {
  // The index side of build_basis_tables: the same trio / tbf walks and
  // counting sorts, keeping for every entry where its coefficients go.
  const int n = static_cast<int>(geo.size());
  std::vector<int> icon1(geo.icon1.begin(), geo.icon1.end()), icon2(geo.icon2.begin(), geo.icon2.end());
  SegView sv{icon1.data(), icon2.data(), seg.s.data(), seg.b.data()};
  int jco[kTableCap];
  double ax[kTableCap], bx[kTableCap], cx[kTableCap];
  std::vector<int> tk, sj;
  out->tptr.assign(n + 1, 0);
  out->bptr.assign(n + 1, 0);
  std::vector<int> ccount(n + 1, 0), scount(n + 1, 0);
  for (int j = 0; j < n; ++j) {
    const int cnt = trio(sv, j + 1, jco, ax, bx, cx, kTableCap);
    if (cnt < 0) throw NecError("TRIO - SEGMENT CONNECTION ERROR FOR SEGMENT " + std::to_string(j + 1));
    for (int q = 0; q < cnt; ++q) {
      tk.push_back(jco[q] - 1);
      ++ccount[jco[q]];
    }
    out->tptr[j + 1] = out->tptr[j] + cnt;
  }
  for (int i = 0; i < n; ++i) {
    const int cnt = tbf(sv, i + 1, 1, jco, ax, bx, cx, kTableCap);
    if (cnt < 0) throw NecError("TBF - SEGMENT CONNECTION ERROR FOR SEGMENT " + std::to_string(i + 1));
    for (int q = 0; q < cnt; ++q) {
      sj.push_back(jco[q] - 1);
      ++scount[jco[q]];
    }
    out->bptr[i + 1] = out->bptr[i] + cnt;
  }
  out->colptr.assign(n + 1, 0);
  out->segptr.assign(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    out->colptr[k + 1] = out->colptr[k] + ccount[k + 1];
    out->segptr[k + 1] = out->segptr[k] + scount[k + 1];
  }
  out->colj.assign(tk.size(), 0);
  out->cpos.assign(tk.size(), 0);
  out->segbasis.assign(sj.size(), 0);
  out->spos.assign(sj.size(), 0);
  std::vector<int> cfill(out->colptr.begin(), out->colptr.end() - 1), sfill(out->segptr.begin(), out->segptr.end() - 1);
  for (int j = 0; j < n; ++j)
    for (int e = out->tptr[j]; e < out->tptr[j + 1]; ++e) {
      const int p = cfill[tk[e]]++;
      out->colj[p] = j;
      out->cpos[e] = p;
    }
  for (int i = 0; i < n; ++i)
    for (int e = out->bptr[i]; e < out->bptr[i + 1]; ++e) {
      const int p = sfill[sj[e]]++;
      out->segbasis[p] = i;
      out->spos[e] = p;
    }
}

DevBuf::~DevBuf() { release(); }

void DevBuf::release() {
  if (ptr_) (void)hipFree(ptr_);
  ptr_ = nullptr;
  bytes_ = 0;
}

void* DevBuf::ensure(size_t bytes) {
  if (bytes <= bytes_ && ptr_) return ptr_;
  release();
  const size_t fail_at = forced_alloc_failure_threshold();
  if (fail_at && bytes >= fail_at)
    throw NecError("HIP error in hipMalloc: out of memory (forced by NECGPU_FAIL_ALLOC)");
  hip_check(hipMalloc(&ptr_, bytes ? bytes : 1), "hipMalloc");
  bytes_ = bytes;
  return ptr_;
}

DeviceSolver::DeviceSolver() {}

DeviceSolver::~DeviceSolver() {
  if (stream_) {
    (void)hipStreamSynchronize(stream_);
    for (auto& e : ev_)
      if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(stream_);
  }
}

void DeviceSolver::ensure_device() {
  if (stream_) return;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    throw NecError("no HIP device available: the GPU path needs an MI355X (gfx950)");
  hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  for (auto& e : ev_) hip_check(hipEventCreate(&e), "hipEventCreate");
}

float DeviceSolver::elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
  return ms;
}

DevSegments DeviceSolver::dev_segments() const {
  const double* p = seg_.as<double>();
  DevSegments g;
  g.x = p;
  g.y = p + n_;
  g.z = p + 2 * n_;
  g.s = p + 3 * n_;
  g.b = p + 4 * n_;
  g.cab = p + 5 * n_;
  g.sab = p + 6 * n_;
  g.salp = p + 7 * n_;
  return g;
}

void DeviceSolver::load(const Geometry& geo, const ScaledSegments& seg, bool use_ek) {
  ensure_device();
  const int n = static_cast<int>(geo.size());
  n_ = n;
  ld_ = n;
  use_ek_ = use_ek;
  factored_ = false;

  std::vector<double> segv(8 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    segv[i] = seg.x[i];
    segv[n + i] = seg.y[i];
    segv[2 * n + i] = seg.z[i];
    segv[3 * n + i] = seg.s[i];
    segv[4 * n + i] = seg.b[i];
    segv[5 * n + i] = seg.cab[i];
    segv[6 * n + i] = seg.sab[i];
    segv[7 * n + i] = seg.salp[i];
  }

  BasisTables t;
  build_basis_tables(geo, seg, &t);
  std::vector<int> ek(2 * (size_t)n, 2);
  if (use_ek)
    for (int j = 0; j < n; ++j) ek_ends(geo, seg, j, &ek[2 * j], &ek[2 * j + 1]);

  // The fill's field order: segments sorted along a Morton (Z-order) curve
  // of their centres, so that the bases a gather wave handles, and the
  // sources their supports share at junctions, sit close together in E.
  // NECGPU_FILL_ORDER=0 keeps the natural order (for measurements).
  std::vector<int> ord(n), pos(n), cjp(t.colj.size());
  for (int i = 0; i < n; ++i) ord[i] = i;
  const char* fo = std::getenv("NECGPU_FILL_ORDER");
  if (!(fo && fo[0] == '0') && n > 1) {
    double lo[3] = {seg.x[0], seg.y[0], seg.z[0]}, hi[3] = {lo[0], lo[1], lo[2]};
    for (int i = 0; i < n; ++i) {
      const double c[3] = {seg.x[i], seg.y[i], seg.z[i]};
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], c[a]);
        hi[a] = std::max(hi[a], c[a]);
      }
    }
    const double span = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    std::vector<uint64_t> key(n, 0);
    for (int i = 0; i < n; ++i) {
      const double c[3] = {seg.x[i], seg.y[i], seg.z[i]};
      uint64_t q[3];
      for (int a = 0; a < 3; ++a) {
        const double u = span > 0.0 ? (c[a] - lo[a]) / span : 0.0;
        q[a] = (uint64_t)std::min(2097151.0, std::max(0.0, u * 2097151.0));
      }
      uint64_t m = 0;
      for (int bit = 20; bit >= 0; --bit)
        for (int a = 0; a < 3; ++a) m = (m << 1) | ((q[a] >> bit) & 1u);
      key[i] = m;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
  }
  for (int p = 0; p < n; ++p) pos[ord[p]] = p;
  for (size_t q = 0; q < t.colj.size(); ++q) cjp[q] = pos[t.colj[q]];

  // Clusters of the one-pass fill: runs of T bases of the Morton order (a
  // compact patch of the structure), each with the sorted list of every
  // source its bases' supports draw on.  T = 512 up to 8,192 segments, else
  // 1,024 (NECGPU_FILL_TILE overrides it, for measurements): on the planar
  // grids a source list is then 1.11-1.2 times the cluster, the fields
  // evaluated twice at the clusters' edges.  Measured fill times, T = 128 ..
  // 2,048: 3,280 segments 1.10, 1.02, 0.91, 1.16, 1.20 ms; 16,020 segments
  // (T = 256 .. 2,048) 17.4, 21.9, 16.4, 28.7 ms.
  {
    const char* tv = std::getenv("NECGPU_FILL_TILE");
    const int T = std::max(8, (tv && *tv) ? std::atoi(tv) : (n <= 8192 ? 512 : 1024));
    std::vector<int> srcp{0}, src, basep{0}, base, cjl(t.colj.size());
    int smax = 0;
    for (int p0 = 0; p0 < n; p0 += T) {
      const int p1 = std::min(n, p0 + T);
      std::vector<int> bases(ord.begin() + p0, ord.begin() + p1);
      std::sort(bases.begin(), bases.end());
      std::vector<int> s;
      for (int k : bases)
        for (int q = t.colptr[k]; q < t.colptr[k + 1]; ++q) s.push_back(t.colj[q]);
      std::sort(s.begin(), s.end());
      s.erase(std::unique(s.begin(), s.end()), s.end());
      for (int k : bases)
        for (int q = t.colptr[k]; q < t.colptr[k + 1]; ++q)
          cjl[q] = static_cast<int>(std::lower_bound(s.begin(), s.end(), t.colj[q]) - s.begin());
      smax = std::max(smax, static_cast<int>(s.size()));
      src.insert(src.end(), s.begin(), s.end());
      srcp.push_back(static_cast<int>(src.size()));
      base.insert(base.end(), bases.begin(), bases.end());
      basep.push_back(static_cast<int>(base.size()));
    }
    fill_clusters_ = static_cast<int>(basep.size()) - 1;
    fill_smax_ = smax;
    fill_rb_ = 0;
    for (int rb : {4, 2, 1})
      if ((size_t)rb * smax * 48 <= kFillTileLdsMax) {
        fill_rb_ = rb;
        break;
      }
    upload(cl_src_ptr_, srcp, stream_);
    upload(cl_src_, src, stream_);
    upload(cl_base_ptr_, basep, stream_);
    upload(cl_base_, base, stream_);
    upload(cjl_, cjl, stream_);
  }

  upload(seg_, segv, stream_);
  upload(ford_, ord, stream_);
  upload(fcjp_, cjp, stream_);
  upload(colptr_, t.colptr, stream_);
  upload(colj_, t.colj, stream_);
  upload(colabc_, t.colabc, stream_);
  upload(segptr_, t.segptr, stream_);
  upload(segbasis_, t.segbasis, stream_);
  upload(segabc_, t.segabc, stream_);
  upload(ekind_, ek, stream_);
  hip_check(hipStreamSynchronize(stream_), "load");
}

void DeviceSolver::set_loads(const std::vector<zc>& z) {
  has_loads_ = !z.empty();
  if (!has_loads_) return;
  if ((int)z.size() != n_) throw NecError("loads: one impedance per segment expected");
  std::vector<double> v(2 * z.size());
  for (size_t i = 0; i < z.size(); ++i) {
    v[2 * i] = z[i].re;
    v[2 * i + 1] = z[i].im;
  }
  upload(zload_dev_, v, stream_);
  hip_check(hipStreamSynchronize(stream_), "loads");  // v is released on return
}

std::vector<zc> DeviceSolver::near_field(bool magnetic, const std::vector<double>& pts,
                                         const std::vector<double>& ax, double rkh) {
  const int np = static_cast<int>(ax.size());
  std::vector<zc> out;
  if (np == 0) return out;
  DevBuf dpts, dax, dpart, dout;
  upload(dpts, pts, stream_);
  upload(dax, ax, stream_);
  dpart.ensure((size_t)np * ((n_ + 255) / 256) * 6 * sizeof(double));
  dout.ensure((size_t)np * 6 * sizeof(double));
  if (launch_near_field(magnetic, dev_segments(), n_, dpts.as<double>(), dax.as<double>(), np, abc_.as<double>(),
                        use_ek_ ? 1 : 0, ekind_.as<int>(), rkh, ground_, dpart.as<double>(), dout.as<double>(),
                        stream_))
    hip_check(hipGetLastError(), "near field launch");
  std::vector<double> h((size_t)np * 6);
  hip_check(hipMemcpyAsync(h.data(), dout.as<void>(), h.size() * sizeof(double), hipMemcpyDeviceToHost, stream_),
            "near field");
  hip_check(hipStreamSynchronize(stream_), "near field");
  out.resize((size_t)np * 3);
  for (size_t k = 0; k < out.size(); ++k) out[k] = zmk(h[2 * k], h[2 * k + 1]);
  return out;
}

void DeviceSolver::enqueue_fill(double rkh) {
  const int n = n_;
  last_rkh_ = rkh;
  Z_.ensure((size_t)ld_ * n * 16);
  // one pass with the fields in LDS (NECGPU_FILL_FUSED=0 keeps the field
  // buffer and the gather, for measurements)
  const char* fv = std::getenv("NECGPU_FILL_FUSED");
  if (fill_rb_ > 0 && !(fv && fv[0] == '0')) {
    if (launch_fill_tile(dev_segments(), n, rkh, use_ek_ ? 1 : 0, ekind_.as<int>(), ground_,
                         has_loads_ ? zload_dev_.as<double>() : nullptr, fill_clusters_, cl_src_ptr_.as<int>(),
                         cl_src_.as<int>(), cl_base_ptr_.as<int>(), cl_base_.as<int>(), colptr_.as<int>(),
                         colj_.as<int>(), cjl_.as<int>(), colabc_.as<double>(), fill_smax_, fill_rb_, Z_.as<double>(),
                         ld_, stream_))
      hip_check(hipGetLastError(), "fill launch");
    return;
  }
  const int rows = static_cast<int>(std::min<size_t>(n, std::max<size_t>(1, kFieldBudget / ((size_t)n * 48))));
  E_.ensure((size_t)rows * n * 48);
  const DevSegments g = dev_segments();
  for (int i0 = 0; i0 < n; i0 += rows) {
    const int r = std::min(rows, n - i0);
    if (launch_fill(g, n, i0, r, rkh, use_ek_ ? 1 : 0, ekind_.as<int>(), ground_,
                    has_loads_ ? zload_dev_.as<double>() : nullptr, E_.as<double>(), colptr_.as<int>(),
                    colj_.as<int>(), fcjp_.as<int>(), ford_.as<int>(), colabc_.as<double>(), Z_.as<double>(), ld_,
                    stream_))
      hip_check(hipGetLastError(), "fill launch");
  }
}

void DeviceSolver::enqueue_factor() {
  const int n = n_;
  piv_.ensure((size_t)2 * n * sizeof(int));
  info_.ensure(sizeof(int));
  lu_work_.ensure(lu_workspace_bytes(n));
  hip_check(hipMemsetAsync(info_.as<void>(), 0, sizeof(int), stream_), "memset");
  if (precision_ == 1) {
    // mixed: factor an FP32 copy (block-local pivots first); Z keeps the
    // FP64 matrix for refinement
    const char* v = std::getenv("NECGPU_MIXED_PIVOT");
    enqueue_factor_f32((v && v[0] == '0') ? 0 : 1);
    return;
  }
  const int rc = launch_lu_fp64(Z_.as<double>(), ld_, n, piv_.as<int>(), info_.as<int>(), lu_work_.as<void>(), stream_,
                                lookahead_);
  if (rc == 2) throw NecError("LU: matrix order " + std::to_string(n) + " exceeds the panel exchange area");
  if (rc) hip_check(hipGetLastError(), "lu launch");
  ensure_solve_work();
  if (launch_lu_solve_prepare(Z_.as<double>(), ld_, n, solve_work_.as<void>(), stream_))
    hip_check(hipGetLastError(), "solve prepare launch");
}

void DeviceSolver::enqueue_factor_f32(int pivot_mode) {
  const int n = n_;
  mixed_pivot_ = pivot_mode;
  hip_check(hipMemsetAsync(info_.as<void>(), 0, sizeof(int), stream_), "memset");
  float* z32 = static_cast<float*>(Z32_.ensure((size_t)n * n * 8));
  if (launch_to_f32(Z_.as<double>(), ld_, z32, n, n, stream_)) hip_check(hipGetLastError(), "copy launch");
  const int rc =
      launch_lu_fp32(z32, n, n, piv_.as<int>(), info_.as<int>(), lu_work_.as<void>(), stream_, pivot_mode, lookahead_);
  if (rc == 2) throw NecError("LU: matrix order " + std::to_string(n) + " exceeds the panel exchange area");
  if (rc) hip_check(hipGetLastError(), "lu launch");
  ensure_solve_work();
  if (launch_lu_solve_prepare_f32(z32, n, n, solve_work_.as<void>(), stream_))
    hip_check(hipGetLastError(), "solve prepare launch");
}

void DeviceSolver::refine_solve() {
  // x_ holds b on entry.  x = LU32 \ b, then x += LU32 \ (b - Z x) until
  // |b - Z x| <= refine_tol_ |b|.  The norms come back to the host after
  // every residual, so this stage synchronises once per step.
  const int n = n_;
  double* b = static_cast<double*>(b_.ensure((size_t)n * 16));
  double* r = static_cast<double*>(r_.ensure((size_t)n * 16));
  double* acc = static_cast<double*>(acc_.ensure(64));
  const float* z32 = Z32_.as<float>();
  hip_check(hipMemcpyAsync(b, x_.as<void>(), (size_t)n * 16, hipMemcpyDeviceToDevice, stream_), "rhs copy");
  if (launch_lu_solve_apply_f32(z32, n, n, piv_.as<int>(), x_.as<double>(), solve_work_.as<void>(), stream_))
    hip_check(hipGetLastError(), "solve launch");
  double norms[2] = {0.0, 0.0};
  refine_steps_ = 0;
  double prev = 0.0;
  for (int it = 0;; ++it) {
    if (launch_residual(Z_.as<double>(), ld_, n, x_.as<double>(), b, r, acc, it == 0, stream_))
      hip_check(hipGetLastError(), "residual launch");
    hip_check(hipMemcpyAsync(norms, acc, it == 0 ? 16 : 8, hipMemcpyDeviceToHost, stream_), "norms");
    hip_check(hipStreamSynchronize(stream_), "refinement");
    const double bn = std::sqrt(norms[1]);
    refine_residual_ = bn > 0.0 ? std::sqrt(norms[0]) / bn : std::sqrt(norms[0]);
    if (!(refine_residual_ > refine_tol_)) break;  // converged (or no right-hand side)
    // Block-local pivots that converge slowly (or not at all): factor again
    // with partial pivoting and start over from b.
    if (mixed_pivot_ == 1 && (it >= kMaxLocalRefine || (it >= 2 && !(refine_residual_ < 0.5 * prev)))) {
      ++refine_fallbacks_;
      enqueue_factor_f32(0);
      const int info = factor_outcome();
      if (info != 0)
        throw NecError("mixed-precision LU: zero pivot in column " + std::to_string(info) +
                       " of the partial-pivot refactorisation");
      hip_check(hipMemcpyAsync(x_.as<void>(), b, (size_t)n * 16, hipMemcpyDeviceToDevice, stream_), "rhs copy");
      if (launch_lu_solve_apply_f32(z32, n, n, piv_.as<int>(), x_.as<double>(), solve_work_.as<void>(), stream_))
        hip_check(hipGetLastError(), "solve launch");
      it = -1;
      prev = 0.0;
      refine_steps_ = 0;
      continue;
    }
    prev = refine_residual_;
    if (it >= kMaxRefine)
      throw NecError("mixed-precision refinement did not reach the tolerance in " + std::to_string(kMaxRefine) +
                     " steps (relative residual " + std::to_string(refine_residual_) + ")");
    if (launch_lu_solve_apply_f32(z32, n, n, piv_.as<int>(), r, solve_work_.as<void>(), stream_))
      hip_check(hipGetLastError(), "solve launch");
    if (launch_update(x_.as<double>(), r, n, stream_)) hip_check(hipGetLastError(), "update launch");
    ++refine_steps_;
  }
}

void DeviceSolver::ensure_solve_work() {
  solve_work_.ensure(solve_workspace_bytes(n_));
}

void DeviceSolver::enqueue_solve(const void* rhs_dev, hipMemcpyKind kind, const void* rhs_host) {
  const int n = n_;
  x_.ensure((size_t)n * 16);
  abc_.ensure((size_t)6 * n * sizeof(double));
  curx_.ensure((size_t)n * 16);
  ensure_solve_work();
  hip_check(hipMemcpyAsync(x_.as<void>(), rhs_dev ? rhs_dev : rhs_host, (size_t)n * 16, kind, stream_), "rhs");
  if (precision_ == 1) {
    refine_solve();
  } else {
    if (launch_lu_solve_apply(Z_.as<double>(), ld_, n, piv_.as<int>(), x_.as<double>(), solve_work_.as<void>(),
                              stream_))
      hip_check(hipGetLastError(), "solve launch");
  }
  if (launch_currents(n, x_.as<double>(), segptr_.as<int>(), segbasis_.as<int>(), segabc_.as<double>(),
                      abc_.as<double>(), curx_.as<double>(), stream_))
    hip_check(hipGetLastError(), "currents launch");
}

PatternOut DeviceSolver::pattern_buffers(int nd) {
  const size_t per = (size_t)nd;
  double* base = static_cast<double*>(pat_out_.ensure(per * 15 * sizeof(double)));
  int* sense = static_cast<int*>(pat_sense_.ensure(per * sizeof(int)));
  PatternOut po;
  po.gain = base;
  po.pg_vert = base + per;
  po.pg_horiz = base + 2 * per;
  po.pg_tot = base + 3 * per;
  po.pg_rhcp = base + 4 * per;
  po.pg_lhcp = base + 5 * per;
  po.axial = base + 6 * per;
  po.tilt = base + 7 * per;
  po.avg_scale = base + 8 * per;
  po.pint = base + 9 * per;
  po.e_theta = base + 10 * per;
  po.e_phi = base + 12 * per;
  po.sense = sense;
  return po;
}

void DeviceSolver::enqueue_pattern(const PatternSpec& spec, const PatternOut& po) {
  const int nd = spec.n_theta * spec.n_phi;
  pat_part_.ensure((size_t)pattern_slices(n_, nd) * nd * (spec.ground ? 12 : 6) * sizeof(double));
  if (launch_pattern(dev_segments(), n_, abc_.as<double>(), spec, trig_.as<double>(), pat_part_.as<double>(), po,
                     stream_))
    hip_check(hipGetLastError(), "pattern launch");
}

void DeviceSolver::fill(double rkh) {
  hip_check(hipEventRecord(ev_[0], stream_), "event");
  enqueue_fill(rkh);
  hip_check(hipEventRecord(ev_[1], stream_), "event");
  hip_check(hipEventSynchronize(ev_[1]), "fill");
  times_.fill_ms = elapsed(ev_[0], ev_[1]);
  factored_ = false;
}

int DeviceSolver::factor_wait(bool* timed_out) {
  // waits for the factorisation on the stream; returns the zero-pivot word
  // (1 + first zero column)
  int info = 0, status = 0;
  hip_check(hipMemcpyAsync(&info, info_.as<int>(), sizeof(int), hipMemcpyDeviceToHost, stream_), "info");
  hip_check(hipMemcpyAsync(&status, lu_status(lu_work_.as<void>(), n_), sizeof(int), hipMemcpyDeviceToHost, stream_),
            "status");
  hip_check(hipStreamSynchronize(stream_), "factor");
  *timed_out = status != 0;
  return info;
}

int DeviceSolver::factor_outcome() {
  bool timed_out = false;
  const int info = factor_wait(&timed_out);
  if (timed_out) throw NecError("LU: a panel exchange timed out");
  return info;
}

void DeviceSolver::factor() {
  hip_check(hipEventRecord(ev_[0], stream_), "event");
  enqueue_factor();
  hip_check(hipEventRecord(ev_[1], stream_), "event");
  bool timed_out = false;
  int info = factor_wait(&timed_out);
  // NECGPU_TEST_PANEL_TIMEOUT=1 makes this factorisation count as timed out
  // (the kernels themselves ran normally), to exercise the fallback
  const char* force = std::getenv("NECGPU_TEST_PANEL_TIMEOUT");
  if (force && force[0] == '1') timed_out = true;
  if (timed_out) {
    // The cooperative panel's workgroups wait on each other; if they could
    // not all be resident beside the look-ahead update (a device shared with
    // other work), their bounded wait gave up and the factors are wrong.
    // Factor again with every launch on one stream, where a panel never
    // shares the device with an update; the FP64 path refills Z first (the
    // factorisation overwrote it).
    ++lu_fallbacks_;
    if (precision_ == 0) enqueue_fill(last_rkh_);
    lookahead_ = false;
    hip_check(hipEventRecord(ev_[0], stream_), "event");
    try {
      if (precision_ == 1)
        enqueue_factor_f32(mixed_pivot_);
      else
        enqueue_factor();
    } catch (...) {
      lookahead_ = true;
      throw;
    }
    lookahead_ = true;
    hip_check(hipEventRecord(ev_[1], stream_), "event");
    info = factor_outcome();
  }
  if (precision_ == 1 && mixed_pivot_ == 1 && info != 0) {
    // a zero pivot inside a diagonal block: partial pivoting instead
    ++refine_fallbacks_;
    enqueue_factor_f32(0);
    hip_check(hipEventRecord(ev_[1], stream_), "event");
    info = factor_outcome();
  }
  times_.factor_ms = elapsed(ev_[0], ev_[1]);
  lu_info_ = info;
  factored_ = true;
}

void DeviceSolver::solve(const std::vector<zc>& e, std::vector<zc>* curx, std::vector<double>* abc) {
  if (!factored_) throw NecError("solve requested before the matrix was factored");
  const int n = n_;
  hip_check(hipEventRecord(ev_[0], stream_), "event");
  enqueue_solve(nullptr, hipMemcpyHostToDevice, e.data());
  hip_check(hipEventRecord(ev_[1], stream_), "event");
  if (curx) {
    curx->resize(n);
    hip_check(hipMemcpyAsync(curx->data(), curx_.as<void>(), (size_t)n * 16, hipMemcpyDeviceToHost, stream_), "curx");
  }
  if (abc) {
    abc->resize(6 * (size_t)n);
    hip_check(hipMemcpyAsync(abc->data(), abc_.as<void>(), (size_t)6 * n * 8, hipMemcpyDeviceToHost, stream_), "abc");
  }
  int status = 0;
  hip_check(hipMemcpyAsync(&status, solve_status(solve_work_.as<void>(), n), sizeof(int), hipMemcpyDeviceToHost, stream_),
            "status");
  hip_check(hipStreamSynchronize(stream_), "solve");
  if (status) throw NecError("triangular solve: block wait timed out");
  times_.solve_ms = elapsed(ev_[0], ev_[1]);
}

void DeviceSolver::adjoint(double rkh, const std::vector<zc>& g, std::vector<zc>* lam, std::vector<zc>* x,
                           std::vector<double>* gi, std::vector<double>* gj) {
  if (!factored_) throw NecError("adjoint requested before the matrix was factored");
  if (precision_ != 0) throw NecError("the adjoint needs the FP64 factors (precision mode 0)");
  if (use_ek_) throw NecError("the adjoint supports the thin-wire kernel only (no EK card)");
  if (ground_.type != 0) throw NecError("the adjoint supports free space only (no GN card)");
  if (has_loads_) throw NecError("the adjoint supports unloaded structures only (no LD card)");
  const int n = n_;
  if ((int)g.size() != n) throw NecError("adjoint: right-hand side has the wrong length");
  adj_g_.ensure((size_t)n * 16);
  adj_lam_.ensure((size_t)n * 16);
  adj_acc_.ensure((size_t)n * 16);
  adj_gi_.ensure((size_t)n * 7 * 8);
  adj_gj_.ensure((size_t)n * 13 * 8);
  const int rows = adjoint_rows_per_chunk(n, kFieldBudget / 4);
  adj_pj_.ensure((size_t)rows * n * 13 * 8);
  hip_check(hipMemcpyAsync(adj_g_.as<void>(), g.data(), (size_t)n * 16, hipMemcpyHostToDevice, stream_), "adjoint g");
  hip_check(hipMemsetAsync(adj_gj_.as<void>(), 0, (size_t)n * 13 * 8, stream_), "adjoint zero");
  hip_check(hipEventRecord(ev_[0], stream_), "event");
  if (launch_lu_solve_transposed(Z_.as<double>(), ld_, n, piv_.as<int>(), adj_g_.as<double>(), adj_lam_.as<double>(),
                                 adj_acc_.as<double>(), stream_))
    hip_check(hipGetLastError(), "transposed solve launch");
  const DevSegments sg = dev_segments();
  for (int i0 = 0; i0 < n; i0 += rows) {
    const int r = std::min(rows, n - i0);
    if (launch_adjoint_fill(sg, n, i0, r, rkh, adj_lam_.as<double>(), abc_.as<double>(), adj_gi_.as<double>(),
                            adj_pj_.as<double>(), adj_gj_.as<double>(), stream_))
      hip_check(hipGetLastError(), "adjoint fill launch");
  }
  hip_check(hipEventRecord(ev_[1], stream_), "event");
  lam->resize(n);
  x->resize(n);
  gi->resize((size_t)n * 7);
  gj->resize((size_t)n * 13);
  hip_check(hipMemcpyAsync(lam->data(), adj_lam_.as<void>(), (size_t)n * 16, hipMemcpyDeviceToHost, stream_), "lam");
  hip_check(hipMemcpyAsync(x->data(), x_.as<void>(), (size_t)n * 16, hipMemcpyDeviceToHost, stream_), "x");
  hip_check(hipMemcpyAsync(gi->data(), adj_gi_.as<void>(), (size_t)n * 7 * 8, hipMemcpyDeviceToHost, stream_), "gi");
  hip_check(hipMemcpyAsync(gj->data(), adj_gj_.as<void>(), (size_t)n * 13 * 8, hipMemcpyDeviceToHost, stream_), "gj");
  hip_check(hipStreamSynchronize(stream_), "adjoint");
  adjoint_ms_ = elapsed(ev_[0], ev_[1]);
}

void DeviceSolver::pattern(const PatternSpec& spec, const std::vector<double>& trig, HostPattern* out) {
  const int nd = spec.n_theta * spec.n_phi;
  upload(trig_, trig, stream_);
  const PatternOut po = pattern_buffers(nd);
  hip_check(hipMemsetAsync(pat_out_.as<void>(), 0, (size_t)nd * 15 * sizeof(double), stream_), "memset");
  hip_check(hipEventRecord(ev_[0], stream_), "event");
  enqueue_pattern(spec, po);
  hip_check(hipEventRecord(ev_[1], stream_), "event");
  auto get = [&](std::vector<double>& v, const double* src, size_t cnt) {
    v.resize(cnt);
    hip_check(hipMemcpyAsync(v.data(), src, cnt * sizeof(double), hipMemcpyDeviceToHost, stream_), "pattern copy");
  };
  get(out->gain, po.gain, nd);
  get(out->pg_vert, po.pg_vert, nd);
  get(out->pg_horiz, po.pg_horiz, nd);
  get(out->pg_tot, po.pg_tot, nd);
  get(out->pg_rhcp, po.pg_rhcp, nd);
  get(out->pg_lhcp, po.pg_lhcp, nd);
  get(out->axial, po.axial, nd);
  get(out->tilt, po.tilt, nd);
  get(out->avg_scale, po.avg_scale, nd);
  get(out->pint, po.pint, nd);
  get(out->e_theta, po.e_theta, 2 * (size_t)nd);
  get(out->e_phi, po.e_phi, 2 * (size_t)nd);
  out->sense.resize(nd);
  hip_check(hipMemcpyAsync(out->sense.data(), po.sense, (size_t)nd * sizeof(int), hipMemcpyDeviceToHost, stream_),
            "pattern copy");
  hip_check(hipStreamSynchronize(stream_), "pattern");
  times_.pattern_ms = elapsed(ev_[0], ev_[1]);
}

void DeviceSolver::benchmark_pipeline(double rkh, const std::vector<zc>& e, const std::vector<int>& src_seg,
                                      const std::vector<zc>& src_v, double wavelength, const PatternSpec* spec,
                                      const std::vector<double>& trig, int warmup, int steps, double* ms_per_solve,
                                      StageTimes* stage_ms, std::vector<double>* zin_out) {
  const int n = n_;
  const int nsrc = static_cast<int>(src_seg.size());
  upload(rhs_, e, stream_);
  std::vector<double> srcv(2 * (size_t)nsrc);
  for (int s = 0; s < nsrc; ++s) {
    srcv[2 * s] = src_v[s].re;
    srcv[2 * s + 1] = src_v[s].im;
  }
  std::vector<int> srcs(src_seg);
  // voltages first (16-byte aligned), then the segment indices
  src_.ensure(std::max<size_t>(1, (size_t)nsrc * (sizeof(int) + 16)));
  if (nsrc) {
    hip_check(hipMemcpyAsync(src_.as<void>(), srcv.data(), nsrc * 16, hipMemcpyHostToDevice, stream_), "src");
    hip_check(hipMemcpyAsync(static_cast<char*>(src_.as<void>()) + nsrc * 16, srcs.data(), nsrc * sizeof(int),
                             hipMemcpyHostToDevice, stream_),
              "src");
  }
  const double* dvolt = src_.as<double>();
  const int* dseg = reinterpret_cast<const int*>(static_cast<char*>(src_.as<void>()) + nsrc * 16);
  zin_.ensure(std::max<size_t>(1, (size_t)nsrc * 64));
  PatternOut po{};
  if (spec) {
    upload(trig_, trig, stream_);
    po = pattern_buffers(spec->n_theta * spec->n_phi);
  }
  // Stage boundaries get their own events per timed step and every kernel
  // launch carries a start/stop pair (KernelClock), so nothing waits on the
  // host between steps.  The warmup steps run with the clock installed too,
  // so that the timed steps find all the events they need already created.
  struct ClockGuard {
    explicit ClockGuard(KernelClock* c) { g_kernel_clock = c; }
    ~ClockGuard() { g_kernel_clock = nullptr; }
  };
  ClockGuard guard(kernel_timing_ ? &kclock_ : nullptr);
  kclock_.reset();
  auto one_step = [&](hipEvent_t* stage) {
    if (stage) hip_check(hipEventRecord(stage[0], stream_), "event");
    enqueue_fill(rkh);
    if (stage) hip_check(hipEventRecord(stage[1], stream_), "event");
    enqueue_factor();
    if (stage) hip_check(hipEventRecord(stage[2], stream_), "event");
    enqueue_solve(rhs_.as<void>(), hipMemcpyDeviceToDevice, nullptr);
    if (launch_zin(nsrc, dseg, dvolt, curx_.as<double>(), wavelength, zin_.as<double>(), stream_))
      hip_check(hipGetLastError(), "zin launch");
    if (stage) hip_check(hipEventRecord(stage[3], stream_), "event");
    if (spec) enqueue_pattern(*spec, po);
  };
  for (int it = 0; it < warmup; ++it) one_step(nullptr);
  hip_check(hipStreamSynchronize(stream_), "warmup");
  const size_t per_step = kclock_.pending() / (size_t)std::max(1, warmup);
  kclock_.reset();
  if (!kclock_.reserve(per_step * (size_t)steps)) throw NecError("cannot create kernel timing events");
  std::vector<hipEvent_t> stage(4 * (size_t)std::max(0, steps));
  for (auto& e : stage) hip_check(hipEventCreate(&e), "hipEventCreate");
  struct EventsGuard {
    std::vector<hipEvent_t>& v;
    ~EventsGuard() {
      for (hipEvent_t e : v) (void)hipEventDestroy(e);
    }
  } events_guard{stage};
  hipEvent_t t0 = ev_[4], t1 = ev_[5];
  hip_check(hipEventRecord(t0, stream_), "event");
  for (int it = 0; it < steps; ++it) one_step(&stage[4 * (size_t)it]);
  hip_check(hipEventRecord(t1, stream_), "event");
  hip_check(hipEventSynchronize(t1), "event");
  const double total = elapsed(t0, t1);
  StageTimes acc;
  for (int it = 0; it < steps; ++it) {
    const hipEvent_t* s = &stage[4 * (size_t)it];
    acc.fill_ms += elapsed(s[0], s[1]);
    acc.factor_ms += elapsed(s[1], s[2]);
    acc.solve_ms += elapsed(s[2], s[3]);
  }
  if (kclock_.collect() != 0) throw NecError("kernel timing events failed: " + kclock_.failure);
  *ms_per_solve = steps > 0 ? total / steps : 0.0;
  if (stage_ms) {
    stage_ms->fill_ms = acc.fill_ms / std::max(1, steps);
    stage_ms->factor_ms = acc.factor_ms / std::max(1, steps);
    stage_ms->solve_ms = acc.solve_ms / std::max(1, steps);
    stage_ms->pattern_ms = *ms_per_solve - stage_ms->fill_ms - stage_ms->factor_ms - stage_ms->solve_ms;
  }
  if (zin_out) {
    zin_out->resize(8 * (size_t)nsrc);
    hip_check(hipMemcpy(zin_out->data(), zin_.as<void>(), (size_t)nsrc * 64, hipMemcpyDeviceToHost), "zin");
  }
  (void)n;
  factored_ = true;
}

void DeviceSolver::download_matrix(std::vector<double>* Z) {
  const size_t cnt = (size_t)2 * n_ * n_;
  Z->resize(cnt);
  if (ld_ == n_) {
    hip_check(hipMemcpyAsync(Z->data(), Z_.as<void>(), cnt * 8, hipMemcpyDeviceToHost, stream_), "matrix copy");
  } else {
    hip_check(hipMemcpy2DAsync(Z->data(), (size_t)n_ * 16, Z_.as<void>(), (size_t)ld_ * 16, (size_t)n_ * 16, n_,
                               hipMemcpyDeviceToHost, stream_),
              "matrix copy");
  }
  hip_check(hipStreamSynchronize(stream_), "matrix copy");
}

}  // namespace necgpu
