// kernels.h -- launch interface of the HIP kernels (host-callable).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace necgpu {

// ------------------------------------------------------------ kernel clock
// Kernel families whose launches can be timed one by one.
enum KernelKind {
  K_FILL_FIELDS,
  K_FILL_ASSEMBLE,
  K_LU_INIT,
  K_LU_PANEL,
  K_LU_LASWP,
  K_LU_TRSM,
  K_ZGEMM,
  K_SOLVE_RESET,
  K_DIAG_INVERSE,
  K_TRSV_LOWER,
  K_TRSV_UPPER,
  K_CURRENTS,
  K_ZIN,
  K_PATTERN_PARTIAL,
  K_PATTERN_FINISH,
  K_REFINE,
  K_RESIDUAL,
  K_BATCH_FILL,
  K_BATCH_LU,
  K_BATCH_SOLVE,
  K_BATCH_PATTERN,
  K_ADJOINT_FILL,
  K_ADJOINT_SOLVE,
  K_ADJOINT_REDUCE,
  K_LU_NEXT,
  K_NEAR_FIELD,
  K_BATCH_TABLES,
  K_FILL_TILE,
  K_NKINDS
};
const char* kernel_kind_name(int kind);

// While a KernelClock is installed on the calling thread, every launch made
// through launch_k carries a start/stop event pair (hipExtLaunchKernelGGL),
// so the events bracket exactly that kernel's execution on its stream.
class KernelClock {
 public:
  KernelClock() = default;
  KernelClock(const KernelClock&) = delete;
  KernelClock& operator=(const KernelClock&) = delete;
  ~KernelClock();
  bool next(int kind, hipEvent_t* start, hipEvent_t* stop);
  // After the timed work has completed: add every pending pair to the totals.
  int collect();
  std::string failure;  // the first failed pair of the last collect()
  void reset();
  size_t pending() const { return used_; }
  // pre-create event pairs so that later launches create none
  bool reserve(size_t slots);
  double total_ms[K_NKINDS] = {};
  long launches[K_NKINDS] = {};

 private:
  std::vector<hipEvent_t> ev_;  // 2 per slot
  std::vector<int> kind_;
  size_t used_ = 0;
};
extern thread_local KernelClock* g_kernel_clock;

template <typename F, typename... Args>
inline void launch_k(int kind, F kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t stream, Args... args) {
  hipEvent_t a, b;
  if (g_kernel_clock && g_kernel_clock->next(kind, &a, &b))
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, a, b, 0u, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, args...);
}

// Ground of the fill and the pattern (nec_ground): type 0 free space, 1
// perfect ground, 2 finite ground by reflection coefficients; zr = zrati
// (1 / sqrt of the complex relative permittivity), fr = frati (1 for both).
struct GroundSpec {
  int type = 0;
  double zr_re = 0.0, zr_im = 0.0;
  double fr_re = 1.0, fr_im = 0.0;
};

// Segment tables on the device, in wavelengths (structure of arrays).
struct DevSegments {
  const double *x, *y, *z, *s, *b, *cab, *sab, *salp;
};

// ------------------------------------------------------------ batches
// One solve of a batch of small structures: offsets into the batch's
// concatenated arrays (segment tables, basis CSRs, matrices, vectors).
struct BatchSolve {
  int n;            // segments
  int src;          // source segment (0-based)
  long long row0;   // sum of n over earlier solves (x, rp, abc offsets)
  long long seg0;   // into seg: 8 doubles per segment, SoA per solve
  long long cptr0;  // into colptr (n + 1 entries per solve)
  long long csr0;   // into colj (and 3x into colabc)
  long long sptr0;  // into segptr (n + 1 entries per solve)
  long long ssr0;   // into segbasis (and 3x into segabc)
  long long z0;     // into Z, complex entries (n x n row-major)
  double wavelength;
  double er, ei;    // right-hand side at the source: -V / (s lambda)
  double vr, vi;    // source voltage
  int design;       // index into the BatchDesign table
  double fr;        // frequency scale: wavelengths per metre
};
constexpr int kBatchMaxN = 512;  // largest structure of the batched path

// One design of a batch: what does not change with the frequency.  Offsets
// into the concatenated design arrays: geometry g0 (8 n doubles: x, y, z,
// length, radius in metres, then cab, sab, salp), connections ic0 (icon1 n,
// then icon2 n), the column CSR cptr0 (colptr n + 1) / csr0 (colj and cpos,
// nnz each), trio's entry starts tptr0 (n + 1), the segment CSR sptr0
// (segptr n + 1) / ssr0 (segbasis and spos, nnz each) and tbf's entry
// starts bptr0 (n + 1).  cpos / spos give, for the q-th entry of trio(j) /
// tbf(i), its position in the column / segment CSR.
struct BatchDesign {
  int n;
  long long g0, ic0, cptr0, csr0, tptr0, sptr0, ssr0, bptr0;
};

// batch.hip: every solve's tables on the device, one workgroup per solve:
// the segment table in wavelengths (seg, 8 n doubles at sol.seg0), the CSR
// index arrays copied from its design, and the basis coefficients (colabc,
// segabc) from trio / tbf at the solve's frequency.  status: set to 1 if a
// trio / tbf count differs from the design's (never for a valid design).
int launch_batch_tables(const BatchSolve* sol, int S, const BatchDesign* des, const double* dgeo, const int* dicon,
                        const int* dcolptr, const int* dcolj, const int* dtptr, const int* dcpos, const int* dsegptr,
                        const int* dsegbasis, const int* dbptr, const int* dspos, double* seg, int* colptr, int* colj,
                        double* colabc, int* segptr, int* segbasis, double* segabc, int* status, hipStream_t stream);

// fill.hip: one workgroup per block of R consecutive rows of one solve
// (R = batch_fill_rows(largest n)); blk holds (solve, first row) of each of
// the nblk blocks.  LDS: R * maxn * 48 bytes.
int batch_fill_rows(int maxn);
int launch_batch_fill(const BatchSolve* sol, const int* blk, int nblk, int R, int maxn, const double* seg,
                      const int* colptr, const int* colj, const double* colabc, double* Z, double rkh,
                      hipStream_t stream);

// batch.hip: one workgroup per solve, bt = 256 or 512 threads (bt >= the
// largest n).  LU with row permutation rp (logical row i of the factors is
// physical row rp[i]; no rows are moved), panel width P (8, 16 or 32; 8 or
// 16 for bt = 512); then the solve, current coefficients abc (6 per
// segment), Z_in and input power (zin: 4 per solve); then the total gain
// (dB) in n_dir directions (dir: sin th, cos th, sin ph, cos ph each).
int launch_batch_lu(const BatchSolve* sol, int S, double* Z, int* rp, int* info, int P, int bt, hipStream_t stream);
int launch_batch_solve(const BatchSolve* sol, int S, const double* Z, const int* rp, const double* seg,
                       const int* segptr, const int* segbasis, const double* segabc, double* x, double* abc,
                       double* zin, int bt, hipStream_t stream);
int launch_batch_pattern(const BatchSolve* sol, int S, const double* seg, const double* abc, const double* zin,
                         int n_dir, const double* dir, double* gain, int bt, hipStream_t stream);
// Batched adjoint.  batch.hip: Z^T lam = g per solve with the batched
// factors (g, lam: 2 per row).  fill.hip: the fill's adjoint contraction per
// solve over blocks of batch_adjoint_rows() observation rows (blk: solve and
// first row of each of the nblk blocks; fb: S + 1 entries, the first block
// of each solve), gi (7 per row: the observation side) and gj (13 per row:
// the source side and lam_i E_ij), both written; part: device scratch of
// nblk * 13 * maxn doubles.
int launch_batch_solve_transposed(const BatchSolve* sol, int S, const double* Z, const int* rp, const double* g,
                                  double* lam, int bt, hipStream_t stream);
int batch_adjoint_rows();
int launch_batch_adjoint_fill(const BatchSolve* sol, int S, const int* blk, const int* fb, int nblk, int maxn,
                              const double* seg, const double* lam, const double* abc, double rkh, double* gi,
                              double* part, double* gj, hipStream_t stream);

// fill.hip: fields of rows [i0, i0+rows) into E (rows*n*6 doubles), then
// gather into Z (row-major complex, leading dimension ldz complex entries).
// With a ground each pair adds the image term of efld.  zload (2 doubles per
// segment, or null): loads, subtracted on each loaded segment's row right
// after its own term (cmset's order).  ord: the field order (E's column t
// holds source ord[t]; the gather takes basis ord[t] on thread t); cjp: for
// every column-CSR entry, the position of its source in that order.
// Near electric (nefld; magnetic false) or magnetic (nhfld; magnetic true)
// field of wires, in free space or over the ground, at npts points (3
// doubles each, wavelengths); ax: per point the radius of the segment it
// lies on or 0; abc: the current coefficients (6 planes of n); part:
// npts * ceil(n/256) * 6 doubles of scratch; out: 6 doubles per point.
int launch_near_field(bool magnetic, const DevSegments& g, int n, const double* pts, const double* ax, int npts,
                      const double* abc, int use_ek, const int* ekind, double rkh, const GroundSpec& ground,
                      double* part, double* out, hipStream_t stream);

int launch_fill(const DevSegments& g, int n, int i0, int rows, double rkh, int use_ek, const int* ekind,
                const GroundSpec& ground, const double* zload, double* E, const int* colptr, const int* colj,
                const int* cjp, const int* ord, const double* colabc, double* Z, long long ldz, hipStream_t stream);

// The fill in one pass (no field buffer): workgroups of rb rows (1, 2 or 4)
// and one cluster of bases each.  Cluster q: sources cl_src[cl_src_ptr[q] ..
// cl_src_ptr[q+1]) (ascending), bases cl_base[cl_base_ptr[q] ..) (ascending);
// cjl[e]: position, in its basis' cluster source list, of the source of
// column-CSR entry e.  rb * smax * 48 bytes of LDS, at most kFillTileLdsMax.
constexpr size_t kFillTileLdsMax = 64 * 1024;
int launch_fill_tile(const DevSegments& g, int n, double rkh, int use_ek, const int* ekind, const GroundSpec& ground,
                     const double* zload, int nclusters, const int* cl_src_ptr, const int* cl_src,
                     const int* cl_base_ptr, const int* cl_base, const int* colptr, const int* colj, const int* cjl,
                     const double* colabc, int smax, int rb, double* Z, long long ldz, hipStream_t stream);

// lu.hip: in-place partial-pivot LU of the row-major complex n x n matrix A
// (leading dimension lda complex).  piv (2n ints): piv[k] = row exchanged
// with row k at step k, then the composed permutation.  info (device int)
// receives 1 + the first exactly-zero pivot, or stays 0.  work: device
// scratch of lu_workspace_bytes(n) (no initialisation needed); lu_status
// points at its status int, set to 1 if a panel exchange timed out.
// allow_lookahead false keeps every launch on `stream` (no panel runs beside
// a trailing update), the fallback after a timed-out exchange.
size_t lu_workspace_bytes(int n);
int* lu_status(void* work, int n);
int launch_lu_fp64(double* A, long long lda, int n, int* piv, int* info, void* work, hipStream_t stream,
                   bool allow_lookahead = true);
// The same LU with its workspace in a fixed device array of the library
// (lu_noalloc_scratch_bytes() per device, for n <= kNoAllocMaxN) and no side
// stream: no allocation of any kind.  Returns 3 for n > kNoAllocMaxN, 4 if
// the array's device address cannot be found, else as launch_lu_fp64.
constexpr int kNoAllocMaxN = 4096;
size_t lu_noalloc_scratch_bytes();
int launch_lu_fp64_noalloc(double* A, long long lda, int n, int* piv, int* info, hipStream_t stream);
// The same factorisation of an FP32 matrix (the mixed-precision path): FP32
// storage and trailing update on v_mfma_f32_16x16x4f32.  pivot_mode 0: the
// same pivot rule (partial pivoting); 1: pivots searched inside each panel's
// 64 x 64 diagonal block only (no grid-wide exchange; the factors of a
// block-permuted matrix, for preconditioning).  work: lu_workspace_bytes(n)
// is large enough for either.
int launch_lu_fp32(float* A, long long lda, int n, int* piv, int* info, void* work, hipStream_t stream,
                   int pivot_mode, bool allow_lookahead = true);
// one-wave check of the FP64 MFMA operand / result lane layout (64 a, 64 b in, 256 d out)
int launch_mfma_probe(const double* a, const double* b, double* d, hipStream_t stream);
int launch_mfma_probe_f32(const float* a, const float* b, float* d, hipStream_t stream);

// solve.hip: x <- A^-1 x with the LU factors of launch_lu_fp64 (or the FP32
// factors of the mixed path; x stays FP64).  work: device scratch of
// solve_workspace_bytes(n) bytes (no initialisation needed: every apply
// clears what it polls).  prepare (inverts the diagonal blocks) runs once
// per factorisation, before any apply.
size_t solve_workspace_bytes(int n);
int* solve_status(void* work, int n);
int launch_lu_solve_prepare(const double* LU, long long lda, int n, void* work, hipStream_t stream);
int launch_lu_solve_prepare_f32(const float* LU, long long lda, int n, void* work, hipStream_t stream);
int launch_lu_solve_apply(const double* LU, long long lda, int n, const int* piv, double* x, void* work,
                          hipStream_t stream);
int launch_lu_solve_apply_f32(const float* LU, long long lda, int n, const int* piv, double* x, void* work,
                              hipStream_t stream);
// prepare + apply
int launch_lu_solve_fp64(const double* LU, long long lda, int n, const int* piv, double* x, void* work,
                         hipStream_t stream);

// Adjoint (gradient) path.  solve.hip: Z^T lam = g with the FP64 factors
// (acc: n complex of scratch).  fill.hip: the fill's adjoint contraction
// for observation rows i0 .. i0+rows-1: gi (7 per segment, written for
// those rows) and gj (13 per segment, accumulated: zero it before the
// first chunk); pj holds rows * n * 13 doubles of pair terms.
int launch_lu_solve_transposed(const double* LU, long long lda, int n, const int* piv, const double* g, double* lam,
                               double* acc, hipStream_t stream);
int adjoint_rows_per_chunk(int n, size_t budget_bytes);
int launch_adjoint_fill(const DevSegments& g, int n, int i0, int rows, double rkh, const double* lam,
                        const double* abc, double* gi, double* pj, double* gj, hipStream_t stream);

// refine.hip: the mixed path's FP32 copy of Z, FP64 residual r = b - Z x
// (acc[0] = |r|^2, acc[1] = |b|^2 when with_b) and update x += d.
int launch_to_f32(const double* Z, long long ldz, float* Z32, long long ld32, int n, hipStream_t stream);
int launch_residual(const double* Z, long long ldz, int n, const double* x, const double* b, double* r, double* acc,
                    bool with_b, hipStream_t stream);
int launch_update(double* x, const double* d, int n, hipStream_t stream);

// current.hip: basis amplitudes x (complex) -> per-segment coefficients
// abc (6*n: air, aii, bir, bii, cir, cii) and segment-centre currents curx.
int launch_currents(int n, const double* x, const int* segptr, const int* segbasis, const double* segabc,
                    double* abc, double* curx, hipStream_t stream);

// current.hip: antenna input parameters of nsrc voltage sources (seg 0-based,
// volt complex): out[8*s..] = Z, Y, I (complex), P.
int launch_zin(int nsrc, const int* seg, const double* volt, const double* curx, double wavelength, double* out,
               hipStream_t stream);

// pattern.hip
struct PatternSpec {
  int n_theta, n_phi;
  // ground: 0 free space, 1 perfect, 2 finite (zrati zr); with a ground the
  // structure image is added and directions flagged below the horizon in
  // trig (theta > 90.01 deg) read as no field
  int ground;
  double zr_re, zr_im;
  int output_format, normalization, ipd, iavp;
  double wavelength, gcon, gcop, exrm, exra, rfld;
  double dphi_rad, half_dtheta_rad;
};
struct PatternOut {
  double *gain, *pg_vert, *pg_horiz, *pg_tot, *pg_rhcp, *pg_lhcp, *axial, *tilt, *avg_scale;
  int* sense;
  double *e_theta, *e_phi, *pint;  // e_*: 2 doubles per direction
};
// Directions d = kph * n_theta + kth.  trig (of the reference's accumulated
// angles): sin theta, cos theta, theta (radians) (n_theta each), sin phi,
// cos phi (n_phi each), then, with a ground, a below-horizon flag per theta
// (1.0 where theta > 90.01 deg, else 0.0).
// Number of segment slices launch_pattern uses (scratch = slices*n_dir*6
// doubles, 12 with a ground).
int pattern_slices(int n, int n_dir);
int launch_pattern(const DevSegments& g, int n, const double* abc, const PatternSpec& spec, const double* trig,
                   double* eth_eph_scratch, const PatternOut& out, hipStream_t stream);

}  // namespace necgpu
