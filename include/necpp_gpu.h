/*
 * necpp_gpu.h -- entry points of the MI355X engine that have no counterpart
 * in nec2++'s libnecpp.h.  They sit beside the drop-in API and never change
 * the behaviour of the reference calls.  Return conventions follow
 * libnecpp.h: 0 = success, 1 = error (text via nec_error_message()).
 */
#pragma once

#include <stdint.h>

#include "libnecpp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Library and device identification: writes a NUL-terminated text of at
 * most len bytes ("<gcn arch> <device name> <library version>"). */
long nec_gpu_info(char* out, int len);

/* Number of segments of a completed geometry (-1 before geometry_complete). */
long nec_gpu_segment_count(nec_context* ctx);

/* Segment table after geometry_complete (metres): out receives 8*n doubles,
 * per segment x, y, z (centre), length, radius, cos(alpha)cos(beta),
 * cos(alpha)sin(beta), sin(alpha); conn receives 3*n ints: icon1, icon2,
 * tag.  Either pointer may be NULL. */
long nec_gpu_segments(nec_context* ctx, double* out, int* conn, long n);

/* Precision of the factorisation for the following XQ / RP cards:
 * 0 = FP64 LU (default), 1 = complex FP32 LU on MFMA with FP64 iterative
 * refinement against the FP64 matrix. */
long nec_gpu_set_precision(nec_context* ctx, int mode);

/* Mixed precision: relative residual target of the refinement (default
 * 1e-12, i.e. |b - Z x| <= tol |b| with the FP64 Z), and, after a solve,
 * the refinement steps taken and the final relative residual. */
long nec_gpu_set_refine_tolerance(nec_context* ctx, double tol);
long nec_gpu_refinement(nec_context* ctx, int* steps, double* rel_residual);

/* How often this context's factorisations were repeated: lu = after a panel
 * exchange timed out (the cooperative panel's workgroups could not all be
 * resident beside the look-ahead update, e.g. on a device shared with other
 * work; the repeat keeps every launch on one stream), pivot = mixed
 * precision's block-local pivots replaced by partial pivoting.  Either
 * pointer may be NULL.  No counterpart in the reference (new entry point). */
long nec_gpu_factor_fallbacks(nec_context* ctx, long* lu, long* pivot);

/* Adjoint gradient of the last solve (XQ or RP card; one voltage source,
 * FP64 path, thin-wire kernel) for the design objective
 *   J = -G(theta, phi) + weight * |Z_in - z0|^2,
 * G the total power gain (dBi) in direction (theta, phi) and Z_in the input
 * impedance of the source.  The transposed solve Z^T lam = dJ/dx reuses the
 * LU factors; the fill derivative lam^T (dZ/dp) x is contracted on the GPU
 * without storing dZ/dp.  Writes J to *J_out, dJ with respect to every
 * segment's endpoints and radius to grad_out (7 doubles per segment: x1 y1 z1
 * x2 y2 z2 radius, per metre; the endpoints after junction resolution), and
 * dJ/d(frequency in MHz) to *dfreq_out (may be NULL).  info (may be NULL)
 * receives 6 doubles: gain (dBi), Z_in (re, im), host ms, device ms,
 * dJ/d(wavelength) at fixed wavelength-scaled geometry.  No counterpart in
 * the reference (new entry point). */
long nec_gpu_adjoint_gradient(nec_context* ctx, double theta_deg, double phi_deg, double weight, double z0_re,
                              double z0_im, double* J_out, double* grad_out, double* dfreq_out, double* info);

/* Segment currents (amperes, complex interleaved re,im) of antenna-input
 * result freq_index: writes 2*n doubles, n = nec_gpu_segment_count(). */
long nec_gpu_currents(nec_context* ctx, int freq_index, double* out, long n);

/* Stage times in milliseconds of the most recent solve, measured with HIP
 * events on the engine's stream: out[0] fill, out[1] factor, out[2] solve
 * + current coefficients, out[3] pattern. */
long nec_gpu_stage_times(nec_context* ctx, double* out);

/* Benchmark: repeat the whole solve of the most recent XQ / RP at its last
 * frequency (fill, LU, solve, current coefficients, Z_in and, after RP, the
 * pattern) warmup + steps times entirely on the device, inputs resident in
 * HBM, no host round trip inside a solve.  out[0] = ms per solve over the
 * timed steps; out[1..4] = mean fill / factor / solve / pattern ms; out[5],
 * out[6] = Z_in of the first feed from the last repetition. */
long nec_gpu_benchmark_solve(nec_context* ctx, int warmup, int steps, double* out);

/* Per-kernel times of the timed steps of the last nec_gpu_benchmark_solve:
 * every launch there carries its own start/stop HIP events
 * (hipExtLaunchKernelGGL) on the engine's stream.  For kernel family k
 * (0 <= k < the return value): out[2k] = launches, out[2k+1] = summed
 * milliseconds.  n = length of out in doubles.  Returns the number of
 * families, or -1 on error.  nec_gpu_kernel_name(k) names family k. */
long nec_gpu_kernel_times(nec_context* ctx, double* out, long n);

/* Per-kernel events on (1, the default) or off (0) for the following
 * nec_gpu_benchmark_solve calls.  Each event pair adds marker packets to the
 * queue (about 5 % of an N = 3,280 solve), so a wall-clock measurement of
 * the solve runs with them off and nec_gpu_kernel_times reports nothing new. */
long nec_gpu_set_kernel_timing(nec_context* ctx, int on);
const char* nec_gpu_kernel_name(int kind);

/* Impedance matrix of the completed geometry at freq_mhz (FP64 fill only,
 * no factorisation): out receives 2*n*n doubles, Z[i][k] (observation
 * segment i, basis k) at out[2*(i*n+k)], i.e. nec2++'s "cm" layout.  Leaves
 * the context needing a refill before the next solve. */
long nec_gpu_fill_matrix(nec_context* ctx, double freq_mhz, double* out, long n);

/* Pattern arrays of pattern result freq_index.  which: 0 gain, 1 vertical
 * (or major-axis) gain, 2 horizontal (or minor-axis) gain, 3 total gain,
 * 4 RHCP, 5 LHCP, 6 axial ratio, 7 tilt, 8 sense (as doubles), 9 averaging
 * scale, 10 E_theta, 11 E_phi (complex: 2 doubles per point).  Points are
 * ordered phi_index * n_theta + theta_index.  len = capacity in doubles. */
long nec_gpu_pattern_array(nec_context* ctx, int freq_index, int which, double* out, long len);

/* Average power gain and its solid angle / pi of pattern result freq_index
 * (A = 1 or 2 on the RP card). */
long nec_gpu_pattern_average(nec_context* ctx, int freq_index, double* avg_gain, double* solid_angle);

/* ------------------------------------------------------------------ */
/* Batched solves: many small structures (straight wires in free space,   */
/* at most 256 segments each), each at n_freq frequencies, one voltage    */
/* source each, with an optional pattern cut.  No counterpart in nec2++,  */
/* which runs one structure per nec_context; each (design, frequency) is  */
/* what an EX 0 + FR + XQ/RP deck would compute.                          */
/* ------------------------------------------------------------------ */
typedef struct nec_gpu_batch {
  int n_designs;
  const int* wire_ptr;      /* n_designs + 1: wires of design d are [wire_ptr[d], wire_ptr[d+1]) */
  const int* wire_tag;      /* per wire (GW tag) */
  const int* wire_ns;       /* per wire (GW segment count) */
  const double* wire_geom;  /* 7 per wire: x1 y1 z1 x2 y2 z2 radius, metres */
  const int* src_tag;       /* per design: EX 0 tag */
  const int* src_seg;       /* per design: EX 0 segment within the tag (1-based) */
  const double* src_volt;   /* 2 per design: source voltage */
  int n_freq;               /* frequencies per design */
  const double* freq_mhz;   /* n_designs * n_freq, design-major */
  int n_dir;                /* pattern directions (0: no pattern) */
  const double* theta_deg;  /* n_dir */
  const double* phi_deg;    /* n_dir */
  double rkh;               /* KH card distance in wavelengths (0: default 1) */
  int panel_width;          /* batched-LU panel width 8, 16 or 32; 0: autotune */
  int repeats;              /* timed replays of the captured pipeline graph */
} nec_gpu_batch;

/* Runs the batch: zin receives 2 doubles per solve (solve s = d * n_freq + f),
 * gain n_dir doubles per solve (total power gain, dB; may be NULL when
 * n_dir == 0), stats 12 doubles: solves, segments, host ms, fill ms, LU ms,
 * solve+pattern ms, ms per graph replay, chosen panel width, autotune ms for
 * widths 8 / 16 / 32, graph replays equal to the eager pass (1/0/-1).
 * Returns 0, or 1 with nec_error_message() set. */
long nec_gpu_batch_solve(const nec_gpu_batch* batch, double* zin, double* gain, double* stats);

/* Run a NEC card deck (the text of a .nec input file) on in_context, as
 * nec2++ -i deck.nec would (src/nec2cpp.cpp:290-440 job loop; comment
 * cards, the geometry section GW/GC/GS/GE, then data cards until EN or the
 * end of the text).  Results are read back with the libnecpp getters.
 * Cards outside the GPU path stop the deck with their "not supported"
 * message, as the API reports them; results computed before that card stay
 * in the context.  Returns 0, or 1 with nec_error_message() naming the deck
 * line and card. */
long nec_gpu_deck(nec_context* in_context, const char* deck);

/* The near field of NE or NH result number `index` (0-based, in the order
 * computed): for each point x, y, z (metres) then the field's x, y, z
 * components (re, im; V/m for NE, A/m for NH), 9 doubles, at most
 * max_points written to out.  Returns the number of points, or -1 when there
 * is no such result.  NE / NH (src/nec_context.cpp:896-955, nfpat
 * :5068-5188, nefld :4304-4450, nhfld :5189-5262) run on the GPU for wires
 * in free space or over a perfect or reflection-coefficient ground; the
 * reference exposes near fields only through its printed output. */
long nec_gpu_near_field(nec_context* in_context, int index, double* out, long max_points);

/* Which field near-field result `index` holds: 0 electric (NE card), 1
 * magnetic (NH card), -1 no such result.  No counterpart in the reference
 * (its printed output names the field). */
long nec_gpu_near_field_kind(nec_context* in_context, int index);

/* Adjoint gradients of a whole batch (the structures and frequencies of
 * `batch`; its pattern and repeat fields are not used): for every (design,
 * frequency) solve s, in the order of nec_gpu_batch_solve, the objective of
 * nec_gpu_adjoint_gradient, J[s], its gradient with respect to every
 * segment's endpoints and radius (grad: 7 doubles per segment, per metre,
 * solves one after another) and dJ/d(frequency in MHz) (dfreq[s]).  One
 * batched forward pass and one batched adjoint pass on the GPU (transposed
 * solves with the batch's LU factors, fill contraction); the per-solve host
 * work runs on the host's cores.  info (may be NULL): 3 doubles per solve,
 * gain (dBi), Z_in (re, im).  stats (may be NULL): 10 doubles, solves,
 * segments, host preparation ms, fill ms, LU ms, solve ms, host objective ms,
 * device adjoint ms, host gradient ms, total ms.  N <= 512 per structure.
 * No counterpart in the reference (new entry point). */
long nec_gpu_batch_adjoint(const nec_gpu_batch* batch, double theta_deg, double phi_deg, double weight, double z0_re,
                           double z0_im, double* J, double* grad, double* dfreq, double* info, double* stats);

/* ------------------------------------------------------------------ */
/* Device-pointer entry points (no host round trip).  Matrices are      */
/* row-major complex128 (interleaved), leading dimension lda entries.   */
/* ------------------------------------------------------------------ */

/* In-place partial-pivot LU of the n x n matrix at device pointer a.
 * piv: device int buffer of 2*n entries (transpositions, then the composed
 * row permutation).  info: device int, set to 1 + first zero pivot or 0.
 * work: device scratch of necgpu_zgetrf_workspace(n) bytes (no
 * initialisation needed).  stream: hipStream_t (NULL = default stream).
 * Returns 1 if n exceeds the largest order the panel kernel supports
 * (28,672). */
long long necgpu_zgetrf_workspace(int n);
long necgpu_zgetrf(double* a, long long lda, int n, int* piv, int* info, void* work, void* stream);

/* necgpu_zgetrf without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.
 * The same factorisation and result conventions as necgpu_zgetrf, on the
 * caller's stream only (no look-ahead stream), with the workspace held in a
 * fixed array of the library on each device (about 4.8 MiB, kept until the
 * process ends).  n <= 4,096: a larger n returns 1 with a message.  Calls
 * that overlap on one device share that workspace and race; serialise them
 * (one stream, or events).  piv and info as for necgpu_zgetrf. */
long necgpu_zgetrf_noalloc(double* a, long long lda, int n, int* piv, int* info, void* stream);

/* The mixed path's complex FP32 LU on caller device memory: a is n x n
 * complex64 (row-major, leading dimension lda), factored in place; piv,
 * info as for necgpu_zgetrf; work: necgpu_zgetrf_workspace(n) bytes.
 * pivot_mode 0: partial pivoting (first maximum of |z| in the column);
 * 1: pivots searched inside each 64-row diagonal block only, so P A = L U
 * with P permuting inside 64-row blocks (the default of the mixed solve,
 * whose FP64 refinement makes up for the weaker pivoting).  Returns 1 if n
 * exceeds the largest order the panel supports (57,344). */
long necgpu_cgetrf(float* a, long long lda, int n, int* piv, int* info, void* work, int pivot_mode, void* stream);

/* Solve with the factors of necgpu_zgetrf: x (device, n complex) is
 * overwritten.  work: device buffer of necgpu_zgetrs_workspace(n) bytes
 * (no initialisation needed; calls on one workspace are ordered by their
 * stream).  epoch: ignored (each call clears what its sweeps poll); kept so
 * that callers of the earlier epoch-chained sweeps still link. */
long long necgpu_zgetrs_workspace(int n);
long necgpu_zgetrs(const double* lu, long long lda, int n, const int* piv, double* x, void* work, unsigned epoch,
                   void* stream);

/* Self-check of the FP64 matrix-core lane layout the LU relies on: one wave
 * feeds a[l], b[l] (device, 64 doubles each) to v_mfma_f64_16x16x4f64 and
 * writes the 4 result registers of every lane to d (device, 256 doubles). */
long necgpu_check_mfma_f64(const double* a, const double* b, double* d, void* stream);
/* The same for v_mfma_f32_16x16x4f32 (64 floats in each, 256 out), the
 * instruction of the FP32 trailing update of the mixed-precision path. */
long necgpu_check_mfma_f32(const float* a, const float* b, float* d, void* stream);

#ifdef __cplusplus
}
#endif
