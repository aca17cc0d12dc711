// lu.hip -- blocked right-looking complex LU with partial pivoting (FP64).
//
// Replaces lu_decompose / Eigen PartialPivLU (src/matrix_algebra.cpp:217-266,
// src/eigen/Eigen/src/LU/PartialPivLU.h:338-465).  The matrix is the fill's
// row-major Z (row i = observation segment), which is the reference's "cm"
// storage; pivoting exchanges whole rows, which are contiguous here.  Pivot
// rule: first maximum of |z| = hypot(re, im) in the column, as Eigen's
// scalar_score_coeff_op; ties go to the lower row.
//
// Per panel of NB = 64 columns:
//   lu_panel   G cooperating workgroups, each holding up to ~112 rows of
//              the panel in LDS.  Per column: local pivot candidate, one
//              grid-wide exchange (candidate rows published to global
//              memory, one atomic arrival counter), row exchange between
//              the two owners, complex-division scaling and rank-1 update
//              in LDS.
//   lu_laswp   the panel's row exchanges applied to the columns outside it.
//   lu_trsm    U12 = L11^-1 A12 with L11 and a 64-column block in LDS.
//   zgemm      A22 -= L21 * U12 on FP64 MFMA (v_mfma_f64_16x16x4f64), the
//              complex product in 4M form; 64x64 complex tile per
//              workgroup, operands staged in LDS as split re/im planes.
#include <climits>
#include <cstdlib>
#include <mutex>

#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

constexpr int NB = 64;
constexpr int PANEL_THREADS = 1024;
constexpr int GMAX = 256;     // most cooperating panel workgroups
constexpr int RMAX = 112;     // most FP64 panel rows per workgroup (LDS: RMAX*(NB+1)*16 B)

// Panel rows per workgroup for element type T (FP32 rows take half the LDS).
template <typename T>
constexpr int rmax_of() {
  return sizeof(T) == 8 ? RMAX : 2 * RMAX;
}

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// Exchange area of the cooperative panel.  Every slot is written once per
// panel, so a reader knows a value is current when it differs from the
// sentinel the slot was cleared to (all-ones bits: a NaN payload that no
// arithmetic produces from the canonical NaNs the panel loads).  Two areas
// alternate by panel parity; each panel clears the other one for its
// successor, which runs only after it has finished.
// Layout of one area (Gcap = most panel workgroups of this n):
//   val [NB][Gcap] doubles   candidate |z| of group g for column c
//   idx [NB][Gcap] ints      candidate row (INT_MAX: none)
//   row [NB][Gcap][2 NB]     candidate row (whole panel row)
//   diag[NB][2 NB]           row k, published by its owner
struct Xchg {
  double* val;
  int* idx;
  double* row;
  double* diag;
};

__host__ __device__ constexpr size_t xchg_area_bytes(int Gcap) {
  return ((size_t)NB * Gcap * 8 + 255) / 256 * 256 + ((size_t)NB * Gcap * 4 + 255) / 256 * 256 +
         (size_t)NB * Gcap * 2 * NB * 8 + (size_t)NB * 2 * NB * 8;
}

__host__ __device__ inline Xchg xchg_area(char* base, int Gcap, int parity) {
  char* p = base + (size_t)parity * xchg_area_bytes(Gcap);
  Xchg x;
  x.val = reinterpret_cast<double*>(p);
  p += ((size_t)NB * Gcap * 8 + 255) / 256 * 256;
  x.idx = reinterpret_cast<int*>(p);
  p += ((size_t)NB * Gcap * 4 + 255) / 256 * 256;
  x.row = reinterpret_cast<double*>(p);
  p += (size_t)NB * Gcap * 2 * NB * 8;
  x.diag = reinterpret_cast<double*>(p);
  return x;
}

struct LuWork {
  int* status;
  int* ticket;  // arrival count of lu_next_update_kernel (zero between launches)
  char* xbase;  // two exchange areas
  int Gcap;
};

constexpr long long kSentinelBits = -1LL;

__device__ __forceinline__ void xstore(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double xload(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xstore_i(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int xload_i(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool is_sentinel(double v) { return __double_as_longlong(v) == kSentinelBits; }

// Clear the exchange slots that a panel of G groups will use (threads
// tid = 0 .. nthreads-1 share the work).
__device__ void clear_xchg(const Xchg& x, int Gcap, int G, int tid, int nthreads) {
  const double s = __longlong_as_double(kSentinelBits);
  const int nv = NB * G;
  for (int e = tid; e < nv; e += nthreads) {
    const int c = e / G, g = e % G;
    xstore(&x.val[c * Gcap + g], s);
    xstore_i(&x.idx[c * Gcap + g], -1);
  }
  const int nr = NB * G * 2 * NB;
  for (int e = tid; e < nr; e += nthreads) {
    const int c = e / (G * 2 * NB), rem = e % (G * 2 * NB);
    xstore(&x.row[(size_t)c * Gcap * 2 * NB + rem], s);
  }
  for (int e = tid; e < NB * 2 * NB; e += nthreads) xstore(&x.diag[e], s);
}

// Bounded poll of one exchange word; a timeout sets the sticky status.
__device__ __forceinline__ double poll(const double* p, int* status) {
  double v = xload(p);
  long spins = 0;
  while (is_sentinel(v)) {
    if (++spins > (1L << 22) || (spins % 64 == 0 && xload_i(status) != 0)) {
      __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return 0.0;
    }
    __builtin_amdgcn_s_sleep(1);
    v = xload(p);
  }
  return v;
}

__device__ __forceinline__ bool better(double ob, int oi, double best, int idx) {
  return ob > best || (ob == best && oi < idx);
}

__device__ __forceinline__ double canonical(double v) {
  return (v != v) ? __longlong_as_double(0x7FF8000000000000LL) : v;
}

__global__ void lu_init_kernel(int* perm, int n, int* status, int* ticket, int* info, char* xbase, int Gcap, int G0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) perm[i] = i;
  if (i == 0) {
    *status = 0;
    *ticket = 0;
    *info = 0;
  }
  clear_xchg(xchg_area(xbase, Gcap, 0), Gcap, G0, i, gridDim.x * blockDim.x);
}

// The wave's best (v, i): largest v, ties to the smallest i; v >= 0 or the
// sentinel -1 (no candidate, with i = INT_MAX), so the bit patterns of v
// order as signed 64-bit integers.  DPP row shifts then row broadcasts, a
// few cycles per step where a shuffle is an LDS round trip; the result is
// wave-uniform.
__device__ __forceinline__ void wave_argmax(double& v, int& i) {
  long long b = __double_as_longlong(v);
  int hi = (int)(b >> 32), lo = (int)(b & 0xffffffffll), ix = i;
#define NECGPU_ARGMAX_STEP(CTRL, ROWMASK)                                           \
  {                                                                               \
    const int h2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWMASK, 0xf, false); \
    const int l2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWMASK, 0xf, false); \
    const int i2 = __builtin_amdgcn_update_dpp(ix, ix, CTRL, ROWMASK, 0xf, false); \
    const long long a = ((long long)hi << 32) | (unsigned int)lo;                  \
    const long long o = ((long long)h2 << 32) | (unsigned int)l2;                  \
    const bool take = o > a || (o == a && i2 < ix);                                 \
    hi = take ? h2 : hi;                                                            \
    lo = take ? l2 : lo;                                                            \
    ix = take ? i2 : ix;                                                            \
  }
  NECGPU_ARGMAX_STEP(0x111, 0xf)  // row_shr:1
  NECGPU_ARGMAX_STEP(0x112, 0xf)  // row_shr:2
  NECGPU_ARGMAX_STEP(0x114, 0xf)  // row_shr:4
  NECGPU_ARGMAX_STEP(0x118, 0xf)  // row_shr:8
  NECGPU_ARGMAX_STEP(0x142, 0xa)  // row_bcast:15
  NECGPU_ARGMAX_STEP(0x143, 0xc)  // row_bcast:31
#undef NECGPU_ARGMAX_STEP
  hi = __builtin_amdgcn_readlane(hi, 63);
  lo = __builtin_amdgcn_readlane(lo, 63);
  i = __builtin_amdgcn_readlane(ix, 63);
  v = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Factor panel columns [k0, k0+nb) over rows [k0, n) with G workgroups; the
// active workgroup g is launch block g*stride.  Rows owned by g:
// [k0 + g*R, min(n, k0 + (g+1)*R)), at most rmax_of<T>(), one thread per
// row.  Rows are held in LDS as T; arithmetic is FP64.
//
// Column c (k = k0 + c):
//  A  reduction of the column-c candidates (|z| from step E): wave max,
//     first lane by ballot (rows follow thread order), then two waves.
//  B  wave 0 applies the update still pending from column c-1 (columns
//     c+1..) to the candidate row and to row k, publishes both with the
//     candidate, and polls the other groups' candidates; the other waves
//     meanwhile apply that pending update to every other owned row, which
//     takes the bulk of the rank-1 work off the exchange's critical path.
//  D  the pivot row is fetched; rows k and p are exchanged.
//  E  one thread per row: scale column c, update column c+1 only, and form
//     the column-(c+1) candidate; the update of columns c+2.. is left to B.
template <typename T>
__global__ void __launch_bounds__(PANEL_THREADS) lu_panel_kernel(T* __restrict__ A, long long lda, int n, int k0,
                                                                  int nb, int G, int R, int stride, int* piv, int* info, char* xbase, int Gcap, int parity,
                                                                  int Gnext, int* status) {
  extern __shared__ double lds_raw[];
  constexpr int RMAX_T = rmax_of<T>();
  constexpr int NRW = (RMAX_T + 63) / 64;  // waves that hold rows
  if (blockIdx.x % stride != 0) return;
  constexpr int NW = PANEL_THREADS / 64;  // waves
  const int g = blockIdx.x / stride;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int LDR = NB + 1;  // padded complex row length in LDS
  T* rows = reinterpret_cast<T*>(lds_raw);  // [R][LDR] complex
  T* su = rows + 2 * (size_t)RMAX_T * LDR;   // current pivot row, nb complex
  __shared__ double s_best[NRW];
  __shared__ int s_idx[NRW];
  __shared__ int s_win[2];     // winner row, winner group
  __shared__ double s_piv[4];  // pivot row, columns c and c+1
  __shared__ double s_dg[4];   // row k, columns c and c+1
  const Xchg X = xchg_area(xbase, Gcap, parity);
  // the next panel's area: this panel's predecessor used it and has ended
  if (Gnext > 0) clear_xchg(xchg_area(xbase, Gcap, parity ^ 1), Gcap, Gnext, g * PANEL_THREADS + t, G * PANEL_THREADS);
  const int r0 = k0 + g * R;
  const int r1 = min(n, r0 + R);
  const int nrow = max(0, r1 - r0);
  {
    // every load of a thread in flight at once, then the LDS stores
    constexpr int PER = (RMAX_T * NB + PANEL_THREADS - 1) / PANEL_THREADS;
    double vr[PER], vi[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = t + PANEL_THREADS * u;
      vr[u] = vi[u] = 0.0;
      if (e < nrow * nb) {
        const int r = e / nb, c = e % nb;
        const T* v = A + 2 * ((size_t)(r0 + r) * lda + k0 + c);
        vr[u] = v[0];
        vi[u] = v[1];
      }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = t + PANEL_THREADS * u;
      if (e < nrow * nb) {
        const int r = e / nb, c = e % nb;
        rows[2 * (r * LDR + c)] = static_cast<T>(canonical(vr[u]));
        rows[2 * (r * LDR + c) + 1] = static_cast<T>(canonical(vi[u]));
      }
    }
  }
  __syncthreads();
  // candidate of column 0 (thread t holds row t)
  double best = -1.0;
  if (t < nrow) best = hypot(rows[2 * (t * LDR)], rows[2 * (t * LDR) + 1]);
  bool has = t < nrow;
  for (int c = 0; c < nb; ++c) {
    const int k = k0 + c;
    // A. candidate reduction (rows are held by waves 0 .. NRW-1 only)
    if (wave < NRW) {
      // rows follow thread order, so the lowest row at the maximum is the
      // first; NaN magnitudes are never candidates
      double m = (has && best >= 0.0) ? best : -1.0;
      int mi = (has && best >= 0.0) ? r0 + t : INT_MAX;
      wave_argmax(m, mi);
      if (lane == 0) {
        s_best[wave] = m;
        s_idx[wave] = mi;
      }
    }
    __syncthreads();
    // first maximum over the waves in row order
    int bidx = INT_MAX;
    double bval = -1.0;
#pragma unroll
    for (int w = 0; w < NRW; ++w)
      if (s_idx[w] != INT_MAX && (bidx == INT_MAX || s_best[w] > bval)) {
        bval = s_best[w];
        bidx = s_idx[w];
      }
    const int rcand = (bidx != INT_MAX) ? bidx - r0 : -1;  // local candidate row
    const int rk = (k >= r0 && k < r1) ? k - r0 : -1;       // local row k
    // columns c+1.. still carry the pending update of column c-1 with
    // multipliers rows[r][c-1] (rows below k-1, i.e. local rows >= k - r0)
    const int pend_lo = c + 1;
    const bool pending = (c > 0) && (pend_lo < nb);
    if (wave == 0) {
      // B0. candidate row and row k brought up to date, then published
      if (pending && lane >= pend_lo && lane < nb) {
        const zc u = zmk(su[2 * lane], su[2 * lane + 1]);
        for (int which = 0; which < 2; ++which) {
          const int r = which == 0 ? rcand : rk;
          if (r < 0 || (which == 1 && r == rcand)) continue;
          const zc l = zmk(rows[2 * (r * LDR + c - 1)], rows[2 * (r * LDR + c - 1) + 1]);
          const zc prod = zmul(l, u);
          rows[2 * (r * LDR + lane)] = static_cast<T>(rows[2 * (r * LDR + lane)] - prod.re);
          rows[2 * (r * LDR + lane) + 1] = static_cast<T>(rows[2 * (r * LDR + lane) + 1] - prod.im);
        }
      }
      // Slots go straight to the device coherence point (agent-scope
      // atomics); each is written once per panel, so no ordering between
      // them is needed: readers poll until a slot leaves the sentinel.
      if (rcand >= 0 && lane < nb) {
        double* dst = X.row + ((size_t)c * Gcap + g) * 2 * NB;
        xstore(&dst[2 * lane], rows[2 * (rcand * LDR + lane)]);
        xstore(&dst[2 * lane + 1], rows[2 * (rcand * LDR + lane) + 1]);
      }
      if (rk >= 0 && lane < nb) {
        xstore(&X.diag[(size_t)c * 2 * NB + 2 * lane], rows[2 * (rk * LDR + lane)]);
        xstore(&X.diag[(size_t)c * 2 * NB + 2 * lane + 1], rows[2 * (rk * LDR + lane) + 1]);
      }
      if (lane == 0) {
        xstore(&X.val[c * Gcap + g], bval);
        xstore_i(&X.idx[c * Gcap + g], bidx);
      }
      // a group without a candidate still fills the pivot words that every
      // reader polls (columns c and c+1 of its row slot)
      if (rcand < 0 && 2 * c + lane < 2 * min(c + 2, nb) && lane < 4)
        xstore(X.row + ((size_t)c * Gcap + g) * 2 * NB + 2 * c + lane, 0.0);
      // Winner over all groups (every group computes the same).  Each lane
      // polls one group's (val, idx) and, with them, columns c and c+1 of
      // that group's candidate row and of row k, so that the scaling step
      // can start before the whole pivot row has arrived.
      const int c1 = (c + 1 < nb) ? c + 1 : c;
      const double* dsrc = X.diag + (size_t)c * 2 * NB;
      double wv = -1.0, e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0;
      double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
      int wi = INT_MAX, wgl = 0;
      for (int h0 = 0; h0 < G; h0 += 64) {
        const int h = h0 + lane;
        const bool act = h < G;
        const double* rs = X.row + ((size_t)c * Gcap + (act ? h : 0)) * 2 * NB;
        double v = -1.0, a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int ix = INT_MAX;
        long spins = 0;
        for (;;) {
          bool ok = true;
          if (act) {
            v = xload(&X.val[c * Gcap + h]);
            ix = xload_i(&X.idx[c * Gcap + h]);
            a0 = xload(&rs[2 * c]);
            a1 = xload(&rs[2 * c + 1]);
            a2 = xload(&rs[2 * c1]);
            a3 = xload(&rs[2 * c1 + 1]);
            ok = !(is_sentinel(v) || ix == -1 || is_sentinel(a0) || is_sentinel(a1) || is_sentinel(a2) ||
                   is_sentinel(a3));
          }
          if (h0 == 0) {
            // row k's columns c, c+1 (same address in every lane)
            d0 = xload(&dsrc[2 * c]);
            d1 = xload(&dsrc[2 * c + 1]);
            d2 = xload(&dsrc[2 * c1]);
            d3 = xload(&dsrc[2 * c1 + 1]);
            ok = ok && !(is_sentinel(d0) || is_sentinel(d1) || is_sentinel(d2) || is_sentinel(d3));
          }
          if (ok) break;
          // bounded wait; a timeout is sticky so every later wait returns at once
          if (++spins > (1L << 22) || (spins % 64 == 0 && xload_i(status) != 0)) {
            __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v = -1.0;
            ix = INT_MAX;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (act && better(v, ix, wv, wi)) {
          wv = v;
          wi = ix;
          wgl = h;
          e0 = a0;
          e1 = a1;
          e2 = a2;
          e3 = a3;
        }
      }
      // first maximum over the groups (lowest row at it), then the winner's
      // words from its lane
      const int mine = wi;
      wave_argmax(wv, wi);
      const unsigned long long hit = __ballot(wi != INT_MAX && mine == wi);
      const int wl = hit ? (__ffsll((long long)hit) - 1) : 0;
      if (!hit) {
        wv = -1.0;
        wi = INT_MAX;
      }
      int wg = __builtin_amdgcn_readlane(wgl, wl);
      e0 = readlane_d(e0, wl);
      e1 = readlane_d(e1, wl);
      e2 = readlane_d(e2, wl);
      e3 = readlane_d(e3, wl);
      if (lane == 0) {
        int p = wi;
        if (p == INT_MAX || !(wv > 0.0)) {
          if (wv == 0.0 && g == 0 && *info == 0) *info = k + 1;
          p = k;
          wg = -1;
        }
        s_win[0] = p;
        s_win[1] = wg;
        // pivot row columns c, c+1: the winner's candidate row, or row k
        if (p != k) {
          s_piv[0] = e0;
          s_piv[1] = e1;
          s_piv[2] = e2;
          s_piv[3] = e3;
        } else {
          s_piv[0] = d0;
          s_piv[1] = d1;
          s_piv[2] = d2;
          s_piv[3] = d3;
        }
        s_dg[0] = d0;
        s_dg[1] = d1;
        s_dg[2] = d2;
        s_dg[3] = d3;
        if (g == 0) piv[k] = p;
      }
    } else if (pending) {
      // B1. the pending update for every other owned row below k - 1
      const int w = nb - pend_lo;
      const zc u = (lane < w) ? zmk(su[2 * (pend_lo + lane)], su[2 * (pend_lo + lane) + 1]) : zmk(0.0, 0.0);
      const int rstart = max(0, k - r0);
      for (int r = rstart + (wave - 1); r < nrow; r += 4 * (NW - 1)) {
        double lr[4], li[4], ar[4], ai[4];
        bool on[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rr = r + q * (NW - 1);
          on[q] = rr < nrow && rr != rcand && rr != rk && lane < w;
          lr[q] = li[q] = ar[q] = ai[q] = 0.0;
          if (on[q]) {
            lr[q] = rows[2 * (rr * LDR + c - 1)];
            li[q] = rows[2 * (rr * LDR + c - 1) + 1];
            ar[q] = rows[2 * (rr * LDR + pend_lo + lane)];
            ai[q] = rows[2 * (rr * LDR + pend_lo + lane) + 1];
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rr = r + q * (NW - 1);
          if (on[q]) {
            const zc prod = zmul(zmk(lr[q], li[q]), u);
            rows[2 * (rr * LDR + pend_lo + lane)] = static_cast<T>(ar[q] - prod.re);
            rows[2 * (rr * LDR + pend_lo + lane) + 1] = static_cast<T>(ai[q] - prod.im);
          }
        }
      }
    }
    __syncthreads();
    const int p = s_win[0];
    const int wg = s_win[1];
    const bool mine_p = (p != k) && p >= r0 && p < r1;
    // D. loads of the whole pivot row (and, in the winner's group, of row
    //    k's content for row p) are issued now and used after step E
    double fsr = 0.0, fsi = 0.0, fpr = 0.0, fpi = 0.0;
    const double* fsrc = nullptr;
    if (t < nb) {
      if (p != k) {
        fsrc = X.row + ((size_t)c * Gcap + wg) * 2 * NB;
      } else if (rk < 0) {
        fsrc = X.diag + (size_t)c * 2 * NB;
      }
      if (fsrc) {
        fsr = xload(&fsrc[2 * t]);
        fsi = xload(&fsrc[2 * t + 1]);
      } else {
        fsr = rows[2 * (rk * LDR + t)];
        fsi = rows[2 * (rk * LDR + t) + 1];
      }
      if (mine_p) {
        fpr = xload(&X.diag[(size_t)c * 2 * NB + 2 * t]);
        fpi = xload(&X.diag[(size_t)c * 2 * NB + 2 * t + 1]);
      }
    }
    // E. one thread per row below k: scale column c (Smith's division as
    //    g++'s complex divide, with the pivot's ratio and denominator formed
    //    once), update column c+1, and keep the column-(c+1) candidate.
    //    Row p (winner's group) takes row k's columns c, c+1 from the poll.
    const zc d = zmk(s_piv[0], s_piv[1]);
    const bool scale = (d.re != 0.0) || (d.im != 0.0);
    const bool swp = fabs(d.re) < fabs(d.im);
    const double ratio = swp ? d.re / d.im : d.im / d.re;
    const double denom = swp ? (d.re * ratio) + d.im : (d.im * ratio) + d.re;
    const bool next = c + 1 < nb;
    const zc u1 = next ? zmk(s_piv[2], s_piv[3]) : zmk(0.0, 0.0);
    has = false;
    const int r = t;
    if (r < nrow && r0 + r > k) {
      T* v = rows + 2 * (r * LDR + c);
      const bool is_p = (r0 + r == p);
      double lr = is_p ? s_dg[0] : v[0];
      double li = is_p ? s_dg[1] : v[1];
      if (scale) {
        const double ar = lr, ai = li;
        if (swp) {
          lr = ((ar * ratio) + ai) / denom;
          li = ((ai * ratio) - ar) / denom;
        } else {
          lr = ((ai * ratio) + ar) / denom;
          li = (ai - (ar * ratio)) / denom;
        }
      }
      v[0] = static_cast<T>(lr);
      v[1] = static_cast<T>(li);
      if (next) {
        // the multiplier as stored (FP32 rows round it)
        const zc prod = zmul(zmk(v[0], v[1]), u1);
        const double nr = (is_p ? s_dg[2] : v[2]) - prod.re;
        const double ni = (is_p ? s_dg[3] : v[3]) - prod.im;
        v[2] = static_cast<T>(nr);
        v[3] = static_cast<T>(ni);
        best = hypot(static_cast<double>(v[2]), static_cast<double>(v[3]));
        has = true;
      }
    }
    // the whole pivot row into su; rows k and p exchanged (row p keeps the
    // columns c, c+1 step E has just written)
    if (t < nb) {
      if (fsrc && (is_sentinel(fsr) || is_sentinel(fsi))) {
        fsr = poll(&fsrc[2 * t], status);
        fsi = poll(&fsrc[2 * t + 1], status);
      }
      su[2 * t] = static_cast<T>(fsr);
      su[2 * t + 1] = static_cast<T>(fsi);
      if (p != k && rk >= 0) {
        rows[2 * (rk * LDR + t)] = static_cast<T>(fsr);
        rows[2 * (rk * LDR + t) + 1] = static_cast<T>(fsi);
      }
      if (mine_p && t != c && !(next && t == c + 1)) {
        if (is_sentinel(fpr) || is_sentinel(fpi)) {
          fpr = poll(&X.diag[(size_t)c * 2 * NB + 2 * t], status);
          fpi = poll(&X.diag[(size_t)c * 2 * NB + 2 * t + 1], status);
        }
        rows[2 * ((p - r0) * LDR + t)] = static_cast<T>(fpr);
        rows[2 * ((p - r0) * LDR + t) + 1] = static_cast<T>(fpi);
      }
    }
    __syncthreads();
  }
  for (int e = t; e < nrow * nb; e += PANEL_THREADS) {
    const int r = e / nb, c = e % nb;
    T* v = A + 2 * ((size_t)(r0 + r) * lda + k0 + c);
    v[0] = rows[2 * (r * LDR + c)];
    v[1] = rows[2 * (r * LDR + c) + 1];
  }
}

// Apply the exchanges piv[k0 .. k0+nb) to columns [0, k0) and [k0+nb, n).
// Row exchanges of one panel (piv[k0 .. k0+nb)) applied to the columns
// [c0, c1) and [c2, c3).  The nb sequential exchanges touch at most 2 nb
// rows; wave 0 composes them into a list of (destination row, source row)
// moves (a ballot search over the rows outside the panel's first nb), then
// every move is done in one gather: all touched rows of the workgroup's 32
// columns are read into registers, then written to their destinations.
constexpr int LASWP_COLS = 32;
template <typename T>
__global__ void __launch_bounds__(256) lu_laswp_kernel(T* __restrict__ A, long long lda, int k0, int nb,
                                                       const int* __restrict__ piv, int c0, int c1, int c2, int c3,
                                                       int* __restrict__ perm) {
  __shared__ int s_dst[2 * NB], s_src[2 * NB];
  __shared__ int s_moves;
  const int t = threadIdx.x;
  if (t < 64) {
    // slot q < nb: row k0 + q; slot nb + e: e-th row outside [k0, k0+nb)
    // touched (held by lane e).  cont = original row whose content the
    // slot's row holds after the exchanges so far.
    const int lane = t;
    int cont_in = (lane < nb) ? k0 + lane : -1;  // panel rows
    int ext_row = -1, ext_cont = -1;              // outside rows
    int next_ext = 0;
    const int my_piv = (lane < nb) ? piv[k0 + lane] : 0;  // one load per lane, up front
    // every source lane below is wave-uniform (the step, a pivot read from
    // a lane, a ballot's first lane), so the reads are readlanes, not LDS
    // permutes: the chain of 64 steps is the launch's critical path
    for (int q = 0; q < nb; ++q) {
      const int k = k0 + q;
      const int p = __builtin_amdgcn_readlane(my_piv, q);
      if (p == k) continue;
      int ck = __builtin_amdgcn_readlane(cont_in, q);
      int cp;
      if (p < k0 + nb) {
        cp = __builtin_amdgcn_readlane(cont_in, p - k0);
        if (lane == p - k0) cont_in = ck;
      } else {
        const unsigned long long hit = __ballot(ext_row == p);
        int e;
        if (hit) {
          e = __ffsll((long long)hit) - 1;
        } else {
          e = next_ext++;
          if (lane == e) {
            ext_row = p;
            ext_cont = p;
          }
        }
        cp = __builtin_amdgcn_readlane(ext_cont, e);
        if (lane == e) ext_cont = ck;
      }
      if (lane == q) cont_in = cp;
    }
    // moves: slot rows whose content changed
    const bool m1 = lane < nb && cont_in != k0 + lane;
    const bool m2 = ext_row >= 0 && ext_cont != ext_row;
    const unsigned long long b1 = __ballot(m1), b2 = __ballot(m2);
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int n1 = __popcll(b1);
    if (m1) {
      const int i = __popcll(b1 & below);
      s_dst[i] = k0 + lane;
      s_src[i] = cont_in;
    }
    if (m2) {
      const int i = n1 + __popcll(b2 & below);
      s_dst[i] = ext_row;
      s_src[i] = ext_cont;
    }
    if (lane == 0) s_moves = n1 + __popcll(b2);
  }
  __syncthreads();
  const int moves = s_moves;
  if (perm && blockIdx.x == 0) {
    // the composed permutation takes the same moves (one launch per panel
    // passes it): every source entry read before any destination is written
    const int v = t < moves ? perm[s_src[t]] : 0;
    __syncthreads();
    if (t < moves) perm[s_dst[t]] = v;
  }
  const int w1 = c1 - c0;
  const int ncols = w1 + (c3 - c2);
  const int cbase = blockIdx.x * LASWP_COLS;
  const int cl = t % LASWP_COLS, rg = t / LASWP_COLS;  // 8 row groups
  const int idx = cbase + cl;
  const bool on = idx < ncols;
  const int col = idx < w1 ? c0 + idx : c2 + idx - w1;
  // at most 2 nb = 128 moves: 16 per row group, loads issued together
  {
    constexpr int RG = 256 / LASWP_COLS;
    T vr[2 * NB / RG], vi[2 * NB / RG];
#pragma unroll
    for (int u = 0; u < 2 * NB / RG; ++u) {
      const int m = rg + RG * u;
      vr[u] = vi[u] = T(0);
      if (on && m < moves) {
        const T* v = A + 2 * ((size_t)s_src[m] * lda + col);
        vr[u] = v[0];
        vi[u] = v[1];
      }
    }
    // every source row of this workgroup's columns is read before any
    // destination is written (destinations are sources of other moves)
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 2 * NB / RG; ++u) {
      const int m = rg + RG * u;
      if (on && m < moves) {
        T* v = A + 2 * ((size_t)s_dst[m] * lda + col);
        v[0] = vr[u];
        v[1] = vi[u];
      }
    }
  }
}

// U12 = L11^-1 A12: rows [k0, k0+nb), columns [cb, min(cb + 64, cend))
// with cb = cbeg + 64 * blockIdx.x, on FP64 MFMA.  L11 (unit lower) is cut
// into 16 x 16 blocks; wave w inverts diagonal block w, then every wave
// runs the blocked forward substitution for its own 16 columns,
//   X_b = D_b^-1 (A_b - sum_{j<b} L_bj X_j),   b = 0..3,
// with the 16 x 16 complex products as 4 real MFMAs (4M form).  Operand
// layout of v_mfma_f64_16x16x4f64: A[i][k] from lane i + 16 k, B[k][j] from
// lane j + 16 k; result register v of lane l holds D[l / 16 + 4 v][l % 16].
// One workgroup of 256 threads: columns [cb, min(cb + 64, cend)).  sL
// (NB x NB complex), sX (NB x 64 complex) and sDi (4 x 16 x 16 complex)
// are LDS; on return sX holds X.  WRITE: X also goes to A as it is formed.
template <typename T, bool WRITE>
__device__ __forceinline__ void trsm_block64(T* __restrict__ A, long long lda, int cend, int k0, int nb, int cb,
                                             double* sL, double* sX, double* sDi) {
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  {
    // all loads of a thread in flight at once, then the LDS stores
    double vr[NB * NB / 256], vi[NB * NB / 256];
#pragma unroll
    for (int u = 0; u < NB * NB / 256; ++u) {
      const int e = t + 256 * u;
      const int r = e / NB, q = e % NB;
      vr[u] = vi[u] = 0.0;
      if (r < nb && q < r) {
        const T* v = A + 2 * ((size_t)(k0 + r) * lda + k0 + q);
        vr[u] = v[0];
        vi[u] = v[1];
      }
    }
#pragma unroll
    for (int u = 0; u < NB * NB / 256; ++u) {
      sL[2 * (t + 256 * u)] = vr[u];
      sL[2 * (t + 256 * u) + 1] = vi[u];
    }
  }
  // right-hand side blocks of this wave's columns, in result layout
  const int li = lane & 15, lk = lane >> 4;
  const int col = cb + wave * 16 + li;
  const bool on = col < cend;
  double4_t tr[4], ti[4];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * b + lk + 4 * v;
      double xr = 0.0, xi = 0.0;
      if (on && r < nb) {
        const T* p = A + 2 * ((size_t)(k0 + r) * lda + col);
        xr = p[0];
        xi = p[1];
      }
      tr[b][v] = xr;
      ti[b][v] = xi;
    }
  __syncthreads();
  // D_w^-1: lane j < 16 forward-substitutes unit vector e_j
  if (lane < 16) {
    double yr[16], yi[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      yr[r] = (r == lane) ? 1.0 : 0.0;
      yi[r] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < 15; ++q) {
#pragma unroll
      for (int r = q + 1; r < 16; ++r) {
        const int gr = 16 * wave + r, gq = 16 * wave + q;
        const double lr = sL[2 * (gr * NB + gq)], lim = sL[2 * (gr * NB + gq) + 1];
        yr[r] -= lr * yr[q] - lim * yi[q];
        yi[r] -= lr * yi[q] + lim * yr[q];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      sDi[2 * ((wave * 16 + r) * 16 + lane)] = yr[r];
      sDi[2 * ((wave * 16 + r) * 16 + lane) + 1] = yi[r];
    }
  }
  __syncthreads();
  const int xc = wave * 16 + li;  // this lane's column within sX
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    // T_b -= L_bj X_j for finished blocks j < b
#pragma unroll
    for (int j = 0; j < b; ++j)
#pragma unroll
      for (int kk = 0; kk < 16; kk += 4) {
        const int lr_ = 16 * b + li, lq = 16 * j + kk + lk;
        const double ar = sL[2 * (lr_ * NB + lq)], ai = sL[2 * (lr_ * NB + lq) + 1];
        const double xr = sX[2 * ((16 * j + kk + lk) * 64 + xc)], xi = sX[2 * ((16 * j + kk + lk) * 64 + xc) + 1];
        tr[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar, xr, tr[b], 0, 0, 0);
        tr[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, xi, tr[b], 0, 0, 0);
        ti[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ar, xi, ti[b], 0, 0, 0);
        ti[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, xr, ti[b], 0, 0, 0);
      }
    // T_b into sX rows [16 b, 16 b + 16) (plain layout), then X_b = D_b^-1 T_b
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      sX[2 * ((16 * b + lk + 4 * v) * 64 + xc)] = tr[b][v];
      sX[2 * ((16 * b + lk + 4 * v) * 64 + xc) + 1] = ti[b][v];
    }
    // the wave's LDS writes precede its reads (in-order LDS); keep the
    // compiler from moving the reads above the writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double4_t xr4 = (double4_t){0.0, 0.0, 0.0, 0.0}, xi4 = (double4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      const double dr = sDi[2 * ((b * 16 + li) * 16 + kk + lk)], di = sDi[2 * ((b * 16 + li) * 16 + kk + lk) + 1];
      const double yr = sX[2 * ((16 * b + kk + lk) * 64 + xc)], yi = sX[2 * ((16 * b + kk + lk) * 64 + xc) + 1];
      xr4 = __builtin_amdgcn_mfma_f64_16x16x4f64(dr, yr, xr4, 0, 0, 0);
      xr4 = __builtin_amdgcn_mfma_f64_16x16x4f64(-di, yi, xr4, 0, 0, 0);
      xi4 = __builtin_amdgcn_mfma_f64_16x16x4f64(dr, yi, xi4, 0, 0, 0);
      xi4 = __builtin_amdgcn_mfma_f64_16x16x4f64(di, yr, xi4, 0, 0, 0);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * b + lk + 4 * v;
      sX[2 * (r * 64 + xc)] = xr4[v];
      sX[2 * (r * 64 + xc) + 1] = xi4[v];
      if (WRITE && on && r < nb) {
        T* p = A + 2 * ((size_t)(k0 + r) * lda + col);
        p[0] = static_cast<T>(xr4[v]);
        p[1] = static_cast<T>(xi4[v]);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

template <typename T>
__global__ void __launch_bounds__(256) lu_trsm_kernel(T* __restrict__ A, long long lda, int cend, int k0, int nb,
                                                      int cbeg) {
  __shared__ double sL[NB * NB * 2];       // L11[r][q], complex, zero outside the strict lower part
  __shared__ double sX[NB * 64 * 2];       // X[r][col], complex
  __shared__ double sDi[4 * 16 * 16 * 2];  // D_b^-1 [b][r][q], complex
  trsm_block64<T, true>(A, lda, cend, k0, nb, cbeg + blockIdx.x * 64, sL, sX, sDi);
}

// The look-ahead update of the next panel's columns [cb, cend) (at most
// 64) after panel [k0, k0+nb), in one launch: every workgroup forms
// U12 = L11^-1 A12 itself (trsm_block64, no global writes), then updates
// its 64-row tile of rows [k0+nb, n): C -= L21 U12 on FP64 MFMA (4M form,
// each wave a 32 x 32 block).  A12 is read by every workgroup, so U12 is
// written over it only by the workgroup that arrives last at the ticket
// (all others have read A12 by then); the ticket returns to zero.
template <typename T>
__global__ void __launch_bounds__(256) lu_next_update_kernel(T* __restrict__ A, long long lda, int n, int k0, int nb,
                                                             int cb, int cend, int* ticket) {
  __shared__ double sL[NB * NB * 2];       // L11, then this tile's L21 rows
  __shared__ double sX[NB * 64 * 2];       // U12[r][col]
  __shared__ double sDi[4 * 16 * 16 * 2];
  __shared__ int s_last;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int row0 = k0 + nb + 64 * blockIdx.x;
  const int cols = cend - cb;
  // the tile's C block and L21 rows: loads issued before the trsm
  constexpr int PER = 64 * 64 / 256;
  double cr[PER], ci[PER], lr[PER], li[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = t + 256 * u;
    const int r = e / 64, q = e % 64;
    cr[u] = ci[u] = lr[u] = li[u] = 0.0;
    if (row0 + r < n) {
      if (q < cols) {
        const T* v = A + 2 * ((size_t)(row0 + r) * lda + cb + q);
        cr[u] = v[0];
        ci[u] = v[1];
      }
      if (q < nb) {
        const T* v = A + 2 * ((size_t)(row0 + r) * lda + k0 + q);
        lr[u] = v[0];
        li[u] = v[1];
      }
    }
  }
  trsm_block64<T, false>(A, lda, cend, k0, nb, cb, sL, sX, sDi);
  __syncthreads();  // every read of A12 and L11 has completed
  if (t == 0) {
    __threadfence();
    const int old = atomicAdd(ticket, 1);
    s_last = (old == (int)gridDim.x - 1);
    if (s_last) atomicExch(ticket, 0);
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = t + 256 * u;
    sL[2 * e] = lr[u];
    sL[2 * e + 1] = li[u];
  }
  __syncthreads();
  if (s_last) {
    for (int e = t; e < nb * 64; e += 256) {
      const int r = e / 64, q = e % 64;
      if (q < cols) {
        T* v = A + 2 * ((size_t)(k0 + r) * lda + cb + q);
        v[0] = static_cast<T>(sX[2 * e]);
        v[1] = static_cast<T>(sX[2 * e + 1]);
      }
    }
  }
  if (row0 >= n) return;
  // C (64 x 64) -= L21 (64 x nb) U12 (nb x 64); wave (wr, wc) owns rows
  // 32 wr.., columns 32 wc..
  const int wr = wave >> 1, wc = wave & 1;
  const int lx = lane & 15, lk = lane >> 4;
  double4_t accr[2][2], acci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      accr[a][b] = (double4_t){0.0, 0.0, 0.0, 0.0};
      acci[a][b] = (double4_t){0.0, 0.0, 0.0, 0.0};
    }
  for (int kk = 0; kk < NB; kk += 4) {
    const int q = kk + lk;
    double xr[2], xi[2], yr[2], yi[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int r = wr * 32 + a * 16 + lx;
      xr[a] = sL[2 * (r * 64 + q)];
      xi[a] = sL[2 * (r * 64 + q) + 1];
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int c = wc * 32 + b * 16 + lx;
      yr[b] = sX[2 * (q * 64 + c)];
      yi[b] = sX[2 * (q * 64 + c) + 1];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[a], yr[b], accr[a][b], 0, 0, 0);
        accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[a], -yi[b], accr[a][b], 0, 0, 0);
        acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[a], yi[b], acci[a][b], 0, 0, 0);
        acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[a], yr[b], acci[a][b], 0, 0, 0);
      }
  }
  // C from registers (element e = t + 256 u is row e / 64, column e % 64)
  // through LDS (sL is free again) into the result layout
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = t + 256 * u;
    sL[2 * e] = cr[u];
    sL[2 * e + 1] = ci[u];
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // result register v of lane l holds D[(l / 16) + 4 v][l % 16]
        const int r = wr * 32 + a * 16 + v * 4 + lk;
        const int c = wc * 32 + b * 16 + lx;
        if (row0 + r < n && c < cols) {
          T* o = A + 2 * ((size_t)(row0 + r) * lda + cb + c);
          o[0] = static_cast<T>(sL[2 * (r * 64 + c)] - accr[a][b][v]);
          o[1] = static_cast<T>(sL[2 * (r * 64 + c) + 1] - acci[a][b][v]);
        }
      }
}

// C[m0.., n0..] (M x N) -= L[m0.., k0..k0+K) * U[k0..k0+K, n0..] on FP64
// MFMA, complex product in 4M form.  Tiles of 64 x 64 complex are taken in
// turn by the workgroups of a 1-D grid; within a tile, 4 waves each own a
// 32 x 32 block (2 x 2 MFMA blocks of 16 x 16, re and im accumulators).  K
// is staged through LDS 16 at a time (split re / im planes, 32 KB), the
// next chunk's global loads issued before the current chunk's MFMAs.
// Workgroups are numbered so that each XCD (launch index mod 8) walks a
// contiguous run of tiles, which keeps the shared L and U rows in its L2.
constexpr int GK = 16;
__global__ void __launch_bounds__(256) zgemm_update_kernel(double* __restrict__ A, long long lda, int m0, int n0, int M,
                                                           int N, int k0, int K) {
  __shared__ double sAr[GK * 64], sAi[GK * 64];  // [k][row]
  __shared__ double sBr[GK * 64], sBi[GK * 64];  // [k][col]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int tiles_x = (N + 63) / 64;
  const int ntiles = tiles_x * ((M + 63) / 64);
  const int G = gridDim.x;
  // XCD-contiguous numbering (dispatch places workgroup b on XCD b % 8)
  const int per = G / 8, extra = G % 8, x = blockIdx.x % 8, y = blockIdx.x / 8;
  const int logical = (x < extra) ? x * (per + 1) + y : extra * (per + 1) + (x - extra) * per + y;
  const int nchunks = (K + GK - 1) / GK;
  // loader coordinates: A chunk 64 rows x 16 k (thread: row t/4, k 4*(t%4)..+4)
  const int ar_ = t >> 2, ak = (t & 3) * 4;
  // B chunk 16 k x 64 cols (thread: k t/16, cols 4*(t%16)..+4)
  const int bq = t >> 4, bc = (t & 15) * 4;
  for (int tile = logical; tile < ntiles; tile += G) {
    const int row0 = (tile / tiles_x) * 64, col0 = (tile % tiles_x) * 64;
    double pa[8], pb[8];
    auto fetch = [&](int ch) {
      const int kb = ch * GK;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = kb + ak + j;
        pa[2 * j] = pa[2 * j + 1] = 0.0;
        if (row0 + ar_ < M && q < K) {
          const double* v = A + 2 * ((size_t)(m0 + row0 + ar_) * lda + k0 + q);
          pa[2 * j] = v[0];
          pa[2 * j + 1] = v[1];
        }
        const int qb = kb + bq;
        pb[2 * j] = pb[2 * j + 1] = 0.0;
        if (col0 + bc + j < N && qb < K) {
          const double* v = A + 2 * ((size_t)(k0 + qb) * lda + n0 + col0 + bc + j);
          pb[2 * j] = v[0];
          pb[2 * j + 1] = v[1];
        }
      }
    };
    double4_t accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        accr[a][b] = (double4_t){0.0, 0.0, 0.0, 0.0};
        acci[a][b] = (double4_t){0.0, 0.0, 0.0, 0.0};
      }
    fetch(0);
    for (int ch = 0; ch < nchunks; ++ch) {
      __syncthreads();  // previous chunk's operands are no longer read
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        sAr[(ak + j) * 64 + ar_] = pa[2 * j];
        sAi[(ak + j) * 64 + ar_] = pa[2 * j + 1];
        sBr[bq * 64 + bc + j] = pb[2 * j];
        sBi[bq * 64 + bc + j] = pb[2 * j + 1];
      }
      __syncthreads();
      if (ch + 1 < nchunks) fetch(ch + 1);
#pragma unroll
      for (int kk = 0; kk < GK; kk += 4) {
        const int q = kk + lk;
        double xr[2], xi[2], yr[2], yi[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          xr[a] = sAr[q * 64 + wr * 32 + a * 16 + li];
          xi[a] = sAi[q * 64 + wr * 32 + a * 16 + li];
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          yr[b] = sBr[q * 64 + wc * 32 + b * 16 + li];
          yi[b] = sBi[q * 64 + wc * 32 + b * 16 + li];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[a], yr[b], accr[a][b], 0, 0, 0);
            accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[a], -yi[b], accr[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xr[a], yi[b], acci[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xi[a], yr[b], acci[a][b], 0, 0, 0);
          }
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          // result register v of lane l holds D[(l / 16) + 4 v][l % 16]
          const int i = row0 + wr * 32 + a * 16 + v * 4 + lk;
          const int j = col0 + wc * 32 + b * 16 + li;
          if (i < M && j < N) {
            double* c = A + 2 * ((size_t)(m0 + i) * lda + n0 + j);
            c[0] -= accr[a][b][v];
            c[1] -= acci[a][b][v];
          }
        }
  }
}

// The same update, C[m0.., n0..] (M x N) -= L * U, as a persistent kernel
// of 256-thread workgroups that stay resident (FP64: one per CU, LDS 133 KB;
// FP32: two per CU, 66.5 KB each).  K is taken in 64-deep chunks (one per
// panel: K = 128 when the one-stream LU updates a pair of panels at once);
// LDS holds one chunk's 64 x 64 strips of L and U.  Per chunk the workgroup
//   - writes the strips it prefetched into registers to LDS and, at a tile's
//     first chunk, stores the previous tile's result and moves the
//     prefetched C block into the accumulators (which carry over from one
//     chunk of a tile to the next),
//   - runs 16 k-steps of MFMAs per wave (32 x 32 complex per wave) with the
//     accumulators counting down from C.  FP32: the complex product in 4M
//     form, 16 MFMAs per k-step: acc_re += (-Lr) Ur + Li Ui, acc_im +=
//     (-Lr) Ui + (-Li) Ur.  FP64: three real products (3M), 12 MFMAs per
//     k-step: X += (-Lr) Ur, Y += Li Ui, Z += -(Lr + Li)(Ur + Ui) from
//     X = Cr, Y = 0, Z = Ci + Cr, and the result Re = X + Y, Im = (Z - X)
//     + Y (the FP32 kernel is not MFMA-bound: three products measured no
//     faster there).  Each k-step also issues three of the 48 loads of the
//     next chunk's strips and C block (C at a tile's first chunk only), so
//     the loads trickle out beside the MFMAs instead of stalling them in
//     one burst.
// Storing a tile's result after the next tile's operands are in LDS keeps
// the waits for the prefetched loads from waiting on stores.  Loads and
// stores are buffer operations: an element outside the matrix gets an offset
// beyond the buffer's range, which reads as zero and drops the write, so the
// tile loop has no branches.  Tiles are taken in bands of 8 tile rows,
// column by column inside a band; each XCD (launch index mod 8) owns a
// contiguous range of that order, so the tiles its CUs hold at one time
// share a few L and U strips in its L2.
static_assert(NB == 64, "zgemm_tile_kernel stages K in chunks of 64");
constexpr int ZT_SP = 65;  // LDS row stride in complex (odd: the transposing L store is conflict-free)
typedef double dbl2_t __attribute__((ext_vector_type(2)));
typedef float flt2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
constexpr int ZT_OUT = 0x7fffffff;   // buffer range (bytes); every in-tile offset is below it
constexpr unsigned ZT_POISON = 0x80000000u;  // an offset beyond the range: reads 0, drops writes
constexpr int ZT_RSRC_W3 = 0x00020000;  // buffer resource word 3 for gfx9-family raw buffers

// Element traits: complex pair, accumulator, MFMA and the accumulator's
// row for register v of lane group lk.
template <typename T>
struct ZtTraits;
template <>
struct ZtTraits<double> {
  typedef dbl2_t C2;
  typedef double4_t Acc;
  static constexpr int kPerCU = 1;
  static constexpr bool k3M = true;  // three real products per complex one
  __device__ static Acc mfma(double a, double b, Acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  // FP64 16x16x4: register v of lane l holds D[(l / 16) + 4 v][l % 16]
  __device__ static int acc_row(int v, int lk) { return v * 4 + lk; }
  __device__ static C2 load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(C2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  }
  __device__ static void store(C2 d, __amdgpu_buffer_rsrc_t r, unsigned off) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, d), r, off, 0, 0);
  }
};
template <>
struct ZtTraits<float> {
  typedef flt2_t C2;
  typedef float4_t Acc;
  static constexpr int kPerCU = 2;
  static constexpr bool k3M = false;
  __device__ static Acc mfma(float a, float b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  // FP32 16x16x4: register v of lane l holds D[4 (l / 16) + v][l % 16]
  __device__ static int acc_row(int v, int lk) { return lk * 4 + v; }
  __device__ static C2 load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(C2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  }
  __device__ static void store(C2 d, __amdgpu_buffer_rsrc_t r, unsigned off) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, d), r, off, 0, 0);
  }
};

template <typename T>
constexpr size_t zt_lds_bytes() {
  return 2u * 64u * ZT_SP * sizeof(typename ZtTraits<T>::C2);
}

__device__ inline void zt_coords(int idx, int tiles_x, int tiles_y, int* ty, int* tx) {
  const int band = idx / (8 * tiles_x);
  const int hb = min(8, tiles_y - 8 * band);
  const int r = idx - band * 8 * tiles_x;
  *tx = r / hb;
  *ty = 8 * band + r % hb;
}

template <typename T>
__device__ inline __amdgpu_buffer_rsrc_t zt_rsrc(const T* base, int range) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), (short)0, range, ZT_RSRC_W3);
}

// A tile's result from its accumulators (see zgemm_tile_kernel), stored
// through the buffer resource rs at offsets os.
template <typename Tr>
__device__ __forceinline__ void zt_store_result(const typename Tr::Acc (&accr)[2][2], const typename Tr::Acc (&acci)[2][2],
                                                const typename Tr::Acc (&accz)[2][2], __amdgpu_buffer_rsrc_t rs,
                                                const unsigned (&os)[16]) {
  typedef typename Tr::C2 C2;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if constexpr (Tr::k3M) {
          const auto x = accr[a][b][v], y = acci[a][b][v], z = accz[a][b][v];
          Tr::store((C2){x + y, (z - x) + y}, rs, os[(a * 2 + b) * 4 + v]);
        } else {
          Tr::store((C2){accr[a][b][v], acci[a][b][v]}, rs, os[(a * 2 + b) * 4 + v]);
        }
      }
}

template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ZtTraits<T>::kPerCU, ZtTraits<T>::kPerCU)))
zgemm_tile_kernel(T* __restrict__ A, long long lda, int m0, int n0, int M, int N, int k0, int K) {
  typedef ZtTraits<T> Tr;
  typedef typename Tr::C2 C2;
  typedef typename Tr::Acc Acc;
  extern __shared__ __align__(16) unsigned char zt_raw[];
  C2* sL = reinterpret_cast<C2*>(zt_raw);  // [k][row]
  C2* sU = sL + 64 * ZT_SP;                 // [k][col]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int tiles_x = (N + 63) / 64, tiles_y = (M + 63) / 64;
  const int ntiles = tiles_x * tiles_y;
  const int G = gridDim.x;
  const int x = blockIdx.x % 8, y = blockIdx.x / 8, per = G / 8, extra = G % 8;
  const int gx = per + (x < extra ? 1 : 0);
  const int before = x * per + min(x, extra);
  const int tend = (int)((long long)ntiles * (before + gx) / G);
  int tile = (int)((long long)ntiles * before / G) + y;
  if (tile >= tend) return;
  const unsigned esz = 2u * sizeof(T);
  const unsigned row_bytes = (unsigned)(lda * esz);
  // A unit of work is one 64-deep chunk of K of one tile (K <= 128: one or
  // two chunks per tile); the accumulators carry over from a tile's first
  // chunk to its second, C is loaded before the first and stored after the
  // last.  Per-unit load sources and offsets.  Slot i (0..15) of this thread:
  //   L: row wave*16+i of the strip, column `lane` of the chunk
  //   U: row wave*16+i of the chunk, column `lane`
  //   C: accumulator element i = (a b v) of the wave's 32 x 32 block (only
  //      before a tile's first chunk; otherwise a zero-range resource)
  const int nch = (K + 63) / 64;
  __amdgpu_buffer_rsrc_t rl, ru, rc;
  unsigned ol[16], ou[16], oc[16];
  auto plan = [&](int tl, int ch, bool live) __attribute__((always_inline)) {
    int ty, tx;
    zt_coords(tl, tiles_x, tiles_y, &ty, &tx);
    const int row0 = ty * 64, col0 = tx * 64;
    const int range = live ? ZT_OUT : 0;
    const int kc = k0 + 64 * ch, Kc = min(64, K - 64 * ch);
    rl = zt_rsrc(A + 2 * ((size_t)(m0 + row0) * lda + kc), range);
    ru = zt_rsrc(A + 2 * ((size_t)kc * lda + n0 + col0), range);
    rc = zt_rsrc(A + 2 * ((size_t)(m0 + row0) * lda + n0 + col0), ch == 0 ? range : 0);
    const int mr = M - row0, nc = N - col0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int r = wave * 16 + i;
      ol[i] = (r < mr && lane < Kc) ? r * row_bytes + lane * esz : ZT_POISON;
      ou[i] = (r < Kc && lane < nc) ? r * row_bytes + lane * esz : ZT_POISON;
      const int a = i >> 3, b = (i >> 2) & 1, v = i & 3;
      const int cr = wr * 32 + a * 16 + Tr::acc_row(v, lk), cc = wc * 32 + b * 16 + li;
      oc[i] = (cr < mr && cc < nc) ? cr * row_bytes + cc * esz : ZT_POISON;
    }
  };
  C2 pl[16], pu[16], pc[16];
  int ch = 0;
  plan(tile, 0, true);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    pl[i] = Tr::load(rl, ol[i]);
    pu[i] = Tr::load(ru, ou[i]);
    pc[i] = Tr::load(rc, oc[i]);
  }
  // The first pass stores through a zero-range resource (dropped).
  // 4M: accr = Cr - sum Lr Ur + Li Ui, acci = Ci - sum Lr Ui + Li Ur.
  // 3M: accr = X = Cr - sum Lr Ur, acci = Y = sum Li Ui,
  //     accz = Z = Ci + Cr - sum (Lr + Li)(Ur + Ui); then the result is
  //     Re = X + Y, Im = (Z - X) + Y.
  Acc accr[2][2], acci[2][2], accz[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      accr[a][b] = (Acc){T(0), T(0), T(0), T(0)};
      acci[a][b] = (Acc){T(0), T(0), T(0), T(0)};
      accz[a][b] = (Acc){T(0), T(0), T(0), T(0)};
    }
  __amdgpu_buffer_rsrc_t rs = zt_rsrc(A, 0);
  unsigned os[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) os[i] = ZT_POISON;
  for (;;) {
    __syncthreads();  // the previous unit's operands are no longer read
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      sL[lane * ZT_SP + wave * 16 + i] = pl[i];
      sU[(wave * 16 + i) * ZT_SP + lane] = pu[i];
    }
    if (ch == 0) {
      // a new tile: the previous tile's result out, this tile's C in
      zt_store_result<Tr>(accr, acci, accz, rs, os);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          const C2* p = pc + (a * 2 + b) * 4;
          accr[a][b] = (Acc){p[0].x, p[1].x, p[2].x, p[3].x};
          if constexpr (Tr::k3M) {
            acci[a][b] = (Acc){T(0), T(0), T(0), T(0)};
            accz[a][b] = (Acc){p[0].y + p[0].x, p[1].y + p[1].x, p[2].y + p[2].x, p[3].y + p[3].x};
          } else {
            acci[a][b] = (Acc){p[0].y, p[1].y, p[2].y, p[3].y};
          }
        }
      // this tile's store target before plan() moves on
      rs = rc;
#pragma unroll
      for (int i = 0; i < 16; ++i) os[i] = oc[i];
    }
    // the next unit: this tile's next chunk, or the next tile's first
    int next = tile, nextch = ch + 1;
    if (nextch == nch) {
      next = tile + gx;
      nextch = 0;
    }
    const bool live = next < tend;
    plan(live ? next : tile, live ? nextch : 0, live);
    __syncthreads();
    C2 xa[2], yb[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) xa[a] = sL[lk * ZT_SP + wr * 32 + a * 16 + li];
#pragma unroll
    for (int b = 0; b < 2; ++b) yb[b] = sU[lk * ZT_SP + wc * 32 + b * 16 + li];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      pl[s] = Tr::load(rl, ol[s]);
      pu[s] = Tr::load(ru, ou[s]);
      pc[s] = Tr::load(rc, oc[s]);
      C2 xn[2], yn[2];
      if (s < 15) {
        const int q = (s + 1) * 4 + lk;
#pragma unroll
        for (int a = 0; a < 2; ++a) xn[a] = sL[q * ZT_SP + wr * 32 + a * 16 + li];
#pragma unroll
        for (int b = 0; b < 2; ++b) yn[b] = sU[q * ZT_SP + wc * 32 + b * 16 + li];
      }
      if constexpr (Tr::k3M) {
        T sa[2], sb[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) sa[a] = xa[a].x + xa[a].y;
#pragma unroll
        for (int b = 0; b < 2; ++b) sb[b] = yb[b].x + yb[b].y;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            accr[a][b] = Tr::mfma(-xa[a].x, yb[b].x, accr[a][b]);
            acci[a][b] = Tr::mfma(xa[a].y, yb[b].y, acci[a][b]);
            accz[a][b] = Tr::mfma(-sa[a], sb[b], accz[a][b]);
          }
      } else {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            accr[a][b] = Tr::mfma(-xa[a].x, yb[b].x, accr[a][b]);
            accr[a][b] = Tr::mfma(xa[a].y, yb[b].y, accr[a][b]);
            acci[a][b] = Tr::mfma(-xa[a].x, yb[b].y, acci[a][b]);
            acci[a][b] = Tr::mfma(-xa[a].y, yb[b].x, acci[a][b]);
          }
      }
      if (s < 15) {
#pragma unroll
        for (int a = 0; a < 2; ++a) xa[a] = xn[a];
#pragma unroll
        for (int b = 0; b < 2; ++b) yb[b] = yn[b];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    if (!live) break;
    tile = next;
    ch = nextch;
  }
  zt_store_result<Tr>(accr, acci, accz, rs, os);
}

// The mixed-precision path's FP32 trailing update: zgemm_update_kernel with
// FP32 operands on v_mfma_f32_16x16x4f32 (FP32 accumulation), whose result
// register v of lane l holds D[4 (l / 16) + v][l % 16].  The complex product
// takes three real products instead of four (T1 = Ar Br, T2 = Ai Bi,
// T3 = (Ar + Ai)(Br + Bi); Re = T1 - T2, Im = T3 - T1 - T2): these factors
// only precondition the FP64 refinement, which absorbs the slightly larger
// rounding of the imaginary part.
__global__ void __launch_bounds__(256) zgemm_update_f32_kernel(float* __restrict__ A, long long lda, int m0, int n0, int M,
                                                           int N, int k0, int K) {
  __shared__ float sAr[GK * 64], sAi[GK * 64];  // [k][row]
  __shared__ float sBr[GK * 64], sBi[GK * 64];  // [k][col]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int tiles_x = (N + 63) / 64;
  const int ntiles = tiles_x * ((M + 63) / 64);
  const int G = gridDim.x;
  // XCD-contiguous numbering (dispatch places workgroup b on XCD b % 8)
  const int per = G / 8, extra = G % 8, x = blockIdx.x % 8, y = blockIdx.x / 8;
  const int logical = (x < extra) ? x * (per + 1) + y : extra * (per + 1) + (x - extra) * per + y;
  const int nchunks = (K + GK - 1) / GK;
  // loader coordinates: A chunk 64 rows x 16 k (thread: row t/4, k 4*(t%4)..+4)
  const int ar_ = t >> 2, ak = (t & 3) * 4;
  // B chunk 16 k x 64 cols (thread: k t/16, cols 4*(t%16)..+4)
  const int bq = t >> 4, bc = (t & 15) * 4;
  for (int tile = logical; tile < ntiles; tile += G) {
    const int row0 = (tile / tiles_x) * 64, col0 = (tile % tiles_x) * 64;
    float pa[8], pb[8];
    auto fetch = [&](int ch) {
      const int kb = ch * GK;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int q = kb + ak + j;
        pa[2 * j] = pa[2 * j + 1] = 0.0f;
        if (row0 + ar_ < M && q < K) {
          const float* v = A + 2 * ((size_t)(m0 + row0 + ar_) * lda + k0 + q);
          pa[2 * j] = v[0];
          pa[2 * j + 1] = v[1];
        }
        const int qb = kb + bq;
        pb[2 * j] = pb[2 * j + 1] = 0.0f;
        if (col0 + bc + j < N && qb < K) {
          const float* v = A + 2 * ((size_t)(k0 + qb) * lda + n0 + col0 + bc + j);
          pb[2 * j] = v[0];
          pb[2 * j + 1] = v[1];
        }
      }
    };
    float4_t t1[2][2], t2[2][2], t3[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        t1[a][b] = (float4_t){0.0f, 0.0f, 0.0f, 0.0f};
        t2[a][b] = (float4_t){0.0f, 0.0f, 0.0f, 0.0f};
        t3[a][b] = (float4_t){0.0f, 0.0f, 0.0f, 0.0f};
      }
    fetch(0);
    for (int ch = 0; ch < nchunks; ++ch) {
      __syncthreads();  // previous chunk's operands are no longer read
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        sAr[(ak + j) * 64 + ar_] = pa[2 * j];
        sAi[(ak + j) * 64 + ar_] = pa[2 * j + 1];
        sBr[bq * 64 + bc + j] = pb[2 * j];
        sBi[bq * 64 + bc + j] = pb[2 * j + 1];
      }
      __syncthreads();
      if (ch + 1 < nchunks) fetch(ch + 1);
#pragma unroll
      for (int kk = 0; kk < GK; kk += 4) {
        const int q = kk + lk;
        float xr[2], xi[2], yr[2], yi[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          xr[a] = sAr[q * 64 + wr * 32 + a * 16 + li];
          xi[a] = sAi[q * 64 + wr * 32 + a * 16 + li];
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          yr[b] = sBr[q * 64 + wc * 32 + b * 16 + li];
          yi[b] = sBi[q * 64 + wc * 32 + b * 16 + li];
        }
        float xs[2], ys[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) xs[a] = xr[a] + xi[a];
#pragma unroll
        for (int b = 0; b < 2; ++b) ys[b] = yr[b] + yi[b];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            t1[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[a], yr[b], t1[a][b], 0, 0, 0);
            t2[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(xi[a], yi[b], t2[a][b], 0, 0, 0);
            t3[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(xs[a], ys[b], t3[a][b], 0, 0, 0);
          }
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          // result register v of lane l holds D[4 (l / 16) + v][l % 16]
          const int i = row0 + wr * 32 + a * 16 + 4 * lk + v;
          const int j = col0 + wc * 32 + b * 16 + li;
          if (i < M && j < N) {
            float* c = A + 2 * ((size_t)(m0 + i) * lda + n0 + j);
            const float p1 = t1[a][b][v], p2 = t2[a][b][v];
            c[0] -= p1 - p2;
            c[1] -= (t3[a][b][v] - p1) - p2;
          }
        }
  }
}

// One wave: lane l feeds a[l] as the A operand and b[l] as the B operand of
// v_mfma_f64_16x16x4f64 and writes its four result registers to d[4*l..].
__global__ void mfma_probe_kernel(const double* a, const double* b, double* d) {
  const int l = threadIdx.x;
  double4_t acc = (double4_t){0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[l], b[l], acc, 0, 0, 0);
  for (int v = 0; v < 4; ++v) d[4 * l + v] = acc[v];
}

// The same for v_mfma_f32_16x16x4f32.
__global__ void mfma_probe_f32_kernel(const float* a, const float* b, float* d) {
  const int l = threadIdx.x;
  float4_t acc = (float4_t){0.0f, 0.0f, 0.0f, 0.0f};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[l], b[l], acc, 0, 0, 0);
  for (int v = 0; v < 4; ++v) d[4 * l + v] = acc[v];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// exchange-area capacity: the FP64 panel's group count, the larger one
constexpr int gcap_for(int n) { return (n + RMAX - 1) / RMAX > 1 ? (n + RMAX - 1) / RMAX : 1; }

LuWork carve(void* work, int n) {
  char* p = static_cast<char*>(work);
  LuWork w;
  w.status = reinterpret_cast<int*>(p);
  w.ticket = reinterpret_cast<int*>(p + 64);
  w.xbase = p + 256;
  w.Gcap = gcap_for(n);
  return w;
}

int panel_stride_override() {
  const char* v = std::getenv("NECGPU_PANEL_STRIDE");
  return (v && *v) ? std::atoi(v) : 0;
}

// Workgroups per free CU for the trailing update that runs beside a panel:
// 1 while the update has fewer than 16 tiles per free CU (the panel beside
// it then runs faster, measured at N = 3,280), else 2 (measured faster at
// N = 16,020).  NECGPU_GEMM_OCC overrides, for measurements.
// Trailing update kernel (NECGPU_ZGEMM, for measurements): 1 = the
// persistent tile kernel everywhere, 0 = the per-tile kernels
// (zgemm_update_kernel, zgemm_update_f32_kernel) everywhere, 2 (default) =
// the tile kernel while the LU runs on one stream and for the large updates
// beside the look-ahead panels, the per-tile kernels for the small ones.
int zgemm_mode() {
  const char* v = std::getenv("NECGPU_ZGEMM");
  return (v && *v) ? std::atoi(v) : 2;
}

// L21 of the block-local FP32 panel: the MFMA kernel (default) or the
// column-by-column one (NECGPU_L21_MFMA=0, for measurements).
bool l21_mfma_enabled() {
  const char* v = std::getenv("NECGPU_L21_MFMA");
  return !(v && v[0] == '0');
}

int gemm_occupancy(int tiles, int free_cus) {
  const char* v = std::getenv("NECGPU_GEMM_OCC");
  const int o = (v && *v) ? std::atoi(v) : 0;
  if (o > 0) return o;
  return tiles < 16 * free_cus ? 1 : 2;
}

// Most 64-row tiles for the fused next-panel update (NECGPU_LU_NEXT_TILES
// overrides, for measurements; 0 turns it off).
int next_fused_max_tiles() {
  const char* v = std::getenv("NECGPU_LU_NEXT_TILES");
  return (v && *v) ? std::atoi(v) : 64;
}

bool lookahead_enabled() {
  const char* v = std::getenv("NECGPU_LU_LOOKAHEAD");
  return !(v && v[0] == '0');
}

// The one-stream phase of the LU in pairs of panels with one 128-deep
// trailing update per pair (default), or one update per panel
// (NECGPU_LU_PAIRS=0, for measurements).
bool lu_pairs_enabled() {
  const char* v = std::getenv("NECGPU_LU_PAIRS");
  return !(v && v[0] == '0');
}

// Largest panel (in workgroups, one CU each) that runs beside the trailing
// update.  A wider panel takes so many CUs from the update, and waits so
// long for CUs the update holds, that the two measured slower together than
// one after the other (N = 16,020: 583 ms against 446 ms for the FP64
// factorisation, 320 against 277 ms for the FP32 one); below it the panel
// hides behind the update (N = 3,280).  Limits of 24, 32, 48 and 64 groups
// measured 448, 447, 446, 444 ms (FP64) and 284, 279, 277, 297 ms (FP32) at
// N = 16,020.  NECGPU_LA_MAX_G overrides, for measurements.
int lookahead_max_groups() {
  const char* v = std::getenv("NECGPU_LA_MAX_G");
  return (v && *v) ? std::atoi(v) : 48;
}

// Second stream (default priority) and two events per device for the
// look-ahead: the next panel runs there while the caller's stream finishes
// the trailing update.  Created once, kept for the life of the process.
struct Lookahead {
  hipStream_t side = nullptr;
  hipEvent_t start = nullptr, panel_done = nullptr, rest_done = nullptr;
};

Lookahead* lookahead_for_device() {
  static Lookahead la[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  Lookahead& l = la[dev];
  if (!l.side) {
    if (hipStreamCreateWithFlags(&l.side, hipStreamNonBlocking) != hipSuccess) return nullptr;
    if (hipEventCreateWithFlags(&l.start, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&l.panel_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&l.rest_done, hipEventDisableTiming) != hipSuccess) {
      (void)hipStreamDestroy(l.side);
      l.side = nullptr;
      return nullptr;
    }
  }
  return &l;
}

}  // namespace

int* lu_status(void* work, int n) { return carve(work, n).status; }

constexpr size_t lu_workspace_bytes_c(int n) { return 256 + 2 * xchg_area_bytes(gcap_for(n)); }
size_t lu_workspace_bytes(int n) { return lu_workspace_bytes_c(n); }

int launch_mfma_probe(const double* a, const double* b, double* d, hipStream_t stream) {
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, stream, a, b, d);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

namespace {

// The mixed path's panel with block-local pivots: the row of column c's
// pivot is searched among the rows of the panel's nb x nb diagonal block
// only (first maximum of |z|, as elsewhere), so no workgroup has to see the
// rows below.  The factors are those of P A with P permuting inside 64-row
// blocks; they only precondition the FP64 refinement (which checks the
// result and falls back to partial pivoting when it does not converge).
// One workgroup factors the diagonal block in LDS and records piv / perm as
// the partial-pivot panel does.
constexpr int DB_THREADS = 256;
constexpr int DB_WAVES = DB_THREADS / 64;
// Value of lane src (uniform) of a wave, for 4- and 8-byte types.
template <typename T>
__device__ __forceinline__ T lane_value(T v, int src) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

// 1024 threads: lanes over columns, 16 waves over rows (4 rows each), so the
// sequential column loop spends little time per wave.  The active rows
// (below the current column) alternate between two LDS buffers: column c
// reads buffer c % 2 and writes its updated rows to the other one, so reads
// and writes need no barrier between them; the finished row c (the pivot row)
// goes to a third buffer.  After the one barrier per column, every wave
// searches the next column's pivot itself (first maximum of |z|^2 over the
// active rows, a preconditioner needs no more), so all waves agree on it
// without another barrier.
// Maximum over the wave of values that are >= 0 or the sentinel -1 (their
// bit patterns order as signed integers), by DPP row shifts and row
// broadcasts: lane 63 ends with the maximum, which readlane makes uniform.
// A few cycles per step, where a shuffle tree takes an LDS round trip each.
__device__ __forceinline__ float wave_max_nonneg(float v) {
  int x = __float_as_int(v);
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x111, 0xf, 0xf, false));  // row_shr:1
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x112, 0xf, 0xf, false));  // row_shr:2
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x114, 0xf, 0xf, false));  // row_shr:4
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x118, 0xf, 0xf, false));  // row_shr:8
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x142, 0xa, 0xf, false));  // row_bcast:15
  x = max(x, __builtin_amdgcn_update_dpp(x, x, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(x, 63));
}

constexpr int DG_THREADS = 1024;
constexpr int DG_WAVES = DG_THREADS / 64;
constexpr int DG_ROWS = NB / DG_WAVES;  // rows per wave
template <typename T>
__global__ void __launch_bounds__(DG_THREADS) lu_diag_block_kernel(T* __restrict__ A, long long lda, int n, int k0,
                                                                    int nb, int* piv, int* perm, int* info) {
  constexpr int LD = NB + 1;
  // [buffer][row][col]: buffers 0 and 1 hold the active rows, 2 the finished ones
  __shared__ T sr[3][NB * LD], si[3][NB * LD];
  // perm[k0 ..] and piv[k0 ..] live here until the end: a global store
  // before a barrier would make every column wait for it to complete
  __shared__ int s_perm[NB], s_piv[NB];
  __shared__ int s_info;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const bool col = lane < nb;
  if (t < nb) s_perm[t] = perm[k0 + t];
  if (t == 0) s_info = 0;
  for (int r = wave; r < nb; r += DG_WAVES)
    if (col) {
      const T* v = A + 2 * ((size_t)(k0 + r) * lda + k0 + lane);
      sr[0][r * LD + lane] = v[0];
      si[0][r * LD + lane] = v[1];
    }
  __syncthreads();
  // first maximum of |z|^2 over rows c .. nb-1 of column c in buffer b
  // (wave-uniform result); a zero column gives -1
  auto search = [&](int b, int c) {
    float m = -1.0f;
    if (lane >= c && lane < nb) {
      const float x = (float)sr[b][lane * LD + c], y = (float)si[b][lane * LD + c];
      m = x * x + y * y;
    }
    const float mx = wave_max_nonneg(m);
    const unsigned long long hit = __ballot(m >= 0.0f && m == mx);
    return (mx > 0.0f && hit) ? (__ffsll((long long)hit) - 1) : -1;
  };
  int p = search(0, 0);
  for (int c = 0; c < nb; ++c) {
    const int cur = c & 1, nxt = cur ^ 1;
    const bool zero = p < 0;
    if (zero) p = c;  // zero column: no exchange, no division
    // pivot row (old row p), 1 / pivot
    const T ur = col ? sr[cur][p * LD + lane] : T(0), ui = col ? si[cur][p * LD + lane] : T(0);
    const T pr = sr[cur][p * LD + c], pi = si[cur][p * LD + c];
    const T den = pr * pr + pi * pi;
    const T inv = (den > T(0)) ? T(1) / den : T(0);
    const T ir = pr * inv, ii = -pi * inv;
    const bool right = col && lane > c, here = lane == c;
    const T vr = right ? ur : T(0), vi = right ? ui : T(0);
    // this wave's rows r = c+1+wave+16q (row p takes old row c)
#pragma unroll
    for (int q = 0; q < DG_ROWS; ++q) {
      const int r = c + 1 + wave + DG_WAVES * q;
      const int rc = min(r, nb - 1);  // clamped: a valid row, result unused
      const int src = (rc == p) ? c : rc;
      const T ar = sr[cur][src * LD + c], ai = si[cur][src * LD + c];
      const T xr = sr[cur][src * LD + lane], xi = si[cur][src * LD + lane];
      const T lr = ar * ir - ai * ii, li = ar * ii + ai * ir;
      if (r < nb && col) {
        sr[nxt][r * LD + lane] = here ? lr : xr - (lr * vr - li * vi);
        si[nxt][r * LD + lane] = here ? li : xi - (lr * vi + li * vr);
      }
    }
    if (wave == (c & (DG_WAVES - 1)) && col) {  // row c is final: the pivot row
      sr[2][c * LD + lane] = ur;
      si[2][c * LD + lane] = ui;
    }
    if (t == 64) {
      s_piv[c] = k0 + p;
      if (p != c) {
        const int tmp = s_perm[c];
        s_perm[c] = s_perm[p];
        s_perm[p] = tmp;
      }
      if (zero && s_info == 0) s_info = k0 + c + 1;
    }
    __syncthreads();
    if (c + 1 < nb) p = search(nxt, c + 1);
  }
  for (int r = wave; r < nb; r += DG_WAVES)
    if (col) {
      T* v = A + 2 * ((size_t)(k0 + r) * lda + k0 + lane);
      v[0] = sr[2][r * LD + lane];
      v[1] = si[2][r * LD + lane];
    }
  if (t < nb) {
    perm[k0 + t] = s_perm[t];
    piv[k0 + t] = s_piv[t];
  }
  if (t == 0 && s_info != 0 && *info == 0) *info = s_info;
}

// L21 = A21 U11^-1 for the rows below the diagonal block (in place): each
// wave holds L21_ROWS rows in registers, lane j holding column j, and runs
// the columns in order: l_c = a_c / u_cc (lane c's value, read by the whole
// wave), then a_j -= l_c u_cj for j > c.  Rows are independent, so the waves
// never wait for each other.
constexpr int L21_ROWS = 16;
template <typename T>
__global__ void __launch_bounds__(DB_THREADS) lu_l21_kernel(T* __restrict__ A, long long lda, int n, int k0, int nb) {
  constexpr int LD = NB + 1;
  __shared__ T ur[NB * LD], ui[NB * LD];  // U11
  __shared__ T dr[NB], di[NB];            // 1 / u_cc
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  for (int e = t; e < nb * NB; e += DB_THREADS) {
    const int r = e / NB, c = e % NB;
    if (c < nb) {
      const T* v = A + 2 * ((size_t)(k0 + r) * lda + k0 + c);
      ur[r * LD + c] = v[0];
      ui[r * LD + c] = v[1];
    }
  }
  __syncthreads();
  if (t < nb) {
    const T a = ur[t * LD + t], b = ui[t * LD + t];
    const T den = a * a + b * b;
    dr[t] = (den > T(0)) ? a / den : T(0);
    di[t] = (den > T(0)) ? -b / den : T(0);
  }
  __syncthreads();
  const int r0 = k0 + nb + (blockIdx.x * DB_WAVES + wave) * L21_ROWS;
  const bool col = lane < nb;
  T xr[L21_ROWS], xi[L21_ROWS];
#pragma unroll
  for (int q = 0; q < L21_ROWS; ++q) {
    xr[q] = xi[q] = T(0);
    if (col && r0 + q < n) {
      const T* v = A + 2 * ((size_t)(r0 + q) * lda + k0 + lane);
      xr[q] = v[0];
      xi[q] = v[1];
    }
  }
  for (int c = 0; c < nb; ++c) {
    const T er = dr[c], ei = di[c];
    // pivot-row entries right of c (zero elsewhere, so other lanes keep
    // their values) and lane c's slot for the multiplier: selects, no
    // branches, so the row arrays stay in place
    const bool right = col && lane > c, here = lane == c;
    const T u_r = right ? ur[c * LD + lane] : T(0), u_i = right ? ui[c * LD + lane] : T(0);
#pragma unroll
    for (int q = 0; q < L21_ROWS; ++q) {
      const T ar = lane_value(xr[q], c), ai = lane_value(xi[q], c);
      const T lr = ar * er - ai * ei, li = ar * ei + ai * er;
      const T nr = xr[q] - (lr * u_r - li * u_i), ni = xi[q] - (lr * u_i + li * u_r);
      xr[q] = here ? lr : nr;
      xi[q] = here ? li : ni;
    }
  }
#pragma unroll
  for (int q = 0; q < L21_ROWS; ++q)
    if (col && r0 + q < n) {
      T* v = A + 2 * ((size_t)(r0 + q) * lda + k0 + lane);
      v[0] = xr[q];
      v[1] = xi[q];
    }
}

// L21 = A21 U11^-1 on FP32 MFMA for nb = 64, one workgroup per 64-row tile
// (rows independent, so tiles never wait for each other).  U11 (upper, from
// the diagonal-block LU) is split into 4 x 4 blocks of 16; the workgroup
// inverts the four 16 x 16 diagonal blocks (one lane per column, back
// substitution), then each wave takes 16 rows of the tile through the column
// blocks in order:
//   T   = A_b - sum_{a<b} X_a U_ab     (MFMA, accumulators start from A_b)
//   X_b = T U_bb^-1                     (MFMA)
// so the sequential part is 4 short steps instead of 64 columns.  The FP64
// refinement that these factors precondition checks the result.
constexpr int L21M_SP = 65;  // LDS row stride (complex)
__global__ void __launch_bounds__(256) lu_l21_mfma_kernel(float* __restrict__ A, long long lda, int n, int k0) {
  __shared__ flt2_t sU[64 * L21M_SP];      // U11 [row][col]
  __shared__ flt2_t sX[64 * L21M_SP];      // the tile's rows [row][col]: A21, then L21
  __shared__ flt2_t sInv[4][16 * 17];      // inverses of the 16 x 16 diagonal blocks [row][col]
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int r0 = k0 + 64 + blockIdx.x * 64;
  const int mr = min(64, n - r0);
  for (int e = t; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    const float* u = A + 2 * ((size_t)(k0 + r) * lda + k0 + c);
    sU[r * L21M_SP + c] = (flt2_t){u[0], u[1]};
    flt2_t x = (flt2_t){0.0f, 0.0f};
    if (r < mr) {
      const float* v = A + 2 * ((size_t)(r0 + r) * lda + k0 + c);
      x = (flt2_t){v[0], v[1]};
    }
    sX[r * L21M_SP + c] = x;
  }
  __syncthreads();
  if (t < 64) {
    // column j of the inverse of diagonal block b: x_i = (delta_ij -
    // sum_{k>i} u_ik x_k) / u_ii for i = 15 .. 0 (zero below the diagonal)
    const int b = t >> 4, j = t & 15;
    const flt2_t* U = sU + (16 * b) * L21M_SP + 16 * b;
    float xr[16], xi[16];
#pragma unroll
    for (int i = 15; i >= 0; --i) {
      float sr = (i == j) ? 1.0f : 0.0f, si = 0.0f;
#pragma unroll
      for (int k = i + 1; k < 16; ++k) {
        const flt2_t u = U[i * L21M_SP + k];
        sr -= u.x * xr[k] - u.y * xi[k];
        si -= u.x * xi[k] + u.y * xr[k];
      }
      const flt2_t d = U[i * L21M_SP + i];
      const float den = d.x * d.x + d.y * d.y;
      const float inv = (den > 0.0f) ? 1.0f / den : 0.0f;
      const float ir = d.x * inv, ii = -d.y * inv;
      xr[i] = sr * ir - si * ii;
      xi[i] = sr * ii + si * ir;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) sInv[b][i * 17 + j] = (flt2_t){xr[i], xi[i]};
  }
  __syncthreads();
  flt2_t* X = sX + (16 * wave) * L21M_SP;  // this wave's 16 rows
  for (int b = 0; b < 4; ++b) {
    // accumulator register v of lane l holds D[4 (l / 16) + v][l % 16]
    float4_t tr, ti;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const flt2_t a = X[(4 * lk + v) * L21M_SP + 16 * b + li];
      tr[v] = a.x;
      ti[v] = a.y;
    }
    for (int kk = 0; kk < 16 * b; kk += 4) {
      const flt2_t x = X[li * L21M_SP + kk + lk];
      const flt2_t u = sU[(kk + lk) * L21M_SP + 16 * b + li];
      tr = __builtin_amdgcn_mfma_f32_16x16x4f32(-x.x, u.x, tr, 0, 0, 0);
      tr = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, u.y, tr, 0, 0, 0);
      ti = __builtin_amdgcn_mfma_f32_16x16x4f32(-x.x, u.y, ti, 0, 0, 0);
      ti = __builtin_amdgcn_mfma_f32_16x16x4f32(-x.y, u.x, ti, 0, 0, 0);
    }
    // T through LDS into the operand layout (this wave's rows only)
#pragma unroll
    for (int v = 0; v < 4; ++v) X[(4 * lk + v) * L21M_SP + 16 * b + li] = (flt2_t){tr[v], ti[v]};
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float4_t xr = (float4_t){0.0f, 0.0f, 0.0f, 0.0f}, xi = xr;
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      const flt2_t x = X[li * L21M_SP + 16 * b + kk + lk];
      const flt2_t w = sInv[b][(kk + lk) * 17 + li];
      xr = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, w.x, xr, 0, 0, 0);
      xr = __builtin_amdgcn_mfma_f32_16x16x4f32(-x.y, w.y, xr, 0, 0, 0);
      xi = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, w.y, xi, 0, 0, 0);
      xi = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, w.x, xi, 0, 0, 0);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int v = 0; v < 4; ++v) X[(4 * lk + v) * L21M_SP + 16 * b + li] = (flt2_t){xr[v], xi[v]};
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  for (int e = t; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    if (r < mr) {
      float* v = A + 2 * ((size_t)(r0 + r) * lda + k0 + c);
      const flt2_t x = sX[r * L21M_SP + c];
      v[0] = x.x;
      v[1] = x.y;
    }
  }
}

}  // namespace

int launch_mfma_probe_f32(const float* a, const float* b, float* d, hipStream_t stream) {
  hipLaunchKernelGGL(mfma_probe_f32_kernel, dim3(1), dim3(64), 0, stream, a, b, d);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

namespace {

template <typename T>
int launch_lu(T* A, long long lda, int n, int* piv, int* info, void* work, hipStream_t stream, bool local_pivot,
              bool allow_lookahead = true) {
  constexpr int RMAX_T = rmax_of<T>();
  static bool attr_set = false;
  static int cu_count = 0;
  static int panel_capacity = 0;  // panel workgroups that can be resident at once
  const size_t lds_bytes = (size_t)(2 * RMAX_T * (NB + 1) + 2 * NB) * sizeof(T);
  if (!attr_set) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(lu_panel_kernel<T>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess)
      return 1;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, lu_panel_kernel<T>, PANEL_THREADS, lds_bytes) !=
        hipSuccess)
      return 1;
    panel_capacity = occ;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 1;
    panel_capacity *= cu_count;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(zgemm_tile_kernel<T>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)zt_lds_bytes<T>()) != hipSuccess)
      return 1;
    attr_set = true;
  }
  if (n <= 0) return 0;
  const int gemm_mode = zgemm_mode();
  // buffer offsets of a 64-row tile stay below 2^31 bytes
  if ((long long)lda * 2 * (long long)sizeof(T) * 64 >= (1ll << 31)) return 2;
  // the cooperative panel's workgroups wait on each other, so all of them
  // must fit on the device at once (one per CU by LDS: the occupancy query
  // says how many); the first (largest) panel bounds n
  const int G0 = (n + RMAX_T - 1) / RMAX_T;
  if (G0 > GMAX || G0 > cu_count || G0 > panel_capacity) return 2;
  int* perm = piv + n;
  const LuWork w = carve(work, n);
  const int npanels = (n + NB - 1) / NB;
  launch_k(K_LU_INIT, lu_init_kernel, dim3(std::max((n + 255) / 256, 64)), dim3(256), 0, stream, perm, n, w.status,
           w.ticket, info, w.xbase, w.Gcap, G0);
  const int forced = panel_stride_override();
  Lookahead* la = (allow_lookahead && npanels > 1 && lookahead_enabled()) ? lookahead_for_device() : nullptr;
  auto panel = [&](int k0, hipStream_t s) {
    const int nb = std::min(NB, n - k0);
    if constexpr (sizeof(T) == 4) {  // block-local pivots: the FP32 path only
      if (local_pivot) {
        launch_k(K_LU_PANEL, lu_diag_block_kernel<T>, dim3(1), dim3(DG_THREADS), 0u, s, A, lda, n, k0, nb, piv,
                 perm, info);
        const int below = n - k0 - nb;
        constexpr int rows_per_block = DB_WAVES * L21_ROWS;
        if (below > 0) {
          if (nb == NB && l21_mfma_enabled())
            launch_k(K_LU_PANEL, lu_l21_mfma_kernel, dim3((below + 63) / 64), dim3(256), 0u, s, A, lda, n, k0);
          else
            launch_k(K_LU_PANEL, lu_l21_kernel<T>, dim3((below + rows_per_block - 1) / rows_per_block),
                     dim3(DB_THREADS), 0u, s, A, lda, n, k0, nb);
        }
        return;
      }
    }
    const int rows = n - k0;
    const int G = (rows + RMAX_T - 1) / RMAX_T;
    const int R = (rows + G - 1) / G;
    // Packing the workgroups onto one XCD (stride 8) measured slower than
    // spreading them, so the default is 1; NECGPU_PANEL_STRIDE overrides.
    int stride = 1;
    if (forced > 0 && forced * G <= cu_count) stride = forced;
    const int Gnext = (k0 + nb < n) ? (n - k0 - nb + RMAX_T - 1) / RMAX_T : 0;
    launch_k(K_LU_PANEL, lu_panel_kernel<T>, dim3(G * stride), dim3(PANEL_THREADS), (uint32_t)lds_bytes, s, A, lda, n,
             k0, nb, G, R, stride, piv, info, w.xbase, w.Gcap, (k0 / NB) & 1, Gnext, w.status);
  };
  // row exchanges of panel k0 on columns [c0, c1) and [c2, c3); with_perm:
  // this launch also applies them to the composed permutation (the cooperative
  // panel leaves that to one launch per panel; the block-local one keeps it)
  auto laswp = [&](int k0, int nb, int c0, int c1, int c2, int c3, hipStream_t s, bool with_perm) {
    const int cols = (c1 - c0) + (c3 - c2);
    int* pm = (with_perm && !local_pivot) ? perm : nullptr;
    if (cols > 0 || pm)
      launch_k(K_LU_LASWP, lu_laswp_kernel<T>, dim3(std::max(1, (cols + LASWP_COLS - 1) / LASWP_COLS)), dim3(256), 0, s,
               A, lda, k0, nb, (const int*)piv, c0, c1, c2, c3, pm);
  };
  auto update = [&](int k0, int nb, int cbeg, int cend, int max_groups, hipStream_t s, bool tile_gemm) {
    // trailing rows [k0+nb, n), columns [cbeg, cend)
    const int cols = cend - cbeg, rows = n - k0 - nb;
    if (cols <= 0) return;
    launch_k(K_LU_TRSM, lu_trsm_kernel<T>, dim3((cols + 63) / 64), dim3(256), 0, s, A, lda, cend, k0, nb, cbeg);
    if (rows <= 0) return;
    const int tiles = ((rows + 63) / 64) * ((cols + 63) / 64);
    if (tile_gemm) {
      // persistent: as many workgroups as stay resident (LDS-bound)
      const int resident = ZtTraits<T>::kPerCU * cu_count;
      const int grid = std::max(1, std::min(tiles, std::min(max_groups, resident)));
      launch_k(K_ZGEMM, zgemm_tile_kernel<T>, dim3(grid), dim3(256), (uint32_t)zt_lds_bytes<T>(), s, A, lda, k0 + nb,
               cbeg, rows, cols, k0, nb);
      return;
    }
    const int grid = std::max(1, std::min(tiles, max_groups));
    if constexpr (sizeof(T) == 8)
      launch_k(K_ZGEMM, zgemm_update_kernel, dim3(grid), dim3(256), 0, s, A, lda, k0 + nb, cbeg, rows, cols, k0, nb);
    else
      launch_k(K_ZGEMM, zgemm_update_f32_kernel, dim3(grid), dim3(256), 0, s, A, lda, k0 + nb, cbeg, rows, cols, k0,
               nb);
  };
  // The next panel's columns: one fused launch while its 64-row tiles fit
  // the CUs the trailing update leaves free (each tile workgroup takes a
  // CU's LDS), else the trsm and the MFMA update as two launches.
  const int fused_max = next_fused_max_tiles();
  auto next_update = [&](int k0, int nb, int cbeg, int cend, hipStream_t s) {
    const int rows = n - k0 - nb;
    const int tiles = std::max(1, (rows + 63) / 64);
    if (tiles <= fused_max)
      launch_k(K_LU_NEXT, lu_next_update_kernel<T>, dim3(tiles), dim3(256), 0, s, A, lda, n, k0, nb, cbeg, cend,
               w.ticket);
    else
      update(k0, nb, cbeg, cend, INT_MAX, s, gemm_mode == 1);
  };
  // Right-looking LU on one stream while the panels are wide (or without
  // look-ahead), then the look-ahead below from panel kla on.
  const int la_max_g = lookahead_max_groups();
  auto groups = [&](int k0) { return (n - k0 + RMAX_T - 1) / RMAX_T; };
  int kla = n;
  panel(0, stream);
  if (local_pivot) la = nullptr;  // cheap panels: one stream
  if (la && groups(0) <= la_max_g) {
    kla = 0;
  } else {
    // Panels in pairs while both stay on the one stream (and the tile
    // kernel does the update): after panel a, only the next panel's columns
    // are updated; after panel b, U12 of both (rows of a by L_aa^-1, rows
    // of b by L_bb^-1 after subtracting L_ba U12_a) and one trailing update
    // 128 deep, which reads and writes the trailing matrix once per pair of
    // panels instead of once per panel.  NECGPU_LU_PAIRS=0 keeps one update
    // per panel, for measurements.
    const bool pairs = gemm_mode != 0 && lu_pairs_enabled();
    const int resident = ZtTraits<T>::kPerCU * cu_count;
    auto tile_update = [&](int m0, int n0, int M, int N, int kk, int K) {
      if (M <= 0 || N <= 0) return;
      const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
      launch_k(K_ZGEMM, zgemm_tile_kernel<T>, dim3(std::max(1, std::min(tiles, resident))), dim3(256),
               (uint32_t)zt_lds_bytes<T>(), stream, A, lda, m0, n0, M, N, kk, K);
    };
    auto trsm = [&](int kk, int nb, int cbeg, int cend) {
      if (cend > cbeg)
        launch_k(K_LU_TRSM, lu_trsm_kernel<T>, dim3((cend - cbeg + 63) / 64), dim3(256), 0, stream, A, lda, cend, kk,
                 nb, cbeg);
    };
    for (int k0 = 0; k0 < n;) {
      const int nb = std::min(NB, n - k0);
      const int k1 = k0 + nb;
      laswp(k0, nb, 0, k0, k1, n, stream, true);
      if (k1 >= n) break;
      const bool b_on_stream = !(la && groups(k1) <= la_max_g);
      const int nb1 = std::min(NB, n - k1);
      const int k2 = k1 + nb1;
      if (!pairs || !b_on_stream || k2 >= n) {
        update(k0, nb, k1, n, INT_MAX, stream, gemm_mode != 0);
        if (!b_on_stream) {
          kla = k1;
          break;
        }
        panel(k1, stream);
        k0 = k1;
        continue;
      }
      next_update(k0, nb, k1, k2, stream);  // panel b's columns
      panel(k1, stream);
      laswp(k1, nb1, 0, k1, k2, n, stream, true);
      trsm(k0, nb, k2, n);                       // U12_a
      tile_update(k1, k2, nb1, n - k2, k0, nb);  // rows of b: -= L_ba U12_a
      trsm(k1, nb1, k2, n);                      // U12_b
      tile_update(k2, k2, n - k2, n - k2, k0, nb + nb1);
      if (la && groups(k2) <= la_max_g) {
        kla = k2;
        break;
      }
      panel(k2, stream);
      k0 = k2;
    }
  }
  if (kla >= n) return hipGetLastError() == hipSuccess ? 0 : 1;
  // Right-looking LU with a look-ahead of one panel, on two streams:
  //   side stream:   panel p -> [rest of p-1 done] row exchanges of the next
  //                  panel's columns -> their U12 and trailing update (one
  //                  launch) -> panel p+1 -> ...
  //   caller stream: [panel p done] row exchanges, U12 and trailing update
  //                  of the columns beyond the next panel, on at most
  //                  occupancy x (CUs - next panel's workgroups) workgroups
  //                  so that the panel finds free CUs -> rest_done.
  // The critical path is panels and the next-columns work on one stream;
  // the two streams never touch the same columns at the same time.
  if (hipEventRecord(la->start, stream) != hipSuccess) return 1;
  if (hipStreamWaitEvent(la->side, la->start, 0) != hipSuccess) return 1;
  if (kla > 0) panel(kla, la->side);
  if (hipEventRecord(la->panel_done, la->side) != hipSuccess) return 1;
  for (int k0 = kla; k0 < n; k0 += NB) {
    const int nb = std::min(NB, n - k0);
    const int next = k0 + nb;
    if (hipStreamWaitEvent(stream, la->panel_done, 0) != hipSuccess) return 1;
    if (next >= n) {
      laswp(k0, nb, 0, k0, n, n, stream, true);
      break;
    }
    const int far = next + std::min(NB, n - next);
    if (k0 > kla && hipStreamWaitEvent(la->side, la->rest_done, 0) != hipSuccess) return 1;
    laswp(k0, nb, next, far, far, far, la->side, false);
    next_update(k0, nb, next, far, la->side);
    panel(next, la->side);
    if (hipEventRecord(la->panel_done, la->side) != hipSuccess) return 1;
    laswp(k0, nb, 0, k0, far, n, stream, true);
    const int Gnext = (n - next + RMAX_T - 1) / RMAX_T;
    const int free_cus = std::max(1, cu_count - Gnext);
    const int tiles = ((n - next + 63) / 64) * ((n - far + 63) / 64);
    // the tile kernel for updates of 16 tiles or more per free CU (N =
    // 16,020: 246.6 against 250.6 ms for the FP64 factorisation), the
    // per-tile kernels for smaller ones (N = 3,280: 14.9 against 15.5 ms)
    const bool tile_k = gemm_mode == 1 || (gemm_mode == 2 && tiles >= 16 * free_cus);
    update(k0, nb, far, n, tile_k ? ZtTraits<T>::kPerCU * free_cus : gemm_occupancy(tiles, free_cus) * free_cus, stream,
           tile_k);
    if (hipEventRecord(la->rest_done, stream) != hipSuccess) return 1;
  }
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace

int launch_lu_fp64(double* A, long long lda, int n, int* piv, int* info, void* work, hipStream_t stream,
                   bool allow_lookahead) {
  return launch_lu<double>(A, lda, n, piv, info, work, stream, false, allow_lookahead);
}

int launch_lu_fp32(float* A, long long lda, int n, int* piv, int* info, void* work, hipStream_t stream,
                   int pivot_mode, bool allow_lookahead) {
  return launch_lu<float>(A, lda, n, piv, info, work, stream, pivot_mode == 1, allow_lookahead);
}

// ---------------------------------------------------------------------
// The FP64 LU without any allocation, for callers that cannot allocate
// device memory: the workspace (status word, panel exchange slots) is the
// file-scope array below, of the size the largest supported order needs.
// Each device holds its own copy for the life of the process; its address
// on the current device comes from hipGetSymbolAddress once per device (the
// array's name in host code is only a stand-in, never a device address).
// lu_init_kernel clears what must start at zero on every call, so a call
// does not depend on the previous call's contents, but two calls that
// overlap on one device share the array and race.  Everything runs on the
// caller's stream (no look-ahead side stream is created).
namespace {
constexpr size_t kNoAllocBytes = lu_workspace_bytes_c(kNoAllocMaxN);
static __device__ __align__(256) char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// the current device's copy of `scratch`, looked up once per device
int scratch_on_current_device(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 1;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return 1;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

size_t lu_noalloc_scratch_bytes() { return kNoAllocBytes; }

int launch_lu_fp64_noalloc(double* A, long long lda, int n, int* piv, int* info, hipStream_t stream) {
  if (n > kNoAllocMaxN) return 3;
  void* work = nullptr;
  if (scratch_on_current_device(&work)) return 4;
  return launch_lu<double>(A, lda, n, piv, info, work, stream, false, false);
}

}  // namespace necgpu
