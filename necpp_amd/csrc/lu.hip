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

#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

constexpr int NB = 64;
constexpr int PANEL_THREADS = 512;
constexpr int GMAX = 256;     // most cooperating panel workgroups
constexpr int RMAX = 112;     // most panel rows per workgroup (LDS: RMAX*(NB+1)*16 B)

typedef double double4_t __attribute__((ext_vector_type(4)));

// Exchange area of the cooperative panel (double-buffered by column parity).
struct PanelXchg {
  double cand_val[2][GMAX];
  int cand_idx[2][GMAX];
  double cand_row[2][GMAX][2 * NB];
  double diag_row[2][2 * NB];
};

struct LuWork {
  unsigned* arrivals;  // one counter per panel
  PanelXchg* x;
  int* status;
};

__device__ __forceinline__ void xstore(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double xload(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool better(double ob, int oi, double best, int idx) {
  return ob > best || (ob == best && oi < idx);
}

__global__ void lu_init_kernel(int* perm, int n, unsigned* arrivals, int npanels, int* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) perm[i] = i;
  if (i < npanels) arrivals[i] = 0u;
  if (i == 0) *status = 0;
}

// Factor panel columns [k0, k0+nb) over rows [k0, n) with G workgroups; the
// active workgroup g is launch block g*stride.  Rows owned by g:
// [k0 + g*R, min(n, k0 + (g+1)*R)).
//
// Column c (k = k0 + c):
//  A  block reduction of the column-c candidates (|z| from step E).
//  B  wave 0 applies the update still pending from column c-1 (columns
//     c+1..) to the candidate row and to row k, publishes both with the
//     candidate and runs the exchange; the other waves meanwhile apply that
//     pending update to every other owned row, which takes the bulk of the
//     rank-1 work off the exchange's critical path.
//  D  the pivot row is fetched; rows k and p are exchanged.
//  E  one thread per row: scale column c, update column c+1 only, and form
//     the column-(c+1) candidate; the update of columns c+2.. is left to B.
__global__ void __launch_bounds__(PANEL_THREADS) lu_panel_kernel(double* __restrict__ A, long long lda, int n, int k0,
                                                                  int nb, int G, int R, int stride, int* piv, int* perm,
                                                                  int* info, unsigned* arrival, PanelXchg* X,
                                                                  int* status) {
  extern __shared__ double lds[];
  if (blockIdx.x % stride != 0) return;
  constexpr int NW = PANEL_THREADS / 64;  // waves
  const int g = blockIdx.x / stride;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int LDR = NB + 1;  // padded complex row length in LDS
  double* rows = lds;                         // [R][LDR] complex
  double* su = lds + 2 * (size_t)RMAX * LDR;  // current pivot row, nb complex
  __shared__ double s_best[NW];
  __shared__ int s_idx[NW];
  __shared__ int s_win[2];  // winner row, winner group
  const int r0 = k0 + g * R;
  const int r1 = min(n, r0 + R);
  const int nrow = max(0, r1 - r0);
  for (int e = t; e < nrow * nb; e += PANEL_THREADS) {
    const int r = e / nb, c = e % nb;
    const double* v = A + 2 * ((size_t)(r0 + r) * lda + k0 + c);
    rows[2 * (r * LDR + c)] = v[0];
    rows[2 * (r * LDR + c) + 1] = v[1];
  }
  __syncthreads();
  // candidates of column 0
  double best = -1.0;
  int bidx = INT_MAX;
  for (int r = t; r < nrow; r += PANEL_THREADS) {
    const double h = hypot(rows[2 * (r * LDR)], rows[2 * (r * LDR) + 1]);
    if (h > best) {
      best = h;
      bidx = r0 + r;
    }
  }
  for (int c = 0; c < nb; ++c) {
    const int k = k0 + c;
    const int par = c & 1;
    // A. block reduction of the candidates
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_down(best, off, 64);
      const int oi = __shfl_down(bidx, off, 64);
      if (better(ob, oi, best, bidx)) {
        best = ob;
        bidx = oi;
      }
    }
    if (lane == 0) {
      s_best[wave] = best;
      s_idx[wave] = bidx;
    }
    __syncthreads();
    best = s_best[0];
    bidx = s_idx[0];
    for (int w = 1; w < NW; ++w)
      if (better(s_best[w], s_idx[w], best, bidx)) {
        best = s_best[w];
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
          rows[2 * (r * LDR + lane)] -= prod.re;
          rows[2 * (r * LDR + lane) + 1] -= prod.im;
        }
      }
      // The exchange area is only touched with agent-scope atomics, which
      // the hardware performs at the device coherence point, so no cache
      // write-back / invalidate fences are needed: the data stores are
      // complete (s_waitcnt vmcnt(0)) before the arrival is counted.
      if (lane == 0) {
        xstore(&X->cand_val[par][g], best);
        __hip_atomic_store(&X->cand_idx[par][g], bidx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (rcand >= 0 && lane < nb) {
        xstore(&X->cand_row[par][g][2 * lane], rows[2 * (rcand * LDR + lane)]);
        xstore(&X->cand_row[par][g][2 * lane + 1], rows[2 * (rcand * LDR + lane) + 1]);
      }
      if (rk >= 0 && lane < nb) {
        xstore(&X->diag_row[par][2 * lane], rows[2 * (rk * LDR + lane)]);
        xstore(&X->diag_row[par][2 * lane + 1], rows[2 * (rk * LDR + lane) + 1]);
      }
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      if (lane == 0) {
        __hip_atomic_fetch_add(arrival, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // bounded wait; a timeout is sticky so every later wait returns at once
        const unsigned target = (unsigned)G * (unsigned)(c + 1);
        long spins = 0;
        while (__hip_atomic_load(arrival, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
          if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
          if (++spins > (1L << 22)) {
            __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      // winner over all groups (every group computes the same)
      double wv = -1.0;
      int wi = INT_MAX, wg = 0;
      for (int h = lane; h < G; h += 64) {
        const double v = xload(&X->cand_val[par][h]);
        const int ix = __hip_atomic_load(&X->cand_idx[par][h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (better(v, ix, wv, wi)) {
          wv = v;
          wi = ix;
          wg = h;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_down(wv, off, 64);
        const int oi = __shfl_down(wi, off, 64);
        const int og = __shfl_down(wg, off, 64);
        if (better(ov, oi, wv, wi)) {
          wv = ov;
          wi = oi;
          wg = og;
        }
      }
      if (lane == 0) {
        int p = wi;
        if (p == INT_MAX || !(wv > 0.0)) {
          if (wv == 0.0 && g == 0 && *info == 0) *info = k + 1;
          p = k;
          wg = -1;
        }
        s_win[0] = p;
        s_win[1] = wg;
        if (g == 0) {
          piv[k] = p;
          if (p != k) {
            const int tmp = perm[k];
            perm[k] = perm[p];
            perm[p] = tmp;
          }
        }
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
            rows[2 * (rr * LDR + pend_lo + lane)] = ar[q] - prod.re;
            rows[2 * (rr * LDR + pend_lo + lane) + 1] = ai[q] - prod.im;
          }
        }
      }
    }
    __syncthreads();
    const int p = s_win[0];
    const int wg = s_win[1];
    // D. the pivot row (whole panel row) and the exchange of rows k and p
    if (t < nb) {
      double ur, ui;
      if (p != k) {
        ur = xload(&X->cand_row[par][wg][2 * t]);
        ui = xload(&X->cand_row[par][wg][2 * t + 1]);
      } else if (rk >= 0) {
        ur = rows[2 * (rk * LDR + t)];
        ui = rows[2 * (rk * LDR + t) + 1];
      } else {
        ur = xload(&X->diag_row[par][2 * t]);
        ui = xload(&X->diag_row[par][2 * t + 1]);
      }
      su[2 * t] = ur;
      su[2 * t + 1] = ui;
    }
    __syncthreads();
    if (p != k && t < nb) {
      if (p >= r0 && p < r1) {
        rows[2 * ((p - r0) * LDR + t)] = xload(&X->diag_row[par][2 * t]);
        rows[2 * ((p - r0) * LDR + t) + 1] = xload(&X->diag_row[par][2 * t + 1]);
      }
      if (rk >= 0) {
        rows[2 * (rk * LDR + t)] = su[2 * t];
        rows[2 * (rk * LDR + t) + 1] = su[2 * t + 1];
      }
    }
    __syncthreads();
    // E. one thread per row below k: scale column c (Smith's division as
    //    g++'s complex divide, with the pivot's ratio and denominator formed
    //    once), update column c+1, and keep the column-(c+1) candidate.
    const zc d = zmk(su[2 * c], su[2 * c + 1]);
    const bool scale = (d.re != 0.0) || (d.im != 0.0);
    const bool swp = fabs(d.re) < fabs(d.im);
    const double ratio = swp ? d.re / d.im : d.im / d.re;
    const double denom = swp ? (d.re * ratio) + d.im : (d.im * ratio) + d.re;
    const bool next = c + 1 < nb;
    const zc u1 = next ? zmk(su[2 * (c + 1)], su[2 * (c + 1) + 1]) : zmk(0.0, 0.0);
    best = -1.0;
    bidx = INT_MAX;
    for (int r = max(0, k + 1 - r0) + t; r < nrow; r += PANEL_THREADS) {
      double* v = rows + 2 * (r * LDR + c);
      double lr = v[0], li = v[1];
      if (scale) {
        const double ar = lr, ai = li;
        if (swp) {
          lr = ((ar * ratio) + ai) / denom;
          li = ((ai * ratio) - ar) / denom;
        } else {
          lr = ((ai * ratio) + ar) / denom;
          li = (ai - (ar * ratio)) / denom;
        }
        v[0] = lr;
        v[1] = li;
      }
      if (next) {
        const zc prod = zmul(zmk(lr, li), u1);
        const double nr = v[2] - prod.re;
        const double ni = v[3] - prod.im;
        v[2] = nr;
        v[3] = ni;
        const double h = hypot(nr, ni);
        if (h > best) {
          best = h;
          bidx = r0 + r;
        }
      }
    }
    __syncthreads();
  }
  for (int e = t; e < nrow * nb; e += PANEL_THREADS) {
    const int r = e / nb, c = e % nb;
    double* v = A + 2 * ((size_t)(r0 + r) * lda + k0 + c);
    v[0] = rows[2 * (r * LDR + c)];
    v[1] = rows[2 * (r * LDR + c) + 1];
  }
}

// Apply the exchanges piv[k0 .. k0+nb) to columns [0, k0) and [k0+nb, n).
__global__ void __launch_bounds__(256) lu_laswp_kernel(double* __restrict__ A, long long lda, int n, int k0, int nb,
                                                       const int* __restrict__ piv) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int outside = n - nb;
  if (idx >= outside) return;
  const int col = idx < k0 ? idx : idx + nb;
  for (int q = 0; q < nb; ++q) {
    const int k = k0 + q;
    const int p = piv[k];
    if (p != k) {
      double* a = A + 2 * ((size_t)k * lda + col);
      double* b = A + 2 * ((size_t)p * lda + col);
      const double ar = a[0], ai = a[1];
      a[0] = b[0];
      a[1] = b[1];
      b[0] = ar;
      b[1] = ai;
    }
  }
}

// U12 = L11^-1 A12: rows [k0, k0+nb), columns [c0, c0+64) of the trailing part.
__global__ void __launch_bounds__(256) lu_trsm_kernel(double* __restrict__ A, long long lda, int n, int k0, int nb) {
  __shared__ double sL[NB * NB * 2];  // L11[r][q], complex
  __shared__ double sX[NB * 64 * 2];  // X[r][c], complex
  const int t = threadIdx.x;
  const int c0 = k0 + nb + blockIdx.x * 64;
  for (int e = t; e < nb * nb; e += 256) {
    const int r = e / nb, q = e % nb;
    const double* v = A + 2 * ((size_t)(k0 + r) * lda + k0 + q);
    sL[2 * (r * NB + q)] = v[0];
    sL[2 * (r * NB + q) + 1] = v[1];
  }
  for (int e = t; e < nb * 64; e += 256) {
    const int r = e / 64, c = e % 64;
    double vr = 0.0, vi = 0.0;
    if (c0 + c < n) {
      const double* v = A + 2 * ((size_t)(k0 + r) * lda + c0 + c);
      vr = v[0];
      vi = v[1];
    }
    sX[2 * (r * 64 + c)] = vr;
    sX[2 * (r * 64 + c) + 1] = vi;
  }
  __syncthreads();
  const int c = t & 63;
  const int rg = t >> 6;  // row group 0..3
  for (int q = 0; q < nb - 1; ++q) {
    const zc xq = zmk(sX[2 * (q * 64 + c)], sX[2 * (q * 64 + c) + 1]);
    for (int r = q + 1 + ((rg - (q + 1)) & 3); r < nb; r += 4) {
      const zc l = zmk(sL[2 * (r * NB + q)], sL[2 * (r * NB + q) + 1]);
      const zc prod = zmul(l, xq);
      sX[2 * (r * 64 + c)] -= prod.re;
      sX[2 * (r * 64 + c) + 1] -= prod.im;
    }
    __syncthreads();
  }
  for (int e = t; e < nb * 64; e += 256) {
    const int r = e / 64, cc = e % 64;
    if (c0 + cc < n) {
      double* v = A + 2 * ((size_t)(k0 + r) * lda + c0 + cc);
      v[0] = sX[2 * (r * 64 + cc)];
      v[1] = sX[2 * (r * 64 + cc) + 1];
    }
  }
}

// C[m0+i][n0+j] -= sum_q A[m0+i][k0+q] * A[k0+q][n0+j], i < M, j < N, q < K.
__global__ void __launch_bounds__(256) zgemm_update_kernel(double* __restrict__ A, long long lda, int m0, int n0, int M,
                                                           int N, int k0, int K) {
  __shared__ double sLr[NB * 64], sLi[NB * 64];  // [q][row]
  __shared__ double sUr[NB * 64], sUi[NB * 64];  // [q][col]
  const int t = threadIdx.x;
  const int row0 = blockIdx.y * 64, col0 = blockIdx.x * 64;
  const int Kp = (K + 3) & ~3;
  for (int e = t; e < 64 * Kp; e += 256) {
    const int r = e / Kp, q = e % Kp;
    double vr = 0.0, vi = 0.0;
    if (row0 + r < M && q < K) {
      const double* v = A + 2 * ((size_t)(m0 + row0 + r) * lda + k0 + q);
      vr = v[0];
      vi = v[1];
    }
    sLr[q * 64 + r] = vr;
    sLi[q * 64 + r] = vi;
  }
  for (int e = t; e < 64 * Kp; e += 256) {
    const int q = e / 64, c = e % 64;
    double vr = 0.0, vi = 0.0;
    if (col0 + c < N && q < K) {
      const double* v = A + 2 * ((size_t)(k0 + q) * lda + n0 + col0 + c);
      vr = v[0];
      vi = v[1];
    }
    sUr[q * 64 + c] = vr;
    sUi[q * 64 + c] = vi;
  }
  __syncthreads();
  const int lane = t & 63, wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  double4_t accr[2][2], acci[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      accr[a][b] = (double4_t){0.0, 0.0, 0.0, 0.0};
      acci[a][b] = (double4_t){0.0, 0.0, 0.0, 0.0};
    }
  for (int kk = 0; kk < Kp; kk += 4) {
    const int q = kk + lk;
    double ar[2], ai[2], br[2], bi[2];
    for (int a = 0; a < 2; ++a) {
      ar[a] = sLr[q * 64 + wr * 32 + a * 16 + li];
      ai[a] = sLi[q * 64 + wr * 32 + a * 16 + li];
    }
    for (int b = 0; b < 2; ++b) {
      br[b] = sUr[q * 64 + wc * 32 + b * 16 + li];
      bi[b] = sUi[q * 64 + wc * 32 + b * 16 + li];
    }
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], accr[a][b], 0, 0, 0);
        accr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], -bi[b], accr[a][b], 0, 0, 0);
        acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], acci[a][b], 0, 0, 0);
        acci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], acci[a][b], 0, 0, 0);
      }
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
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

// One wave: lane l feeds a[l] as the A operand and b[l] as the B operand of
// v_mfma_f64_16x16x4f64 and writes its four result registers to d[4*l..].
__global__ void mfma_probe_kernel(const double* a, const double* b, double* d) {
  const int l = threadIdx.x;
  double4_t acc = (double4_t){0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[l], b[l], acc, 0, 0, 0);
  for (int v = 0; v < 4; ++v) d[4 * l + v] = acc[v];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

LuWork carve(void* work, int n) {
  const int npanels = (n + NB - 1) / NB;
  char* p = static_cast<char*>(work);
  LuWork w;
  w.arrivals = reinterpret_cast<unsigned*>(p);
  p += align256((size_t)npanels * sizeof(unsigned));
  w.status = reinterpret_cast<int*>(p);
  p += 256;
  w.x = reinterpret_cast<PanelXchg*>(p);
  return w;
}

int panel_stride_override() {
  const char* v = std::getenv("NECGPU_PANEL_STRIDE");
  return (v && *v) ? std::atoi(v) : 0;
}

}  // namespace

int* lu_status(void* work, int n) { return carve(work, n).status; }

size_t lu_workspace_bytes(int n) {
  const int npanels = (n + NB - 1) / NB;
  return align256((size_t)npanels * sizeof(unsigned)) + 256 + align256(sizeof(PanelXchg));
}

int launch_mfma_probe(const double* a, const double* b, double* d, hipStream_t stream) {
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, stream, a, b, d);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_lu_fp64(double* A, long long lda, int n, int* piv, int* info, void* work, hipStream_t stream) {
  static bool attr_set = false;
  static int cu_count = 0;
  const size_t lds_bytes = (size_t)(2 * RMAX * (NB + 1) + 2 * NB) * sizeof(double);
  if (!attr_set) {
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(lu_panel_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess)
      return 1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 1;
    attr_set = true;
  }
  int* perm = piv + n;
  const LuWork w = carve(work, n);
  const int npanels = (n + NB - 1) / NB;
  launch_k(K_LU_INIT, lu_init_kernel, dim3((std::max(n, npanels) + 255) / 256), dim3(256), 0, stream, perm, n,
                     w.arrivals, npanels, w.status);
  const int forced = panel_stride_override();
  for (int k0 = 0, pi = 0; k0 < n; k0 += NB, ++pi) {
    const int nb = (n - k0 < NB) ? (n - k0) : NB;
    const int rows = n - k0;
    int G = (rows + RMAX - 1) / RMAX;
    // every active workgroup needs its own CU at the same time (one per CU
    // by LDS), so G may not exceed the CU count
    if (G > GMAX || G > cu_count) return 2;
    const int R = (rows + G - 1) / G;
    // Packing the workgroups onto one XCD (stride 8) measured slower than
    // spreading them, so the default is 1; NECGPU_PANEL_STRIDE overrides.
    int stride = 1;
    if (forced > 0 && forced * G <= cu_count) stride = forced;
    launch_k(K_LU_PANEL, lu_panel_kernel, dim3(G * stride), dim3(PANEL_THREADS), lds_bytes, stream, A, lda, n, k0, nb, G,
                       R, stride, piv, perm, info, w.arrivals + pi, w.x, w.status);
    if (n > nb)
      launch_k(K_LU_LASWP, lu_laswp_kernel, dim3((n - nb + 255) / 256), dim3(256), 0, stream, A, lda, n, k0, nb, piv);
    const int rest = n - k0 - nb;
    if (rest <= 0) break;
    launch_k(K_LU_TRSM, lu_trsm_kernel, dim3((rest + 63) / 64), dim3(256), 0, stream, A, lda, n, k0, nb);
    launch_k(K_ZGEMM, zgemm_update_kernel, dim3((rest + 63) / 64, (rest + 63) / 64), dim3(256), 0, stream, A, lda,
                       k0 + nb, k0 + nb, rest, rest, k0, nb);
  }
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace necgpu
