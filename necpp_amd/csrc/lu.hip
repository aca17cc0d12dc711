// lu.hip -- blocked right-looking complex LU with partial pivoting (FP64).
//
// Replaces lu_decompose / Eigen PartialPivLU (src/matrix_algebra.cpp:217-266,
// src/eigen/Eigen/src/LU/PartialPivLU.h:338-465).  The matrix is the fill's
// row-major Z (row i = observation segment), which is the reference's "cm"
// storage; pivoting exchanges whole rows, which are contiguous here.  Pivot
// rule: first maximum of |z| = hypot(re, im) in the column, as Eigen's
// scalar_score_coeff_op.
//
// Per panel of NB = 64 columns:
//   lu_panel   one workgroup (16 waves): column by column pivot search
//              (fused into the previous column's update), full-row swap,
//              column scaling by complex division, rank-1 update of the
//              panel; every wave walks whole rows so accesses are coalesced.
//   lu_trsm    U12 = L11^-1 A12 with L11 and a 64-column block of A12 in LDS.
//   zgemm_nt   A22 -= L21 * U12 on FP64 MFMA (v_mfma_f64_16x16x4f64), the
//              complex product in 4M form; 64x64 complex tile per workgroup,
//              operands staged in LDS as split re/im planes.
#include <climits>

#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

constexpr int NB = 64;
constexpr int PANEL_THREADS = 1024;

typedef double double4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void better(double& best, int& idx, double ob, int oi) {
  if (ob > best || (ob == best && oi < idx)) {
    best = ob;
    idx = oi;
  }
}

__global__ void init_perm_kernel(int* perm, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) perm[i] = i;
}

// Factor columns [k0, k0+nb) over rows [k0, n).  Row exchanges are applied to
// the whole row (all n columns) and recorded in piv (transpositions) and in
// perm (the composed permutation, perm[i] = original row now at i).
__global__ void __launch_bounds__(PANEL_THREADS) lu_panel_kernel(double* __restrict__ A, long long lda, int n, int k0,
                                                                  int nb, int* __restrict__ piv, int* __restrict__ perm,
                                                                  int* __restrict__ info) {
  __shared__ double s_best[PANEL_THREADS / 64];
  __shared__ int s_idx[PANEL_THREADS / 64];
  __shared__ int s_piv;
  __shared__ double s_prow[2 * NB];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int nwaves = PANEL_THREADS / 64;

  // pivot candidates of the first panel column
  double best = -1.0;
  int bidx = INT_MAX;
  for (int i = k0 + t; i < n; i += PANEL_THREADS) {
    const double* v = A + 2 * ((size_t)i * lda + k0);
    const double h = hypot(v[0], v[1]);
    if (h > best) {
      best = h;
      bidx = i;
    }
  }
  for (int c = 0; c < nb; ++c) {
    const int k = k0 + c;
    // reduce (best, bidx) over the workgroup: larger wins, ties to lower row
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_down(best, off, 64);
      const int oi = __shfl_down(bidx, off, 64);
      better(best, bidx, ob, oi);
    }
    if (lane == 0) {
      s_best[wave] = best;
      s_idx[wave] = bidx;
    }
    __syncthreads();
    if (t == 0) {
      double b = s_best[0];
      int p = s_idx[0];
      for (int w = 1; w < nwaves; ++w) better(b, p, s_best[w], s_idx[w]);
      if (p == INT_MAX || !(b > 0.0)) {
        if (b == 0.0 && *info == 0) *info = k + 1;
        p = k;
      }
      piv[k] = p;
      if (p != k) {
        const int tmp = perm[k];
        perm[k] = perm[p];
        perm[p] = tmp;
      }
      s_piv = p;
    }
    __syncthreads();
    const int p = s_piv;
    if (p != k) {
      double* rk = A + 2 * (size_t)k * lda;
      double* rp = A + 2 * (size_t)p * lda;
      for (int col = t; col < n; col += PANEL_THREADS) {
        const double ar = rk[2 * col], ai = rk[2 * col + 1];
        rk[2 * col] = rp[2 * col];
        rk[2 * col + 1] = rp[2 * col + 1];
        rp[2 * col] = ar;
        rp[2 * col + 1] = ai;
      }
    }
    __syncthreads();
    const int width = nb - c;  // pivot row entries k .. k0+nb-1
    if (t < width) {
      const double* v = A + 2 * ((size_t)k * lda + k + t);
      s_prow[2 * t] = v[0];
      s_prow[2 * t + 1] = v[1];
    }
    __syncthreads();
    const zc d = zmk(s_prow[0], s_prow[1]);
    const bool scale = (d.re != 0.0) || (d.im != 0.0);
    // scale column k and apply the rank-1 update to the rest of the panel;
    // each wave takes whole rows, lane q <-> panel column k + q.
    best = -1.0;
    bidx = INT_MAX;
    const zc u = (lane < width) ? zmk(s_prow[2 * lane], s_prow[2 * lane + 1]) : zmk(0.0, 0.0);
    for (int i = k + 1 + wave; i < n; i += nwaves) {
      double* row = A + 2 * ((size_t)i * lda + k);
      double vr = 0.0, vi = 0.0;
      if (lane < width) {
        vr = row[2 * lane];
        vi = row[2 * lane + 1];
      }
      zc l = zmk(__shfl(vr, 0, 64), __shfl(vi, 0, 64));
      if (scale) l = zdiv(l, d);
      if (lane == 0) {
        row[0] = l.re;
        row[1] = l.im;
      } else if (lane < width) {
        const zc prod = zmul(l, u);
        vr = vr - prod.re;
        vi = vi - prod.im;
        row[2 * lane] = vr;
        row[2 * lane + 1] = vi;
        if (lane == 1) {
          const double h = hypot(vr, vi);
          if (h > best) {
            best = h;
            bidx = i;
          }
        }
      }
    }
    // candidates for the next column live in lane 1 of every wave
    best = __shfl(best, 1, 64);
    bidx = __shfl(bidx, 1, 64);
    __syncthreads();
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
        const int i = row0 + wr * 32 + a * 16 + lk * 4 + v;
        const int j = col0 + wc * 32 + b * 16 + li;
        if (i < M && j < N) {
          double* c = A + 2 * ((size_t)(m0 + i) * lda + n0 + j);
          c[0] -= accr[a][b][v];
          c[1] -= acci[a][b][v];
        }
      }
}

}  // namespace

int launch_lu_fp64(double* A, long long lda, int n, int* piv, int* info, hipStream_t stream) {
  // perm lives right after piv in the caller's int buffer (2*n ints).
  int* perm = piv + n;
  hipLaunchKernelGGL(init_perm_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, perm, n);
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int nb = (n - k0 < NB) ? (n - k0) : NB;
    hipLaunchKernelGGL(lu_panel_kernel, dim3(1), dim3(PANEL_THREADS), 0, stream, A, lda, n, k0, nb, piv, perm, info);
    const int rest = n - k0 - nb;
    if (rest <= 0) break;
    hipLaunchKernelGGL(lu_trsm_kernel, dim3((rest + 63) / 64), dim3(256), 0, stream, A, lda, n, k0, nb);
    hipLaunchKernelGGL(zgemm_update_kernel, dim3((rest + 63) / 64, (rest + 63) / 64), dim3(256), 0, stream, A, lda,
                       k0 + nb, k0 + nb, rest, rest, k0, nb);
  }
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace necgpu
