// solve.hip -- forward / backward substitution with the LU factors (FP64).
//
// Replaces matrix_algebra.cpp solve (:271-307): x <- P b, L y = x (unit
// lower), U x = y.  Each triangular sweep is one launch of one workgroup per
// 64-row block.  Block b accumulates the products with every finished block
// as soon as that block's flag is raised (flags carry an epoch, so no reset
// is needed between calls), then solves its own 64 x 64 diagonal block in LDS
// and raises its flag.  Workgroups are dispatched in block order, so every
// block a workgroup waits on is already resident or done.
#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

constexpr int SB = 64;

__device__ __forceinline__ unsigned load_flag(const unsigned* f) {
  return __hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void permute_kernel(const double* __restrict__ b, const int* __restrict__ perm, double* __restrict__ x,
                               int n, int* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *status = 0;
  if (i < n) {
    const int p = perm[i];
    x[2 * i] = b[2 * p];
    x[2 * i + 1] = b[2 * p + 1];
  }
}

// lower = true: L y = x, unit diagonal, blocks in increasing order.
// lower = false: U x = y, blocks in decreasing order.
template <bool lower>
__global__ void __launch_bounds__(256) trsv_blocked_kernel(const double* __restrict__ LU, long long lda, int n,
                                                           double* x, unsigned* flags, unsigned epoch,
                                                           int* status) {
  __shared__ double sD[SB * SB * 2];
  __shared__ double sAcc[SB * 2];
  __shared__ double sPart[4][SB * 2];
  const int nblk = (n + SB - 1) / SB;
  const int b = lower ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
  const int r0 = b * SB;
  const int rows = (n - r0 < SB) ? (n - r0) : SB;
  const int t = threadIdx.x;
  const int r = t & 63;    // row within the block
  const int part = t >> 6; // column quarter handled by this thread
  // diagonal block into LDS
  for (int e = t; e < rows * rows; e += 256) {
    const int i = e / rows, j = e % rows;
    const double* v = LU + 2 * ((size_t)(r0 + i) * lda + r0 + j);
    sD[2 * (i * SB + j)] = v[0];
    sD[2 * (i * SB + j) + 1] = v[1];
  }
  double accr = 0.0, acci = 0.0;
  // off-diagonal blocks in the order they become available
  for (int step = 0; step < (lower ? b : nblk - 1 - b); ++step) {
    const int c = lower ? step : nblk - 1 - step;
    if (t == 0) {
      // bounded wait; a timeout is sticky so every later wait returns at once
      long spins = 0;
      while (load_flag(flags + c) != epoch) {
        if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (++spins > (1L << 22)) {
          __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    const int c0 = c * SB;
    const int cols = (n - c0 < SB) ? (n - c0) : SB;
    if (r < rows) {
      const double* lrow = LU + 2 * ((size_t)(r0 + r) * lda + c0);
      for (int j = part; j < cols; j += 4) {
        const double yr = __builtin_nontemporal_load(x + 2 * (c0 + j));
        const double yi = __builtin_nontemporal_load(x + 2 * (c0 + j) + 1);
        const double lr = lrow[2 * j], li = lrow[2 * j + 1];
        accr += lr * yr - li * yi;
        acci += lr * yi + li * yr;
      }
    }
  }
  sPart[part][2 * r] = accr;
  sPart[part][2 * r + 1] = acci;
  __syncthreads();
  if (t < rows) {
    const double sr = ((sPart[0][2 * t] + sPart[1][2 * t]) + sPart[2][2 * t]) + sPart[3][2 * t];
    const double si = ((sPart[0][2 * t + 1] + sPart[1][2 * t + 1]) + sPart[2][2 * t + 1]) + sPart[3][2 * t + 1];
    sAcc[2 * t] = x[2 * (r0 + t)] - sr;
    sAcc[2 * t + 1] = x[2 * (r0 + t) + 1] - si;
  }
  __syncthreads();
  // diagonal block: one wave, sequential over rows, lanes over rows
  if (t < 64) {
    if (lower) {
      for (int q = 0; q < rows; ++q) {
        const zc yq = zmk(sAcc[2 * q], sAcc[2 * q + 1]);
        if (t > q && t < rows) {
          const zc l = zmk(sD[2 * (t * SB + q)], sD[2 * (t * SB + q) + 1]);
          const zc p = zmul(l, yq);
          sAcc[2 * t] -= p.re;
          sAcc[2 * t + 1] -= p.im;
        }
        __builtin_amdgcn_wave_barrier();
      }
    } else {
      for (int q = rows - 1; q >= 0; --q) {
        if (t == q) {
          const zc v = zdiv(zmk(sAcc[2 * q], sAcc[2 * q + 1]), zmk(sD[2 * (q * SB + q)], sD[2 * (q * SB + q) + 1]));
          sAcc[2 * q] = v.re;
          sAcc[2 * q + 1] = v.im;
        }
        __builtin_amdgcn_wave_barrier();
        const zc xq = zmk(sAcc[2 * q], sAcc[2 * q + 1]);
        if (t < q) {
          const zc u = zmk(sD[2 * (t * SB + q)], sD[2 * (t * SB + q) + 1]);
          const zc p = zmul(u, xq);
          sAcc[2 * t] -= p.re;
          sAcc[2 * t + 1] -= p.im;
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
    if (t < rows) {
      x[2 * (r0 + t)] = sAcc[2 * t];
      x[2 * (r0 + t) + 1] = sAcc[2 * t + 1];
    }
  }
  __syncthreads();
  if (t == 0) {
    __threadfence();
    __hip_atomic_store(flags + b, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace

size_t solve_workspace_bytes(int n) {
  const int nblk = (n + SB - 1) / SB;
  return (size_t)2 * nblk * sizeof(unsigned) + 64 + (size_t)n * 16 + 256;
}

int launch_lu_solve_fp64(const double* LU, long long lda, int n, const int* piv, double* x, void* work,
                         unsigned epoch, hipStream_t stream) {
  const int nblk = (n + SB - 1) / SB;
  unsigned* flags_f = static_cast<unsigned*>(work);
  unsigned* flags_b = flags_f + nblk;
  int* status = reinterpret_cast<int*>(flags_b + nblk);
  double* tmp = reinterpret_cast<double*>(static_cast<char*>(work) + (((size_t)2 * nblk * sizeof(unsigned) + 64 + 255) & ~(size_t)255));
  const int* perm = piv + n;
  launch_k(K_PERMUTE, permute_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, x, perm, tmp, n, status);
  launch_k(K_TRSV_LOWER, trsv_blocked_kernel<true>, dim3(nblk), dim3(256), 0, stream, LU, lda, n, tmp, flags_f, epoch,
                     status);
  launch_k(K_TRSV_UPPER, trsv_blocked_kernel<false>, dim3(nblk), dim3(256), 0, stream, LU, lda, n, tmp, flags_b, epoch,
                     status);
  if (hipMemcpyAsync(x, tmp, (size_t)n * 16, hipMemcpyDeviceToDevice, stream) != hipSuccess) return 1;
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace necgpu
