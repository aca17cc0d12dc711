// refine.hip -- the mixed-precision path's FP64 side.
//
// The FP64 impedance matrix stays in HBM; an FP32 copy is factored
// (launch_lu_fp32) and iterative refinement wins the accuracy back:
//   x = LU32 \ b;  repeat  r = b - Z x (FP64),  x += LU32 \ r
// until ||r|| <= tol ||b||.  This file holds the copy, the FP64 residual
// (one wave per matrix row, reading the row once) and the update.  There is
// no counterpart in nec2++, which factors in FP64 only (matrix_algebra.cpp
// lu_decompose, :217-266); the solve it refines is matrix_algebra.cpp
// solve (:271-307).
#include "kernels.h"

namespace necgpu {

namespace {

__global__ void __launch_bounds__(256) to_f32_kernel(const double* __restrict__ Z, long long ldz,
                                                     float* __restrict__ Z32, long long ld32, int n) {
  const size_t total = (size_t)n * n;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t i = e / n, k = e % n;
    const double* s = Z + 2 * (i * ldz + k);
    float* d = Z32 + 2 * (i * ld32 + k);
    d[0] = static_cast<float>(s[0]);
    d[1] = static_cast<float>(s[1]);
  }
}

// r = b - Z x, one wave per row; acc[0] += |r|^2, and acc[1] += |b|^2 when
// with_b (the first residual of a solve).
__global__ void __launch_bounds__(256) residual_kernel(const double* __restrict__ Z, long long ldz, int n,
                                                       const double* __restrict__ x, const double* __restrict__ b,
                                                       double* __restrict__ r, double* acc, int with_b) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  const double* zr = Z + 2 * (size_t)row * ldz;
  double sr = 0.0, si = 0.0;
  for (int k = lane; k < n; k += 64) {
    const double ar = zr[2 * k], ai = zr[2 * k + 1];
    const double xr = x[2 * k], xi = x[2 * k + 1];
    sr += ar * xr - ai * xi;
    si += ar * xi + ai * xr;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sr += __shfl_xor(sr, off, 64);
    si += __shfl_xor(si, off, 64);
  }
  if (lane == 0) {
    const double br = b[2 * row], bi = b[2 * row + 1];
    const double rr = br - sr, ri = bi - si;
    r[2 * row] = rr;
    r[2 * row + 1] = ri;
    atomicAdd(&acc[0], rr * rr + ri * ri);
    if (with_b) atomicAdd(&acc[1], br * br + bi * bi);
  }
}

__global__ void __launch_bounds__(256) update_kernel(double* __restrict__ x, const double* __restrict__ d, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    x[2 * i] += d[2 * i];
    x[2 * i + 1] += d[2 * i + 1];
  }
}

__global__ void zero_kernel(double* acc, int m) {
  if ((int)threadIdx.x < m) acc[threadIdx.x] = 0.0;
}

}  // namespace

int launch_to_f32(const double* Z, long long ldz, float* Z32, long long ld32, int n, hipStream_t stream) {
  if (n <= 0) return 0;
  launch_k(K_REFINE, to_f32_kernel, dim3(4096), dim3(256), 0u, stream, Z, ldz, Z32, ld32, n);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_residual(const double* Z, long long ldz, int n, const double* x, const double* b, double* r, double* acc,
                    bool with_b, hipStream_t stream) {
  if (n <= 0) return 0;
  launch_k(K_REFINE, zero_kernel, dim3(1), dim3(64), 0u, stream, acc, with_b ? 2 : 1);
  launch_k(K_RESIDUAL, residual_kernel, dim3((n + 3) / 4), dim3(256), 0u, stream, Z, ldz, n, x, b, r, acc,
           with_b ? 1 : 0);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_update(double* x, const double* d, int n, hipStream_t stream) {
  if (n <= 0) return 0;
  launch_k(K_REFINE, update_kernel, dim3((n + 255) / 256), dim3(256), 0u, stream, x, d, n);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace necgpu
