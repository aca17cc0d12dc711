// solve.hip -- forward / backward substitution with the LU factors.
//
// Replaces matrix_algebra.cpp solve (:271-307): x <- P b, L y = x (unit
// lower), U x = y.  The factors are FP64 (the FP64 path) or FP32 (the
// mixed-precision path); the right-hand side and all arithmetic are FP64.
//
// Once per factorisation, the 64 x 64 diagonal blocks of L and U are
// inverted (one wave per block, all blocks at once).  Each triangular sweep
// is then one launch of one workgroup per 64-row block: block b subtracts
// the products with every finished block as soon as that block's flag is
// raised (flags carry an epoch, so no reset is needed between calls; one
// poll of 64 consecutive flags takes every block finished so far), then
// applies its diagonal block's inverse (a 64 x 64 matrix-vector product
// instead of 64 dependent steps) and raises its flag.  The block that
// finishes last before b (the neighbour) and the inverse are staged in LDS
// while b waits, so that b's critical path is two small products.
// Workgroups are dispatched in block order, so every block a workgroup
// waits on is already resident or done.
//
// Multiplying by a computed inverse leaves a residual of about
// u cond(T) |T| |x| (u = 2^-53, T the diagonal block), where substitution
// leaves u |T| |x|.  The blocks of a well-posed structure have cond(T) of a
// few hundred, but a nearly singular structure (wires thicker than their
// segments are long) has blocks with cond(T) up to 1e17, whose inverse
// would swamp the solution.  So the inversion also estimates
// ||T||_inf ||T^-1||_inf, and a block above kInvCondMax is solved by
// substitution (64 dependent steps in one wave) instead.
#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

constexpr int SB = 64;
constexpr double kInvCondMax = 1.0e3;

// x from lane l of the wave (l uniform)
__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double ld(const double* p) { return *p; }
__device__ __forceinline__ double ld(const float* p) { return static_cast<double>(*p); }

// Inverse of diagonal block b of L (blockIdx.y == 0, unit lower) or U
// (blockIdx.y == 1, upper with its diagonal), padded to 64 x 64 with the
// identity.  Lane j computes column j by substitution on e_j.  subst[y nblk
// + b] = 1 when the block's condition estimate (infinity norms, |re| + |im|
// per entry) exceeds kInvCondMax or is not finite: that block is then
// solved by substitution.
template <typename TF>
__global__ void __launch_bounds__(64) diag_inverse_kernel(const TF* __restrict__ LU, long long lda, int n,
                                                          double* __restrict__ inv, int* __restrict__ subst) {
  __shared__ double sT[SB * SB * 2];  // the block, [r][q]
  __shared__ double sX[SB * SB * 2];  // its inverse, [r][j]
  const int b = blockIdx.x;
  const bool upper = blockIdx.y == 1;
  const int nblk = gridDim.x;
  const int r0 = b * SB;
  const int m = min(SB, n - r0);
  const int j = threadIdx.x;
  for (int e = j; e < SB * SB; e += 64) {
    const int r = e / SB, q = e % SB;
    double vr = 0.0, vi = 0.0;
    if (r < m && q < m) {
      const TF* v = LU + 2 * ((size_t)(r0 + r) * lda + r0 + q);
      vr = ld(v);
      vi = ld(v + 1);
    }
    sT[2 * e] = vr;
    sT[2 * e + 1] = vi;
  }
  __syncthreads();
  if (!upper) {
    for (int r = 0; r < SB; ++r) {
      double sr = (r == j) ? 1.0 : 0.0, si = 0.0;
      if (r < m) {
        for (int q = j; q < r; ++q) {
          const double lr = sT[2 * (r * SB + q)], li = sT[2 * (r * SB + q) + 1];
          const double xr = sX[2 * (q * SB + j)], xi = sX[2 * (q * SB + j) + 1];
          sr -= lr * xr - li * xi;
          si -= lr * xi + li * xr;
        }
      }
      sX[2 * (r * SB + j)] = sr;
      sX[2 * (r * SB + j) + 1] = si;
    }
  } else {
    for (int r = SB - 1; r >= 0; --r) {
      double sr = (r == j) ? 1.0 : 0.0, si = 0.0;
      if (r < m && r <= j) {
        for (int q = r + 1; q <= j && q < m; ++q) {
          const double ur = sT[2 * (r * SB + q)], ui = sT[2 * (r * SB + q) + 1];
          const double xr = sX[2 * (q * SB + j)], xi = sX[2 * (q * SB + j) + 1];
          sr -= ur * xr - ui * xi;
          si -= ur * xi + ui * xr;
        }
        const zc x = zdiv(zmk(sr, si), zmk(sT[2 * (r * SB + r)], sT[2 * (r * SB + r) + 1]));
        sr = x.re;
        si = x.im;
      } else if (r < m) {
        sr = 0.0;  // the inverse of an upper triangle is upper
      }
      sX[2 * (r * SB + j)] = sr;
      sX[2 * (r * SB + j) + 1] = si;
    }
  }
  double* dst = inv + ((size_t)(upper ? nblk : 0) + b) * SB * SB * 2;
  for (int e = j; e < SB * SB * 2; e += 64) dst[e] = sX[e];
  __syncthreads();
  // row sums of |T| and |T^-1| for row j (the unit diagonal of L counted),
  // then their maxima over the wave
  double nt = 0.0, nx = 0.0;
  for (int q = 0; q < SB; ++q) {
    const bool in_t = upper ? (q >= j) : (q < j);
    if (in_t) nt += fabs(sT[2 * (j * SB + q)]) + fabs(sT[2 * (j * SB + q) + 1]);
    nx += fabs(sX[2 * (j * SB + q)]) + fabs(sX[2 * (j * SB + q) + 1]);
  }
  if (!upper || j >= m) nt += 1.0;  // the unit diagonal of L; the identity padding of U
  // a NaN would be dropped by fmax: count it as infinite
  if (!(nt <= 1.0e308)) nt = __builtin_inf();
  if (!(nx <= 1.0e308)) nx = __builtin_inf();
  for (int off = 32; off >= 1; off >>= 1) {
    nt = fmax(nt, __shfl_xor(nt, off));
    nx = fmax(nx, __shfl_xor(nx, off));
  }
  if (j == 0) subst[(upper ? nblk : 0) + b] = (nt * nx <= kInvCondMax) ? 0 : 1;
}

// x values and flags travel through the device coherence point: written
// with agent-scope atomic stores (write-through), the flag after the data
// stores have completed, and read with agent-scope atomic loads, so no
// cache write-back or invalidate is needed.
__device__ __forceinline__ void xstore(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double xload(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One triangular sweep.  lower: out[r] = y_r with L y = P rhs (rhs read
// through perm).  upper: out[r] = x_r with U x = rhs.  Blocks finish in
// order, so one poll of 64 consecutive flags tells how many more finished
// blocks can be taken at once.
template <typename TF, bool lower>
__global__ void __launch_bounds__(256) trsv_kernel(const TF* __restrict__ LU, long long lda, int n,
                                                   const int* __restrict__ perm, const double* rhs, double* out,
                                                   const double* __restrict__ inv_all,
                                                   const int* __restrict__ subst, unsigned* flags,
                                                   unsigned epoch, int* status) {
  __shared__ double sInv[SB * SB * 2];   // inverse of the diagonal block (or the block itself)
  __shared__ double sNb[SB * SB * 2];    // the neighbour block
  __shared__ double sPart[4][SB * 2];
  __shared__ double sAcc[SB * 2];
  __shared__ int s_ready;
  const int nblk = (n + SB - 1) / SB;
  const int b = lower ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
  const int nb = lower ? b - 1 : b + 1;  // neighbour block (last to finish)
  const bool has_nb = lower ? (b > 0) : (b < nblk - 1);
  const int r0 = b * SB;
  const int rows = min(SB, n - r0);
  const int t = threadIdx.x;
  const int r = t & 63;     // row within the block
  const int part = t >> 6;  // column quarter of this thread
  const bool by_subst = subst[(lower ? 0 : nblk) + b] != 0;
  // stage the inverse (or, to substitute, the block padded with the
  // identity) and the neighbour block while earlier blocks finish
  {
    if (by_subst) {
      for (int e = t; e < SB * SB; e += 256) {
        const int i = e / SB, q = e % SB;
        double vr = (i == q) ? 1.0 : 0.0, vi = 0.0;
        if (i < rows && q < rows) {
          const TF* v = LU + 2 * ((size_t)(r0 + i) * lda + r0 + q);
          vr = ld(v);
          vi = ld(v + 1);
        }
        sInv[2 * e] = vr;
        sInv[2 * e + 1] = vi;
      }
    } else {
      const double* src = inv_all + ((size_t)(lower ? 0 : nblk) + b) * SB * SB * 2;
      for (int e = t; e < SB * SB * 2; e += 256) sInv[e] = src[e];
    }
    if (has_nb) {
      const int c0 = nb * SB;
      const int cols = min(SB, n - c0);
      for (int e = t; e < SB * SB; e += 256) {
        const int i = e / SB, q = e % SB;
        double vr = 0.0, vi = 0.0;
        if (i < rows && q < cols) {
          const TF* v = LU + 2 * ((size_t)(r0 + i) * lda + c0 + q);
          vr = ld(v);
          vi = ld(v + 1);
        }
        sNb[2 * e] = vr;
        sNb[2 * e + 1] = vi;
      }
    }
  }
  double accr = 0.0, acci = 0.0;
  // blocks other than the neighbour, taken in finished batches
  const int others = lower ? max(0, b - 1) : max(0, nblk - 2 - b);
  int done = 0;
  long spins = 0;
  while (done < others) {
    if (t < 64) {
      const int step = done + t;
      bool ok = false;
      if (step < others) {
        const int c = lower ? step : nblk - 1 - step;
        ok = __hip_atomic_load(flags + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
      }
      const unsigned long long m = __ballot(ok);
      // count of consecutive finished blocks from `done`
      const int cnt = (~m == 0ull) ? 64 : __ffsll((long long)~m) - 1;
      if (t == 0) {
        int ready = cnt;
        if (cnt == 0) {
          // bounded wait; a timeout is sticky so every later wait returns at once
          if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++spins > (1L << 22)) {
            __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ready = others - done;  // give up waiting: finish with what is there
          } else {
            __builtin_amdgcn_s_sleep(1);
          }
        }
        s_ready = ready;
      }
    }
    __syncthreads();
    const int cnt = s_ready;
    __syncthreads();
    for (int step = done; step < done + cnt; ++step) {
      const int c = lower ? step : nblk - 1 - step;
      const int c0 = c * SB;
      const int cols = min(SB, n - c0);
      if (r < rows) {
        const TF* lrow = LU + 2 * ((size_t)(r0 + r) * lda + c0);
        for (int j = part; j < cols; j += 4) {
          const double yr = xload(out + 2 * (c0 + j));
          const double yi = xload(out + 2 * (c0 + j) + 1);
          const double lr = ld(lrow + 2 * j), li = ld(lrow + 2 * j + 1);
          accr += lr * yr - li * yi;
          acci += lr * yi + li * yr;
        }
      }
    }
    done += cnt;
  }
  if (has_nb) {
    if (t == 0) {
      long sp = 0;
      while (__hip_atomic_load(flags + nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (++sp > (1L << 22)) {
          __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    const int c0 = nb * SB;
    for (int j = part; j < SB; j += 4) {
      const double yr = (c0 + j < n) ? xload(out + 2 * (c0 + j)) : 0.0;
      const double yi = (c0 + j < n) ? xload(out + 2 * (c0 + j) + 1) : 0.0;
      const double lr = sNb[2 * (r * SB + j)], li = sNb[2 * (r * SB + j) + 1];
      accr += lr * yr - li * yi;
      acci += lr * yi + li * yr;
    }
  }
  sPart[part][2 * r] = accr;
  sPart[part][2 * r + 1] = acci;
  __syncthreads();
  if (t < SB) {
    double vr = 0.0, vi = 0.0;
    if (t < rows) {
      const int src = lower ? perm[r0 + t] : r0 + t;
      vr = rhs[2 * src];
      vi = rhs[2 * src + 1];
    }
    const double sr = ((sPart[0][2 * t] + sPart[1][2 * t]) + sPart[2][2 * t]) + sPart[3][2 * t];
    const double si = ((sPart[0][2 * t + 1] + sPart[1][2 * t + 1]) + sPart[2][2 * t + 1]) + sPart[3][2 * t + 1];
    sAcc[2 * t] = (t < rows) ? vr - sr : 0.0;
    sAcc[2 * t + 1] = (t < rows) ? vi - si : 0.0;
  }
  __syncthreads();
  if (!by_subst) {
    // x_b = inverse * acc
    double xr = 0.0, xi = 0.0;
    for (int j = part; j < SB; j += 4) {
      const double ir = sInv[2 * (r * SB + j)], ii = sInv[2 * (r * SB + j) + 1];
      const double ar = sAcc[2 * j], ai = sAcc[2 * j + 1];
      xr += ir * ar - ii * ai;
      xi += ir * ai + ii * ar;
    }
    sPart[part][2 * r] = xr;
    sPart[part][2 * r + 1] = xi;
    __syncthreads();
    if (t < rows) {
      xstore(out + 2 * (r0 + t), ((sPart[0][2 * t] + sPart[1][2 * t]) + sPart[2][2 * t]) + sPart[3][2 * t]);
      xstore(out + 2 * (r0 + t) + 1,
             ((sPart[0][2 * t + 1] + sPart[1][2 * t + 1]) + sPart[2][2 * t + 1]) + sPart[3][2 * t + 1]);
    }
  } else if (t < SB) {
    // substitution in the first wave, lane t holding row t: step q takes
    // x_q from lane q and subtracts its column from the rows still open
    double ar = sAcc[2 * t], ai = sAcc[2 * t + 1];
    if (lower) {
      for (int q = 0; q < SB - 1; ++q) {
        const double yr = readlane_d(ar, q), yi = readlane_d(ai, q);
        if (t > q) {
          const double lr = sInv[2 * (t * SB + q)], li = sInv[2 * (t * SB + q) + 1];
          ar -= lr * yr - li * yi;
          ai -= lr * yi + li * yr;
        }
      }
    } else {
      for (int q = SB - 1; q >= 0; --q) {
        if (t == q) {
          const zc x = zdiv(zmk(ar, ai), zmk(sInv[2 * (q * SB + q)], sInv[2 * (q * SB + q) + 1]));
          ar = x.re;
          ai = x.im;
        }
        const double xr = readlane_d(ar, q), xi = readlane_d(ai, q);
        if (t < q) {
          const double ur = sInv[2 * (t * SB + q)], ui = sInv[2 * (t * SB + q) + 1];
          ar -= ur * xr - ui * xi;
          ai -= ur * xi + ui * xr;
        }
      }
    }
    if (t < rows) {
      xstore(out + 2 * (r0 + t), ar);
      xstore(out + 2 * (r0 + t) + 1, ai);
    }
  }
  // the data stores have completed before the flag is raised
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (t == 0) __hip_atomic_store(flags + b, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Z^T lam = g with the FP64 factors (P Z = L U, perm = piv + n):
// U^T w = g (forward, U^T lower), L^T v = w (backward, unit), lam[perm[i]]
// = v[i].  One workgroup; each step reads one contiguous row of the factors.
// The adjoint's transposed solve for small structures.
__global__ void __launch_bounds__(1024) lu_solve_transposed_kernel(const double* __restrict__ LU, long long lda,
                                                                    int n, const int* __restrict__ perm,
                                                                    const double* __restrict__ g,
                                                                    double* __restrict__ lam, double* acc) {
  const int t = threadIdx.x;
  for (int i = t; i < n; i += blockDim.x) {
    acc[2 * i] = g[2 * i];
    acc[2 * i + 1] = g[2 * i + 1];
  }
  __syncthreads();
  for (int q = 0; q < n; ++q) {
    if (t == 0) {
      const double* u = LU + 2 * ((size_t)q * lda + q);
      const zc w = zdiv(zmk(acc[2 * q], acc[2 * q + 1]), zmk(u[0], u[1]));
      acc[2 * q] = w.re;
      acc[2 * q + 1] = w.im;
    }
    __syncthreads();
    const double wr = acc[2 * q], wi = acc[2 * q + 1];
    const double* row = LU + 2 * (size_t)q * lda;
    for (int i = q + 1 + t; i < n; i += blockDim.x) {
      const double ur = row[2 * i], ui = row[2 * i + 1];
      acc[2 * i] -= ur * wr - ui * wi;
      acc[2 * i + 1] -= ur * wi + ui * wr;
    }
    __syncthreads();
  }
  for (int q = n - 1; q >= 0; --q) {
    const double vr = acc[2 * q], vi = acc[2 * q + 1];
    const double* row = LU + 2 * (size_t)q * lda;
    for (int i = t; i < q; i += blockDim.x) {
      const double lr = row[2 * i], li = row[2 * i + 1];
      acc[2 * i] -= lr * vr - li * vi;
      acc[2 * i + 1] -= lr * vi + li * vr;
    }
    __syncthreads();
  }
  for (int i = t; i < n; i += blockDim.x) {
    lam[2 * perm[i]] = acc[2 * i];
    lam[2 * perm[i] + 1] = acc[2 * i + 1];
  }
}

__global__ void reset_status_kernel(int* status) {
  if (threadIdx.x == 0) *status = 0;
}

struct SolveWork {
  unsigned *flags_f, *flags_b;
  int* status;
  double* tmp;
  int* subst;
  double* inv;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

SolveWork carve_solve(void* work, int n) {
  const int nblk = (n + SB - 1) / SB;
  char* p = static_cast<char*>(work);
  SolveWork w;
  w.flags_f = reinterpret_cast<unsigned*>(p);
  w.flags_b = w.flags_f + nblk;
  p += align256((size_t)2 * nblk * sizeof(unsigned));
  w.status = reinterpret_cast<int*>(p);
  p += 256;
  w.tmp = reinterpret_cast<double*>(p);
  p += align256((size_t)n * 16);
  w.subst = reinterpret_cast<int*>(p);
  p += align256((size_t)2 * nblk * sizeof(int));
  w.inv = reinterpret_cast<double*>(p);
  return w;
}

template <typename TF>
int prepare(const TF* LU, long long lda, int n, void* work, hipStream_t stream) {
  if (n <= 0) return 0;
  const int nblk = (n + SB - 1) / SB;
  const SolveWork w = carve_solve(work, n);
  launch_k(K_DIAG_INVERSE, diag_inverse_kernel<TF>, dim3(nblk, 2), dim3(64), 0u, stream, LU, lda, n, w.inv,
           w.subst);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

template <typename TF>
int apply(const TF* LU, long long lda, int n, const int* piv, double* x, void* work, unsigned epoch,
          hipStream_t stream) {
  if (n <= 0) return 0;
  const int nblk = (n + SB - 1) / SB;
  const SolveWork w = carve_solve(work, n);
  const int* perm = piv + n;
  launch_k(K_SOLVE_RESET, reset_status_kernel, dim3(1), dim3(64), 0u, stream, w.status);
  launch_k(K_TRSV_LOWER, trsv_kernel<TF, true>, dim3(nblk), dim3(256), 0u, stream, LU, lda, n, perm,
           (const double*)x, w.tmp, (const double*)w.inv, (const int*)w.subst, w.flags_f, epoch, w.status);
  launch_k(K_TRSV_UPPER, trsv_kernel<TF, false>, dim3(nblk), dim3(256), 0u, stream, LU, lda, n, perm,
           (const double*)w.tmp, x, (const double*)w.inv, (const int*)w.subst, w.flags_b, epoch, w.status);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace

size_t solve_workspace_bytes(int n) {
  const int nblk = (n + SB - 1) / SB;
  return align256((size_t)2 * nblk * sizeof(unsigned)) + 256 + align256((size_t)n * 16) +
         align256((size_t)2 * nblk * sizeof(int)) + (size_t)2 * nblk * SB * SB * 16;
}

int* solve_status(void* work, int n) { return carve_solve(work, n).status; }

int launch_lu_solve_transposed(const double* LU, long long lda, int n, const int* piv, const double* g, double* lam,
                               double* acc, hipStream_t stream) {
  if (n <= 0) return 0;
  launch_k(K_ADJOINT_SOLVE, lu_solve_transposed_kernel, dim3(1), dim3(1024), 0u, stream, LU, lda, n, piv + n, g, lam,
           acc);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_lu_solve_prepare(const double* LU, long long lda, int n, void* work, hipStream_t stream) {
  return prepare<double>(LU, lda, n, work, stream);
}
int launch_lu_solve_prepare_f32(const float* LU, long long lda, int n, void* work, hipStream_t stream) {
  return prepare<float>(LU, lda, n, work, stream);
}
int launch_lu_solve_apply(const double* LU, long long lda, int n, const int* piv, double* x, void* work,
                          unsigned epoch, hipStream_t stream) {
  return apply<double>(LU, lda, n, piv, x, work, epoch, stream);
}
int launch_lu_solve_apply_f32(const float* LU, long long lda, int n, const int* piv, double* x, void* work,
                              unsigned epoch, hipStream_t stream) {
  return apply<float>(LU, lda, n, piv, x, work, epoch, stream);
}

int launch_lu_solve_fp64(const double* LU, long long lda, int n, const int* piv, double* x, void* work,
                         unsigned epoch, hipStream_t stream) {
  if (prepare<double>(LU, lda, n, work, stream)) return 1;
  return apply<double>(LU, lda, n, piv, x, work, epoch, stream);
}

}  // namespace necgpu
