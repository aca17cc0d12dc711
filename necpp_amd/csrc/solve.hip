// solve.hip -- forward / backward substitution with the LU factors.
//
// Replaces matrix_algebra.cpp solve (:271-307): x <- P b, L y = x (unit
// lower), U x = y.  The factors are FP64 (the FP64 path) or FP32 (the
// mixed-precision path); the right-hand side and all arithmetic are FP64.
//
// Once per factorisation, the 64 x 64 diagonal blocks of L and U are
// inverted (one wave per block, all blocks at once).  Each triangular sweep
// is then one launch of one workgroup per 64-row block: block b subtracts
// the products with every earlier block as soon as that block's values
// appear (the outputs are cleared to a sentinel before the sweeps, so the
// values themselves say they are final), then applies its diagonal block's
// inverse (a 64 x 64 matrix-vector product instead of 64 dependent steps).
// The factor rows of the next block to come are loaded while b waits, so
// that b's critical path after its neighbour is a product with registers,
// one reduction and the inverse.  Workgroups are dispatched in block order,
// so every block a workgroup waits on is already resident or done.
//
// Multiplying by a computed inverse leaves a residual of about
// u cond(T) |T| |x| (u = 2^-53, T the diagonal block), where substitution
// leaves u |T| |x|.  The blocks of a well-posed structure have cond(T) of a
// few hundred, but a nearly singular structure (wires thicker than their
// segments are long) has blocks with cond(T) up to 1e17, whose inverse
// would swamp the solution.  So the inversion also estimates
// ||T||_inf ||T^-1||_inf, and a block above inv_cond_max is solved by
// substitution (64 dependent steps in one wave) instead.  FP32 factors (the
// mixed path) take a far higher limit: their own rounding, about 6e-8
// relative, is what the FP64 refinement corrects, and an inverse formed in
// FP64 from the FP32 block adds only about 1e-16 cond(T) on top of it.
#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

constexpr int SB = 64;
template <typename TF>
constexpr double inv_cond_max() {
  return sizeof(TF) == 4 ? 1.0e8 : 1.0e3;
}

// x from lane l of the wave (l uniform)
__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double ld(const double* p) { return *p; }
__device__ __forceinline__ double ld(const float* p) { return static_cast<double>(*p); }
// one complex entry (re, im adjacent) in one vector load, kept in the
// factors' own type: a conversion right after the load would make the
// load's wait immediate and spoil the prefetch
typedef double sv_dbl2 __attribute__((ext_vector_type(2)));
typedef float sv_flt2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ldc(const double* p, double& re, double& im) {
  const sv_dbl2 v = *reinterpret_cast<const sv_dbl2*>(p);
  re = v.x;
  im = v.y;
}
__device__ __forceinline__ void ldc(const float* p, float& re, float& im) {
  const sv_flt2 v = *reinterpret_cast<const sv_flt2*>(p);
  re = v.x;
  im = v.y;
}

// Inverse of diagonal block b of L (blockIdx.y == 0, unit lower) or U
// (blockIdx.y == 1, upper with its diagonal), padded to 64 x 64 with the
// identity.  Lane j computes column j by substitution on e_j.  subst[y nblk
// + b] = 1 when the block's condition estimate (infinity norms, |re| + |im|
// per entry) exceeds inv_cond_max<TF>() or is not finite: that block is then
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
        // four terms' LDS reads go out together, then the same sums in the
        // same order (one wait per four terms instead of one per term:
        // 200 -> 130 us for the 52 x 2 blocks at N = 3,280; eight measured
        // no better)
        int q = j;
        for (; q + 4 <= r; q += 4) {
          double lr[4], li[4], xr[4], xi[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            lr[u] = sT[2 * (r * SB + q + u)];
            li[u] = sT[2 * (r * SB + q + u) + 1];
            xr[u] = sX[2 * ((q + u) * SB + j)];
            xi[u] = sX[2 * ((q + u) * SB + j) + 1];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            sr -= lr[u] * xr[u] - li[u] * xi[u];
            si -= lr[u] * xi[u] + li[u] * xr[u];
          }
        }
        for (; q < r; ++q) {
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
        const int qe = min(j, m - 1);  // last q of the sum
        int q = r + 1;
        for (; q + 3 <= qe; q += 4) {
          double ur[4], ui[4], xr[4], xi[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            ur[u] = sT[2 * (r * SB + q + u)];
            ui[u] = sT[2 * (r * SB + q + u) + 1];
            xr[u] = sX[2 * ((q + u) * SB + j)];
            xi[u] = sX[2 * ((q + u) * SB + j) + 1];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            sr -= ur[u] * xr[u] - ui[u] * xi[u];
            si -= ur[u] * xi[u] + ui[u] * xr[u];
          }
        }
        for (; q <= qe; ++q) {
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
  if (j == 0) subst[(upper ? nblk : 0) + b] = (nt * nx <= inv_cond_max<TF>()) ? 0 : 1;
}

// x values travel through the device coherence point: written with
// agent-scope atomic stores (write-through) and read with agent-scope
// atomic loads, so no cache write-back or invalidate is needed.
__device__ __forceinline__ void xstore(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double xload(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int xload_i(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A value the sweep has not written yet: all-ones bits, a NaN payload that
// no arithmetic produces (a computed NaN is stored as the canonical one).
constexpr long long kUnwritten = -1LL;
__device__ __forceinline__ bool unwritten(double v) { return __double_as_longlong(v) == kUnwritten; }
__device__ __forceinline__ double canonical_x(double v) {
  return (v != v) ? __longlong_as_double(0x7FF8000000000000LL) : v;
}

// Bounded wait for entry i of a sweep's output (a timeout is sticky).
__device__ __forceinline__ double wait_x(const double* p, int* status) {
  double v = xload(p);
  long spins = 0;
  while (unwritten(v)) {
    if (++spins > (1L << 22) || (spins % 64 == 0 && xload_i(status) != 0)) {
      __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return 0.0;
    }
    __builtin_amdgcn_s_sleep(1);
    v = xload(p);
  }
  return v;
}

// One triangular sweep.  lower: out[r] = y_r with L y = P rhs (rhs read
// through perm).  upper: out[r] = x_r with U x = rhs.  out starts as
// kUnwritten everywhere (solve_reset_kernel), so a reader knows an entry is
// final when it differs from kUnwritten, and no flag is needed; xcopy, if
// given, receives a plain copy of out.  Workgroup b (64 rows) owns rows
// r0 .. r0+63; wave w holds rows r0 + 16 w .. r0 + 16 w + 15 and lane j
// column j of every 64-column block, so every load of the factors is one
// contiguous row piece per wave.
//  - The blocks before b, in the order they finish: each wave, on its own,
//    adds L[r][c0+j] y[c0+j] into its per-lane partial sums (16 rows per
//    lane).  The two blocks just before b are awaited by polling their
//    values, with their factor rows in registers from the start, so the
//    wait for the neighbour's values is the one memory round trip on the
//    chain from block to block; earlier blocks go by their flags.
//  - The partial sums are reduced over the lanes through LDS (row order of
//    the lanes fixed, so the result is deterministic), subtracted from the
//    right-hand side, and multiplied by the diagonal block's inverse (or
//    solved by substitution when the block is ill-conditioned).
//  - x_b goes to out with agent-scope stores (write-through); once they
//    have completed, the block's flag is raised for the batched readers.
template <typename TF, bool lower>
__global__ void __launch_bounds__(256) trsv_kernel(const TF* __restrict__ LU, long long lda, int n,
                                                   const int* __restrict__ perm, const double* rhs, double* out,
                                                   double* xcopy, const double* __restrict__ inv_all,
                                                   const int* __restrict__ subst, int* flags, int* status) {
  constexpr int RS = SB + 1;             // LDS row stride of the reduction (complex)
  constexpr int IS = SB + 1;             // LDS column stride of sInv (complex)
  // inverse of the diagonal block (or the block itself), by columns: entry
  // (i, q) at 2 (q IS + i), so lanes over rows read consecutive words
  __shared__ double sInv[SB * IS * 2];
  __shared__ double sRed[SB * RS * 2];   // per-lane partial sums, [row][lane]
  __shared__ double sPart[4][SB * 2];
  __shared__ double sAcc[SB * 2];
  const int nblk = (n + SB - 1) / SB;
  const int b = lower ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
  const int r0 = b * SB;
  const int rows = min(SB, n - r0);
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const bool by_subst = subst[(lower ? 0 : nblk) + b] != 0;
  // this block's right-hand side, loaded now rather than on the chain
  double rhs_r = 0.0, rhs_i = 0.0;
  if (t < rows) {
    const int src = lower ? perm[r0 + t] : r0 + t;
    rhs_r = rhs[2 * src];
    rhs_i = rhs[2 * src + 1];
  }
  // stage the inverse (or, to substitute, the block padded with the
  // identity) while earlier blocks finish
  if (by_subst) {
    for (int e = t; e < SB * SB; e += 256) {
      const int i = e / SB, q = e % SB;
      double vr = (i == q) ? 1.0 : 0.0, vi = 0.0;
      if (i < rows && q < rows) {
        const TF* v = LU + 2 * ((size_t)(r0 + i) * lda + r0 + q);
        vr = ld(v);
        vi = ld(v + 1);
      }
      sInv[2 * (q * IS + i)] = vr;
      sInv[2 * (q * IS + i) + 1] = vi;
    }
  } else {
    // every load of a thread in flight at once, then the LDS stores
    const sv_dbl2* src = reinterpret_cast<const sv_dbl2*>(inv_all + ((size_t)(lower ? 0 : nblk) + b) * SB * SB * 2);
    sv_dbl2 v[SB * SB / 256];
#pragma unroll
    for (int u = 0; u < SB * SB / 256; ++u) v[u] = src[t + 256 * u];
#pragma unroll
    for (int u = 0; u < SB * SB / 256; ++u) {
      const int e = t + 256 * u, i = e / SB, q = e % SB;
      sInv[2 * (q * IS + i)] = v[u].x;
      sInv[2 * (q * IS + i) + 1] = v[u].y;
    }
  }
  double ar[16], ai[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) ar[q] = ai[q] = 0.0;
  // the blocks before b, in finishing order: step st is block c(st)
  const int before = lower ? b : nblk - 1 - b;
  // this wave's 16 rows: row r0 + 16 w + q for q < rlim; the addresses of
  // rows past n and of columns past n are clamped to row r0, column 0, and
  // their values taken as zero
  const TF* base = LU + 2 * (size_t)r0 * lda;
  const int rlim = rows - wave * 16;
  size_t roff[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) roff[q] = (q < rlim) ? (size_t)(wave * 16 + q) * lda : 0;
  // The last two blocks before b are on the chain: their factor rows are
  // loaded now, and their values polled directly as soon as the earlier
  // blocks are in.  The blocks before them have at least two links of the
  // chain to spare: they are taken in batches as their flags show them
  // finished, with an interval between flag polls that grows with the
  // number of blocks still to come before b, so that the many workgroups
  // far down the chain put little traffic on the flags.
  TF nr[2][16], ni[2][16];  // [0]: block b-1 (the neighbour), [1]: block b-2
  int ncol0[2];
  bool nb_col[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    ncol0[h] = before > h ? ((lower ? b - 1 - h : b + 1 + h) * SB) : 0;
    nb_col[h] = before > h && ncol0[h] + lane < n;
    const int cc = nb_col[h] ? ncol0[h] + lane : 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const TF* v = base + 2 * (roff[q] + cc);
      nr[h][q] = ni[h][q] = TF(0);
      if (nb_col[h] && q < rlim) ldc(v, nr[h][q], ni[h][q]);
    }
  }
  const int* fl = flags + (lower ? 0 : nblk);
  const int others = max(0, before - 2);
  // older blocks one at a time, in finishing order: the rows of the next
  // one are loaded while its flag is awaited (its values only after the
  // flag, so that no load issued after them holds the values back)
  TF pr[16], pi[16];
  auto older_rows = [&](int st) __attribute__((always_inline)) {
    const int c0 = (lower ? st : nblk - 1 - st) * SB;
    const bool col = st < others && c0 + lane < n;
    const int cc = col ? c0 + lane : 0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const TF* v = base + 2 * (roff[q] + cc);
      pr[q] = pi[q] = TF(0);
      if (col && q < rlim) ldc(v, pr[q], pi[q]);
    }
  };
  if (others > 0) older_rows(0);
  long spins = 0;
  for (int st = 0; st < others;) {
    const int c = lower ? st : nblk - 1 - st;
    if (xload_i(fl + c) == 0) {
      // bounded wait; a timeout is sticky so every later wait returns at once
      if (xload_i(status) != 0 || ++spins > (1L << 22)) {
        __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __builtin_amdgcn_s_sleep(2);
        continue;
      }
    }
    const bool col = c * SB + lane < n;
    const double yr = col ? xload(out + 2 * (c * SB + lane)) : 0.0;
    const double yi = col ? xload(out + 2 * (c * SB + lane) + 1) : 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const double lr = pr[q], li = pi[q];
      ar[q] += lr * yr - li * yi;
      ai[q] += lr * yi + li * yr;
    }
    ++st;
    older_rows(st);
  }
  // block b-2, then the neighbour b-1, in finishing order
#pragma unroll
  for (int h = 1; h >= 0; --h) {
    if (nb_col[h]) {
      const double yr = wait_x(out + 2 * (ncol0[h] + lane), status);
      const double yi = wait_x(out + 2 * (ncol0[h] + lane) + 1, status);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const double lr = nr[h][q], li = ni[h][q];
        ar[q] += lr * yr - li * yi;
        ai[q] += lr * yi + li * yr;
      }
    }
  }
  // sums over the lanes: thread (row r, quarter p) adds lanes 16p .. 16p+15
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    sRed[2 * ((wave * 16 + q) * RS + lane)] = ar[q];
    sRed[2 * ((wave * 16 + q) * RS + lane) + 1] = ai[q];
  }
  __syncthreads();
  {
    const int r = t >> 2, p = t & 3;
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      sr += sRed[2 * (r * RS + 16 * p + j)];
      si += sRed[2 * (r * RS + 16 * p + j) + 1];
    }
    sPart[p][2 * r] = sr;
    sPart[p][2 * r + 1] = si;
  }
  __syncthreads();
  if (t < SB) {
    const double vr = rhs_r, vi = rhs_i;
    const double sr = ((sPart[0][2 * t] + sPart[1][2 * t]) + sPart[2][2 * t]) + sPart[3][2 * t];
    const double si = ((sPart[0][2 * t + 1] + sPart[1][2 * t + 1]) + sPart[2][2 * t + 1]) + sPart[3][2 * t + 1];
    sAcc[2 * t] = (t < rows) ? vr - sr : 0.0;
    sAcc[2 * t + 1] = (t < rows) ? vi - si : 0.0;
  }
  __syncthreads();
  const int r = t & 63;     // row within the block
  const int part = t >> 6;  // column quarter of this thread
  if (!by_subst) {
    // x_b = inverse * acc
    double xr = 0.0, xi = 0.0;
    for (int j = part; j < SB; j += 4) {
      const double ir = sInv[2 * (j * IS + r)], ii = sInv[2 * (j * IS + r) + 1];
      const double vr = sAcc[2 * j], vi = sAcc[2 * j + 1];
      xr += ir * vr - ii * vi;
      xi += ir * vi + ii * vr;
    }
    sPart[part][2 * r] = xr;
    sPart[part][2 * r + 1] = xi;
    __syncthreads();
    if (t < rows) {
      const double x0 = canonical_x(((sPart[0][2 * t] + sPart[1][2 * t]) + sPart[2][2 * t]) + sPart[3][2 * t]);
      const double x1 =
          canonical_x(((sPart[0][2 * t + 1] + sPart[1][2 * t + 1]) + sPart[2][2 * t + 1]) + sPart[3][2 * t + 1]);
      xstore(out + 2 * (r0 + t), x0);
      xstore(out + 2 * (r0 + t) + 1, x1);
      if (xcopy) {
        xcopy[2 * (r0 + t)] = x0;
        xcopy[2 * (r0 + t) + 1] = x1;
      }
    }
  } else if (t < SB) {
    // substitution in the first wave, lane t holding row t: step q takes
    // x_q from lane q and subtracts its column from the rows still open
    double vr = sAcc[2 * t], vi = sAcc[2 * t + 1];
    if (lower) {
      for (int q = 0; q < SB - 1; ++q) {
        const double yr = readlane_d(vr, q), yi = readlane_d(vi, q);
        if (t > q) {
          const double lr = sInv[2 * (q * IS + t)], li = sInv[2 * (q * IS + t) + 1];
          vr -= lr * yr - li * yi;
          vi -= lr * yi + li * yr;
        }
      }
    } else {
      for (int q = SB - 1; q >= 0; --q) {
        if (t == q) {
          const zc x = zdiv(zmk(vr, vi), zmk(sInv[2 * (q * IS + q)], sInv[2 * (q * IS + q) + 1]));
          vr = x.re;
          vi = x.im;
        }
        const double xr = readlane_d(vr, q), xi = readlane_d(vi, q);
        if (t < q) {
          const double ur = sInv[2 * (q * IS + t)], ui = sInv[2 * (q * IS + t) + 1];
          vr -= ur * xr - ui * xi;
          vi -= ur * xi + ui * xr;
        }
      }
    }
    if (t < rows) {
      xstore(out + 2 * (r0 + t), canonical_x(vr));
      xstore(out + 2 * (r0 + t) + 1, canonical_x(vi));
      if (xcopy) {
        xcopy[2 * (r0 + t)] = canonical_x(vr);
        xcopy[2 * (r0 + t) + 1] = canonical_x(vi);
      }
    }
  }
  // the flag once this block's values have reached the coherence point
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (t == 0) __hip_atomic_store(flags + (lower ? 0 : nblk) + b, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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

// Status word and flags to zero, both sweeps' outputs to kUnwritten.
__global__ void solve_reset_kernel(int* status, int* flags, int nflags, double* a, double* b, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *status = 0;
  if (i < nflags) flags[i] = 0;
  const double u = __longlong_as_double(kUnwritten);
  if (i < m) {
    a[i] = u;
    b[i] = u;
  }
}

struct SolveWork {
  int* flags;  // per block and sweep: its values are final
  int* status;
  double* tmp;   // the lower sweep's output
  double* tmp2;  // the upper sweep's output (copied to x as it is formed)
  int* subst;
  double* inv;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

SolveWork carve_solve(void* work, int n) {
  const int nblk = (n + SB - 1) / SB;
  char* p = static_cast<char*>(work);
  SolveWork w;
  w.flags = reinterpret_cast<int*>(p);
  p += align256((size_t)2 * nblk * sizeof(int));
  w.status = reinterpret_cast<int*>(p);
  p += 256;
  w.tmp = reinterpret_cast<double*>(p);
  p += align256((size_t)n * 16);
  w.tmp2 = reinterpret_cast<double*>(p);
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
int apply(const TF* LU, long long lda, int n, const int* piv, double* x, void* work, hipStream_t stream) {
  if (n <= 0) return 0;
  const int nblk = (n + SB - 1) / SB;
  const SolveWork w = carve_solve(work, n);
  const int* perm = piv + n;
  launch_k(K_SOLVE_RESET, solve_reset_kernel, dim3((2 * n + 255) / 256), dim3(256), 0u, stream, w.status, w.flags,
           2 * nblk, w.tmp, w.tmp2, 2 * n);
  launch_k(K_TRSV_LOWER, trsv_kernel<TF, true>, dim3(nblk), dim3(256), 0u, stream, LU, lda, n, perm,
           (const double*)x, w.tmp, (double*)nullptr, (const double*)w.inv, (const int*)w.subst, w.flags, w.status);
  launch_k(K_TRSV_UPPER, trsv_kernel<TF, false>, dim3(nblk), dim3(256), 0u, stream, LU, lda, n, perm,
           (const double*)w.tmp, w.tmp2, x, (const double*)w.inv, (const int*)w.subst, w.flags, w.status);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace

size_t solve_workspace_bytes(int n) {
  const int nblk = (n + SB - 1) / SB;
  return align256((size_t)2 * nblk * sizeof(int)) + 256 + 2 * align256((size_t)n * 16) +
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
                          hipStream_t stream) {
  return apply<double>(LU, lda, n, piv, x, work, stream);
}
int launch_lu_solve_apply_f32(const float* LU, long long lda, int n, const int* piv, double* x, void* work,
                              hipStream_t stream) {
  return apply<float>(LU, lda, n, piv, x, work, stream);
}

int launch_lu_solve_fp64(const double* LU, long long lda, int n, const int* piv, double* x, void* work,
                         hipStream_t stream) {
  if (prepare<double>(LU, lda, n, work, stream)) return 1;
  return apply<double>(LU, lda, n, piv, x, work, stream);
}

}  // namespace necgpu
