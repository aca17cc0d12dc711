// batch.hip -- many small solves at once (frequency sweeps, design
// populations): LU, solve + current coefficients + Z_in, and a pattern cut,
// one workgroup per solve.  The fill is batch_fill_kernel in fill.hip.
//
// The algorithms are those of the single-structure path (the reference's
// lu_decompose / solve, matrix_algebra.cpp:217-307; get_current_coefficients,
// c_geometry.cpp:3299-3386; netwk_compute_inputs, nec_context.cpp:4937-5009;
// ffld, nec_context.cpp:6639-6887, and the total gain of
// nec_radiation_pattern.cpp:282-334).  What differs is the layout: a
// structure of at most 512 segments is factored by one workgroup (BT = 256
// or 512 threads, one per row), with the panel in LDS and the pivot rows
// tracked by a permutation instead of being moved.
#include <climits>
#include <cstdlib>

#include "basis.h"
#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

// First maximum over the block: thread t offers value v for row t
// (negative = no candidate); returns the winning row or -1.  Every thread
// must call it.
template <int BT>
__device__ int block_first_max(double v, double* s_val, int* s_row) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double m = v;
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  const unsigned long long hit = __ballot(v >= 0.0 && v == m);
  if (lane == 0) {
    s_val[wave] = m;
    s_row[wave] = hit ? wave * 64 + (__ffsll((long long)hit) - 1) : -1;
  }
  __syncthreads();
  int best = -1;
  double bv = -1.0;
  for (int w = 0; w < BT / 64; ++w)
    if (s_row[w] >= 0 && (best < 0 || s_val[w] > bv)) {
      bv = s_val[w];
      best = s_row[w];
    }
  __syncthreads();
  return best;
}

// Factor the panel of logical rows k0 .. n-1 and columns k0 .. k0+P-1 (the
// matrix's rows through the permutation s_rp) in LDS, column-major
// (sP[c * BT + r]: logical row k0 + r, column k0 + c), with partial pivoting
// (first maximum of |z|), and write it back; sP keeps the factored panel.
template <int BT, int P>
__device__ __forceinline__ void batch_panel(double2* __restrict__ Z, int n, int k0, double2* sP, int* s_rp,
                                            double* s_val, int* s_row, int& first_zero) {
  const int t = threadIdx.x;
  const int pw = min(P, n - k0);
  const int m = n - k0;
  if (t < m) {
    const double2* src = Z + (long long)s_rp[k0 + t] * n + k0;
#pragma unroll
    for (int c = 0; c < P; ++c) sP[c * BT + t] = c < pw ? src[c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int c = 0; c < pw; ++c) {
    const int k = k0 + c;
    double v = -1.0;
    if (t >= c && t < m) v = hypot(sP[c * BT + t].x, sP[c * BT + t].y);
    const int pr = block_first_max<BT>(v, s_val, s_row);
    if (pr >= 0 && pr != c) {
      if (t < P) {
        const double2 a = sP[t * BT + c];
        sP[t * BT + c] = sP[t * BT + pr];
        sP[t * BT + pr] = a;
      }
      if (t == 0) {
        const int tmp = s_rp[k0 + c];
        s_rp[k0 + c] = s_rp[k0 + pr];
        s_rp[k0 + pr] = tmp;
      }
    }
    __syncthreads();
    const double2 dp = sP[c * BT + c];
    const zc dpiv = zmk(dp.x, dp.y);
    const bool zero = (dpiv.re == 0.0 && dpiv.im == 0.0);
    if (zero && first_zero == 0) first_zero = k + 1;
    if (!zero && t > c && t < m) {
      const double2 a = sP[c * BT + t];
      const zc l = zdiv(zmk(a.x, a.y), dpiv);
      sP[c * BT + t] = make_double2(l.re, l.im);
      for (int cc = c + 1; cc < pw; ++cc) {
        const double2 u = sP[cc * BT + c];
        const zc pu = zmul(l, zmk(u.x, u.y));
        double2 e = sP[cc * BT + t];
        e.x -= pu.re;
        e.y -= pu.im;
        sP[cc * BT + t] = e;
      }
    }
    __syncthreads();
  }
  if (t < m) {
    double2* dst = Z + (long long)s_rp[k0 + t] * n + k0;
    for (int c = 0; c < pw; ++c) dst[c] = sP[c * BT + t];
  }
}

// In-place LU with partial pivoting (first maximum of |z|, as the
// single-structure path) of each solve's n x n matrix, n <= BT.
//
// The panel (logical rows k0 .. n-1, columns k0 .. k0+P-1) is held in LDS
// column-major, so a wave reading one column over 64 rows touches 64
// consecutive 16-byte entries (no bank conflicts) and a wave reading one row
// broadcasts.  The trailing update maps lanes to columns: thread t of row
// group g owns column k0 + pw + t % W and rows pw + g, pw + g + G, ... of the
// block, so every global access of a wave is one contiguous row segment.  Each
// thread forms its column of U12 = L11^-1 A12 in registers (the G groups
// repeat that small triangular solve rather than exchange it), then
// subtracts L21 U12 from its rows, four rows at a time.
template <int BT, int P>
__global__ void __launch_bounds__(BT) batch_lu_kernel(const BatchSolve* __restrict__ sol, double* __restrict__ Zall,
                                                      int* __restrict__ rpall, int* __restrict__ info) {
  __shared__ double2 sP[P * BT];  // sP[c * BT + r]: logical row k0 + r, column k0 + c
  __shared__ int s_rp[BT];
  __shared__ double s_val[BT / 64];
  __shared__ int s_row[BT / 64];
  const BatchSolve d = sol[blockIdx.x];
  const int n = d.n;
  const int t = threadIdx.x;
  double2* Z = reinterpret_cast<double2*>(Zall) + d.z0;
  if (t < n) s_rp[t] = t;
  int first_zero = 0;
  __syncthreads();
  for (int k0 = 0; k0 < n; k0 += P) {
    const int pw = min(P, n - k0);
    const int m = n - k0;
    batch_panel<BT, P>(Z, n, k0, sP, s_rp, s_val, s_row, first_zero);
    const int w = m - pw;  // trailing columns
    if (w > 0) {
      const int W = (w + 63) & ~63;
      const int G = BT / W;  // row groups (>= 1: w < n <= BT)
      const int col = t % W, grp = t / W;
      if (col < w && grp < G) {
        const int j = k0 + pw + col;
        double ur[P], ui[P];
#pragma unroll
        for (int c = 0; c < P; ++c) {
          double2 a = make_double2(0.0, 0.0);
          if (c < pw) a = Z[(long long)s_rp[k0 + c] * n + j];
          ur[c] = a.x;
          ui[c] = a.y;
        }
#pragma unroll
        for (int c = 1; c < P; ++c) {
          if (c < pw) {
#pragma unroll
            for (int c2 = 0; c2 < c; ++c2) {
              const double2 l = sP[c2 * BT + c];
              ur[c] -= l.x * ur[c2] - l.y * ui[c2];
              ui[c] -= l.x * ui[c2] + l.y * ur[c2];
            }
          }
        }
        if (grp == 0) {
#pragma unroll
          for (int c = 0; c < P; ++c)
            if (c < pw) Z[(long long)s_rp[k0 + c] * n + j] = make_double2(ur[c], ui[c]);
        }
        int r = pw + grp;
        for (; r + 3 * G < m; r += 4 * G) {
          double2* z0 = Z + (long long)s_rp[k0 + r] * n + j;
          double2* z1 = Z + (long long)s_rp[k0 + r + G] * n + j;
          double2* z2 = Z + (long long)s_rp[k0 + r + 2 * G] * n + j;
          double2* z3 = Z + (long long)s_rp[k0 + r + 3 * G] * n + j;
          double2 a0 = *z0, a1 = *z1, a2 = *z2, a3 = *z3;
#pragma unroll
          for (int c = 0; c < P; ++c) {
            if (c < pw) {
              const double2 l0 = sP[c * BT + r], l1 = sP[c * BT + r + G];
              const double2 l2 = sP[c * BT + r + 2 * G], l3 = sP[c * BT + r + 3 * G];
              a0.x -= l0.x * ur[c] - l0.y * ui[c];
              a0.y -= l0.x * ui[c] + l0.y * ur[c];
              a1.x -= l1.x * ur[c] - l1.y * ui[c];
              a1.y -= l1.x * ui[c] + l1.y * ur[c];
              a2.x -= l2.x * ur[c] - l2.y * ui[c];
              a2.y -= l2.x * ui[c] + l2.y * ur[c];
              a3.x -= l3.x * ur[c] - l3.y * ui[c];
              a3.y -= l3.x * ui[c] + l3.y * ur[c];
            }
          }
          *z0 = a0;
          *z1 = a1;
          *z2 = a2;
          *z3 = a3;
        }
        for (; r < m; r += G) {
          double2* z0 = Z + (long long)s_rp[k0 + r] * n + j;
          double2 a0 = *z0;
#pragma unroll
          for (int c = 0; c < P; ++c) {
            if (c < pw) {
              const double2 l0 = sP[c * BT + r];
              a0.x -= l0.x * ur[c] - l0.y * ui[c];
              a0.y -= l0.x * ui[c] + l0.y * ur[c];
            }
          }
          *z0 = a0;
        }
      }
    }
    __syncthreads();
  }
  if (t < n) rpall[d.row0 + t] = s_rp[t];
  if (t == 0) info[blockIdx.x] = first_zero;
}

// batch_lu_kernel with the trailing update on FP64 MFMA (BT = 256, four
// waves).  After each panel, thread t < w forms column t of U12 = L11^-1 A12
// in registers and writes it to the matrix; then the (m - pw) x w trailing
// block is updated in 16 x 16 tiles of v_mfma_f64_16x16x4f64, four real
// products per complex one (acc_re += (-Lr) Ur + Li Ui, acc_im += (-Lr) Ui +
// (-Li) Ur, from acc = C).  Wave v takes the tile columns v, v + 4, ...,
// holds that column strip's U12 fragments in registers and walks down its
// row tiles: C through the row permutation (lane l, register q: row
// l / 16 + 4 q, column l % 16 of the tile), L21 from the panel in LDS.
template <int BT, int P>
__global__ void __launch_bounds__(BT) batch_lu_mfma_kernel(const BatchSolve* __restrict__ sol,
                                                           double* __restrict__ Zall, int* __restrict__ rpall,
                                                           int* __restrict__ info) {
  static_assert(P % 4 == 0, "MFMA k-steps of 4");
  typedef double double4_t __attribute__((ext_vector_type(4)));
  __shared__ double2 sP[P * BT];
  __shared__ int s_rp[BT];
  __shared__ double s_val[BT / 64];
  __shared__ int s_row[BT / 64];
  const BatchSolve d = sol[blockIdx.x];
  const int n = d.n;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int l16 = lane & 15, l4 = lane >> 4;
  double2* Z = reinterpret_cast<double2*>(Zall) + d.z0;
  if (t < n) s_rp[t] = t;
  int first_zero = 0;
  __syncthreads();
  for (int k0 = 0; k0 < n; k0 += P) {
    const int pw = min(P, n - k0);
    const int m = n - k0;
    batch_panel<BT, P>(Z, n, k0, sP, s_rp, s_val, s_row, first_zero);
    const int w = m - pw;  // trailing columns (and rows)
    if (w > 0) {
      if (t < w) {
        const int j = k0 + pw + t;
        double ur[P], ui[P];
#pragma unroll
        for (int c = 0; c < P; ++c) {
          double2 a = make_double2(0.0, 0.0);
          if (c < pw) a = Z[(long long)s_rp[k0 + c] * n + j];
          ur[c] = a.x;
          ui[c] = a.y;
        }
#pragma unroll
        for (int c = 1; c < P; ++c) {
          if (c < pw) {
#pragma unroll
            for (int c2 = 0; c2 < c; ++c2) {
              const double2 l = sP[c2 * BT + c];
              ur[c] -= l.x * ur[c2] - l.y * ui[c2];
              ui[c] -= l.x * ui[c2] + l.y * ur[c2];
            }
          }
        }
#pragma unroll
        for (int c = 0; c < P; ++c)
          if (c < pw) Z[(long long)s_rp[k0 + c] * n + j] = make_double2(ur[c], ui[c]);
      }
      __syncthreads();  // U12 in the matrix, visible to the workgroup
      const int tiles_r = (w + 15) / 16, tiles_c = (w + 15) / 16;
      for (int tc = wave; tc < tiles_c; tc += BT / 64) {
        // U12 fragments of this column strip: k-step s, lane (j, kk) holds
        // U12[4 s + kk][16 tc + j]
        const int jj = 16 * tc + l16;
        double br[P / 4], bi[P / 4];
#pragma unroll
        for (int s = 0; s < P / 4; ++s) {
          const int c = 4 * s + l4;
          double2 b = make_double2(0.0, 0.0);
          if (c < pw && jj < w) b = Z[(long long)s_rp[k0 + c] * n + k0 + pw + jj];
          br[s] = b.x;
          bi[s] = b.y;
        }
        // C of row tile tr (the loads for tile tr + 2 go out before tile tr's MFMAs)
        auto load_c = [&](int tr, double2(&c)[4]) __attribute__((always_inline)) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int r = pw + 16 * tr + l4 + 4 * q;
            c[q] = make_double2(0.0, 0.0);
            if (r < m && jj < w) c[q] = Z[(long long)s_rp[k0 + r] * n + k0 + pw + jj];
          }
        };
        double2 cv[4], cn[4], cn2[4];
        load_c(0, cv);
        if (1 < tiles_r) load_c(1, cn);
        for (int tr = 0; tr < tiles_r; ++tr) {
          const int r0 = pw + 16 * tr;  // logical row (from k0) of the tile's first row
          if (tr + 2 < tiles_r) load_c(tr + 2, cn2);
          double4_t accr = (double4_t){cv[0].x, cv[1].x, cv[2].x, cv[3].x};
          double4_t acci = (double4_t){cv[0].y, cv[1].y, cv[2].y, cv[3].y};
#pragma unroll
          for (int s = 0; s < P / 4; ++s) {
            // L21 fragment: lane (i, kk) holds L[r0 + i][4 s + kk]
            const int c = 4 * s + l4, r = r0 + l16;
            double2 a = make_double2(0.0, 0.0);
            if (c < pw && r < m) a = sP[c * BT + r];
            accr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.x, br[s], accr, 0, 0, 0);
            accr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, bi[s], accr, 0, 0, 0);
            acci = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.x, bi[s], acci, 0, 0, 0);
            acci = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, br[s], acci, 0, 0, 0);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int r = r0 + l4 + 4 * q;
            if (r < m && jj < w) Z[(long long)s_rp[k0 + r] * n + k0 + pw + jj] = make_double2(accr[q], acci[q]);
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            cv[q] = cn[q];
            cn[q] = cn2[q];
          }
        }
      }
    }
    __syncthreads();
  }
  if (t < n) rpall[d.row0 + t] = s_rp[t];
  if (t == 0) info[blockIdx.x] = first_zero;
}

// Forward / back substitution with the factors and permutation, then the
// current coefficients (gather over the segment CSR in ascending basis order)
// and the antenna input: zin[4 s ..] = Z_in (2), input power, 0.
template <int BT>
__global__ void __launch_bounds__(BT) batch_solve_kernel(const BatchSolve* __restrict__ sol,
                                                         const double* __restrict__ Zall, const int* __restrict__ rpall,
                                                         const double* __restrict__ seg, const int* __restrict__ segptr,
                                                         const int* __restrict__ segbasis,
                                                         const double* __restrict__ segabc, double* __restrict__ xall,
                                                         double* __restrict__ abcall, double* __restrict__ zin) {
  __shared__ double sx[2 * BT];
  __shared__ int s_rp[BT];
  const BatchSolve d = sol[blockIdx.x];
  const int n = d.n;
  const int t = threadIdx.x;
  const double* Z = Zall + 2 * d.z0;
  if (t < n) {
    s_rp[t] = rpall[d.row0 + t];
  }
  __syncthreads();
  // (P b)[i] = b[rp[i]]; b is zero except at the source
  if (t < n) {
    const bool src = s_rp[t] == d.src;
    sx[2 * t] = src ? d.er : 0.0;
    sx[2 * t + 1] = src ? d.ei : 0.0;
  }
  // L y = P b (unit lower), column by column as the reference's solve
  for (int q = 0; q < n; ++q) {
    __syncthreads();
    if (t > q && t < n) {
      const double yr = sx[2 * q], yi = sx[2 * q + 1];
      const double* l = Z + 2 * ((long long)s_rp[t] * n + q);
      sx[2 * t] -= l[0] * yr - l[1] * yi;
      sx[2 * t + 1] -= l[0] * yi + l[1] * yr;
    }
  }
  // U x = y
  for (int q = n - 1; q >= 0; --q) {
    __syncthreads();
    if (t == q) {
      const double* u = Z + 2 * ((long long)s_rp[q] * n + q);
      const zc x = zdiv(zmk(sx[2 * q], sx[2 * q + 1]), zmk(u[0], u[1]));
      sx[2 * q] = x.re;
      sx[2 * q + 1] = x.im;
    }
    __syncthreads();
    if (t < q) {
      const double xr = sx[2 * q], xi = sx[2 * q + 1];
      const double* u = Z + 2 * ((long long)s_rp[t] * n + q);
      sx[2 * t] -= u[0] * xr - u[1] * xi;
      sx[2 * t + 1] -= u[0] * xi + u[1] * xr;
    }
  }
  __syncthreads();
  double* x = xall + 2 * d.row0;
  if (t < n) {
    x[2 * t] = sx[2 * t];
    x[2 * t + 1] = sx[2 * t + 1];
  }
  // current coefficients of segment t
  const int* sp = segptr + d.sptr0;
  const int* sb = segbasis + d.ssr0;
  const double* sv = segabc + 3 * d.ssr0;
  double* abc = abcall + 6 * d.row0;
  if (t < n) {
    double air = 0.0, aii = 0.0, bir = 0.0, bii = 0.0, cir = 0.0, cii = 0.0;
    for (int q = sp[t]; q < sp[t + 1]; ++q) {
      const int i = sb[q];
      const double ar = sx[2 * i], ai = sx[2 * i + 1];
      const double a = sv[3 * q], b = sv[3 * q + 1], c = sv[3 * q + 2];
      air += a * ar;
      aii += a * ai;
      bir += b * ar;
      bii += b * ai;
      cir += c * ar;
      cii += c * ai;
    }
    abc[t] = air;
    abc[n + t] = aii;
    abc[2 * n + t] = bir;
    abc[3 * n + t] = bii;
    abc[4 * n + t] = cir;
    abc[5 * n + t] = cii;
    if (t == d.src) {
      const zc v = zmk(d.vr, d.vi);
      const zc cur = zmulr(zmk(air + cir, aii + cii), d.wavelength);
      const zc z = zdiv(v, cur);
      double* o = zin + 4 * (long long)blockIdx.x;
      o[0] = z.re;
      o[1] = z.im;
      o[2] = 0.5 * zmul(v, zconj(cur)).re;
      o[3] = 0.0;
    }
  }
}

// Z^T lam = g for each solve, with the factors and permutation of
// batch_lu_kernel (P Z = L U, row i of P Z = physical row rp[i]):
// U^T w = g (forward), L^T v = w (backward, unit), lam[rp[i]] = v[i].
// Every step reads one contiguous factor row.
template <int BT>
__global__ void __launch_bounds__(BT) batch_solve_transposed_kernel(const BatchSolve* __restrict__ sol,
                                                                    const double* __restrict__ Zall,
                                                                    const int* __restrict__ rpall,
                                                                    const double* __restrict__ gall,
                                                                    double* __restrict__ lamall) {
  __shared__ double sx[2 * BT];
  __shared__ int s_rp[BT];
  const BatchSolve d = sol[blockIdx.x];
  const int n = d.n;
  const int t = threadIdx.x;
  const double* Z = Zall + 2 * d.z0;
  const double* g = gall + 2 * d.row0;
  if (t < n) {
    s_rp[t] = rpall[d.row0 + t];
    sx[2 * t] = g[2 * t];
    sx[2 * t + 1] = g[2 * t + 1];
  }
  for (int q = 0; q < n; ++q) {
    __syncthreads();
    if (t == q) {
      const double* u = Z + 2 * ((long long)s_rp[q] * n + q);
      const zc w = zdiv(zmk(sx[2 * q], sx[2 * q + 1]), zmk(u[0], u[1]));
      sx[2 * q] = w.re;
      sx[2 * q + 1] = w.im;
    }
    __syncthreads();
    if (t > q && t < n) {
      const double wr = sx[2 * q], wi = sx[2 * q + 1];
      const double* u = Z + 2 * ((long long)s_rp[q] * n + t);
      sx[2 * t] -= u[0] * wr - u[1] * wi;
      sx[2 * t + 1] -= u[0] * wi + u[1] * wr;
    }
  }
  for (int q = n - 1; q >= 0; --q) {
    __syncthreads();
    if (t < q) {
      const double vr = sx[2 * q], vi = sx[2 * q + 1];
      const double* l = Z + 2 * ((long long)s_rp[q] * n + t);
      sx[2 * t] -= l[0] * vr - l[1] * vi;
      sx[2 * t + 1] -= l[0] * vi + l[1] * vr;
    }
  }
  __syncthreads();
  double* lam = lamall + 2 * d.row0;
  if (t < n) {
    lam[2 * s_rp[t]] = sx[2 * t];
    lam[2 * s_rp[t] + 1] = sx[2 * t + 1];
  }
}

// Total power gain (dB) of each solve in n_dir directions: the ffld sum over
// segments staged in LDS, then db10(gcon (|E_theta|^2 + |E_phi|^2)) with
// gcon = lambda^2 2 pi / (eta P_in).
template <int BT>
__global__ void __launch_bounds__(BT) batch_pattern_kernel(const BatchSolve* __restrict__ sol,
                                                           const double* __restrict__ seg,
                                                           const double* __restrict__ abcall,
                                                           const double* __restrict__ zin, int n_dir,
                                                           const double* __restrict__ dir, double* __restrict__ gain) {
  constexpr int F = 13;
  __shared__ double ss[kBatchMaxN * F];
  const BatchSolve d = sol[blockIdx.x];
  const int n = d.n;
  const double* p = seg + d.seg0;
  const double* abc = abcall + 6 * d.row0;
  for (int e = threadIdx.x; e < n * F; e += blockDim.x) {
    const int i = e / F, f = e % F;
    double v;
    switch (f) {
      case 0: v = p[5 * n + i]; break;       // cab
      case 1: v = p[6 * n + i]; break;       // sab
      case 2: v = p[7 * n + i]; break;       // salp
      case 3: v = kPi * p[3 * n + i]; break; // pi * s
      case 4: v = p[i]; break;               // x
      case 5: v = p[n + i]; break;           // y
      case 6: v = p[2 * n + i]; break;       // z
      default: v = abc[(size_t)(f - 7) * n + i]; break;
    }
    ss[e] = v;
  }
  __syncthreads();
  const double pin = zin[4 * (long long)blockIdx.x + 2];
  const double gcon = d.wavelength * d.wavelength * 2.0 * kPi / (em_impedance() * pin);
  for (int q = threadIdx.x; q < n_dir; q += blockDim.x) {
    const double sth = dir[4 * q], cth = dir[4 * q + 1], sph = dir[4 * q + 2], cph = dir[4 * q + 3];
    const double phx = -sph, phy = cph, roz = cth, thz = -sth;
    const double rox = -thz * phy, roy = thz * phx;
    double cxr = 0, cxi = 0, cyr = 0, cyi = 0, czr = 0, czi = 0;
    for (int i = 0; i < n; ++i) {
      const double* sg = ss + i * F;
      const double cab = sg[0], sab = sg[1], salp = sg[2], el = sg[3];
      const double omega = -(rox * cab + roy * sab + roz * salp);
      const double sill = omega * el;
      const double top = el + sill;
      const double bot = el - sill;
      // 2 sin(sill) / omega = 2 el sinc(sill); sin(top) / top; sin(bot) / bot
      const double a = 2. * el * sinc(sill);
      const double too = sinc(top);
      const double boo = sinc(bot);
      const double b = el * (boo - too);
      const double c = el * (boo + too);
      const double rr = a * sg[7] + b * sg[10] + c * sg[11];
      const double ri = a * sg[8] - b * sg[9] + c * sg[12];
      const double arg = kTwoPi * (sg[4] * rox + sg[5] * roy + sg[6] * roz);
      double sa, ca;
      sincos(arg, &sa, &ca);
      const zc exa = zmul(zmk(ca, sa), zmk(rr, ri));
      cxr = cxr + exa.re * cab;
      cxi = cxi + exa.im * cab;
      cyr = cyr + exa.re * sab;
      cyi = cyi + exa.im * sab;
      czr = czr + exa.re * salp;
      czi = czi + exa.im * salp;
    }
    const double thx = roz * phy, thy = -roz * phx;
    const zc const3 = zmk(0.0, -em_impedance() / kFourPi);
    const zc cix = zmk(cxr, cxi), ciy = zmk(cyr, cyi), ciz = zmk(czr, czi);
    const zc eth = zmul(zadd(zadd(zmulr(cix, thx), zmulr(ciy, thy)), zmulr(ciz, thz)), const3);
    const zc eph = zmul(zadd(zmulr(cix, phx), zmulr(ciy, phy)), const3);
    const double ethm2 = eth.re * eth.re + eth.im * eth.im;
    const double ephm2 = eph.re * eph.re + eph.im * eph.im;
    gain[(long long)blockIdx.x * n_dir + q] = db10(gcon * (ethm2 + ephm2));
  }
}

// Every solve's tables (see launch_batch_tables).  The segment table is
// scaled as Geometry::scaled does (each coordinate, length and radius times
// fr), so it is bit-identical to the host's; trio and tbf run the shared
// basis.h code on the solve's lengths and radii (staged in LDS).
constexpr int kTableCapDev = 64;

__global__ void __launch_bounds__(256) batch_tables_kernel(
    const BatchSolve* __restrict__ sol, const BatchDesign* __restrict__ des, const double* __restrict__ dgeo,
    const int* __restrict__ dicon, const int* __restrict__ dcolptr, const int* __restrict__ dcolj,
    const int* __restrict__ dtptr, const int* __restrict__ dcpos, const int* __restrict__ dsegptr,
    const int* __restrict__ dsegbasis, const int* __restrict__ dbptr, const int* __restrict__ dspos,
    double* __restrict__ seg, int* __restrict__ colptr, int* __restrict__ colj, double* __restrict__ colabc,
    int* __restrict__ segptr, int* __restrict__ segbasis, double* __restrict__ segabc, int* status) {
  __shared__ double sS[kBatchMaxN], sB[kBatchMaxN];
  const BatchSolve d = sol[blockIdx.x];
  const BatchDesign D = des[d.design];
  const int n = d.n;
  const int t = threadIdx.x;
  const double fr = d.fr;
  const double* g = dgeo + D.g0;
  double* o = seg + d.seg0;
  for (int i = t; i < n; i += blockDim.x) {
    o[i] = g[i] * fr;
    o[n + i] = g[n + i] * fr;
    o[2 * n + i] = g[2 * n + i] * fr;
    const double s = g[3 * n + i] * fr, b = g[4 * n + i] * fr;
    o[3 * n + i] = s;
    o[4 * n + i] = b;
    o[5 * n + i] = g[5 * n + i];
    o[6 * n + i] = g[6 * n + i];
    o[7 * n + i] = g[7 * n + i];
    sS[i] = s;
    sB[i] = b;
  }
  const int nnzc = dcolptr[D.cptr0 + n], nnzs = dsegptr[D.sptr0 + n];
  for (int e = t; e <= n; e += blockDim.x) {
    colptr[d.cptr0 + e] = dcolptr[D.cptr0 + e];
    segptr[d.sptr0 + e] = dsegptr[D.sptr0 + e];
  }
  for (int e = t; e < nnzc; e += blockDim.x) colj[d.csr0 + e] = dcolj[D.csr0 + e];
  for (int e = t; e < nnzs; e += blockDim.x) segbasis[d.ssr0 + e] = dsegbasis[D.ssr0 + e];
  __syncthreads();
  const SegView sv{dicon + D.ic0, dicon + D.ic0 + n, sS, sB};
  int jco[kTableCapDev];
  double ax[kTableCapDev], bx[kTableCapDev], cx[kTableCapDev];
  for (int j = t; j < n; j += blockDim.x) {
    int cnt = trio(sv, j + 1, jco, ax, bx, cx, kTableCapDev);
    int e0 = dtptr[D.tptr0 + j];
    if (cnt != dtptr[D.tptr0 + j + 1] - e0) {
      __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      for (int q = 0; q < cnt; ++q) {
        const long long p = d.csr0 + dcpos[D.csr0 + e0 + q];
        colabc[3 * p] = ax[q];
        colabc[3 * p + 1] = bx[q];
        colabc[3 * p + 2] = cx[q];
      }
    }
    cnt = tbf(sv, j + 1, 1, jco, ax, bx, cx, kTableCapDev);
    e0 = dbptr[D.bptr0 + j];
    if (cnt != dbptr[D.bptr0 + j + 1] - e0) {
      __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      for (int q = 0; q < cnt; ++q) {
        const long long p = d.ssr0 + dspos[D.ssr0 + e0 + q];
        segabc[3 * p] = ax[q];
        segabc[3 * p + 1] = bx[q];
        segabc[3 * p + 2] = cx[q];
      }
    }
  }
}

}  // namespace

int launch_batch_tables(const BatchSolve* sol, int S, const BatchDesign* des, const double* dgeo, const int* dicon,
                        const int* dcolptr, const int* dcolj, const int* dtptr, const int* dcpos, const int* dsegptr,
                        const int* dsegbasis, const int* dbptr, const int* dspos, double* seg, int* colptr, int* colj,
                        double* colabc, int* segptr, int* segbasis, double* segabc, int* status, hipStream_t stream) {
  if (S <= 0) return 0;
  launch_k(K_BATCH_TABLES, batch_tables_kernel, dim3((unsigned)S), dim3(256), 0u, stream, sol, des, dgeo, dicon,
           dcolptr, dcolj, dtptr, dcpos, dsegptr, dsegbasis, dbptr, dspos, seg, colptr, colj, colabc, segptr, segbasis,
           segabc, status);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The 256-thread LU's trailing update on MFMA (batch_lu_mfma_kernel, the
// default) or on the vector units (NECGPU_BATCH_MFMA=0, for measurements).
bool batch_mfma_enabled() {
  const char* v = std::getenv("NECGPU_BATCH_MFMA");
  return !(v && v[0] == '0');
}

int launch_batch_lu(const BatchSolve* sol, int S, double* Z, int* rp, int* info, int P, int bt, hipStream_t stream) {
  if (S <= 0) return 0;
  if (bt != 512 && batch_mfma_enabled()) {
    if (P == 32)
      launch_k(K_BATCH_LU, batch_lu_mfma_kernel<256, 32>, dim3(S), dim3(256), 0u, stream, sol, Z, rp, info);
    else if (P == 8)
      launch_k(K_BATCH_LU, batch_lu_mfma_kernel<256, 8>, dim3(S), dim3(256), 0u, stream, sol, Z, rp, info);
    else
      launch_k(K_BATCH_LU, batch_lu_mfma_kernel<256, 16>, dim3(S), dim3(256), 0u, stream, sol, Z, rp, info);
    return hipGetLastError() == hipSuccess ? 0 : 1;
  }
  if (bt == 512) {
    // the 512-row panel takes P * 16 KB of LDS: P = 8 or 16
    if (P == 8)
      launch_k(K_BATCH_LU, batch_lu_kernel<512, 8>, dim3(S), dim3(512), 0u, stream, sol, Z, rp, info);
    else
      launch_k(K_BATCH_LU, batch_lu_kernel<512, 16>, dim3(S), dim3(512), 0u, stream, sol, Z, rp, info);
  } else if (P == 32) {
    launch_k(K_BATCH_LU, batch_lu_kernel<256, 32>, dim3(S), dim3(256), 0u, stream, sol, Z, rp, info);
  } else if (P == 8) {
    launch_k(K_BATCH_LU, batch_lu_kernel<256, 8>, dim3(S), dim3(256), 0u, stream, sol, Z, rp, info);
  } else {
    launch_k(K_BATCH_LU, batch_lu_kernel<256, 16>, dim3(S), dim3(256), 0u, stream, sol, Z, rp, info);
  }
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_batch_solve(const BatchSolve* sol, int S, const double* Z, const int* rp, const double* seg,
                       const int* segptr, const int* segbasis, const double* segabc, double* x, double* abc,
                       double* zin, int bt, hipStream_t stream) {
  if (S <= 0) return 0;
  if (bt == 512)
    launch_k(K_BATCH_SOLVE, batch_solve_kernel<512>, dim3(S), dim3(512), 0u, stream, sol, Z, rp, seg, segptr,
             segbasis, segabc, x, abc, zin);
  else
    launch_k(K_BATCH_SOLVE, batch_solve_kernel<256>, dim3(S), dim3(256), 0u, stream, sol, Z, rp, seg, segptr,
             segbasis, segabc, x, abc, zin);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_batch_solve_transposed(const BatchSolve* sol, int S, const double* Z, const int* rp, const double* g,
                                  double* lam, int bt, hipStream_t stream) {
  if (S <= 0) return 0;
  if (bt == 512)
    launch_k(K_ADJOINT_SOLVE, batch_solve_transposed_kernel<512>, dim3(S), dim3(512), 0u, stream, sol, Z, rp, g, lam);
  else
    launch_k(K_ADJOINT_SOLVE, batch_solve_transposed_kernel<256>, dim3(S), dim3(256), 0u, stream, sol, Z, rp, g, lam);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_batch_pattern(const BatchSolve* sol, int S, const double* seg, const double* abc, const double* zin,
                         int n_dir, const double* dir, double* gain, int bt, hipStream_t stream) {
  if (S <= 0 || n_dir <= 0) return 0;
  if (bt == 512)
    launch_k(K_BATCH_PATTERN, batch_pattern_kernel<512>, dim3(S), dim3(512), 0u, stream, sol, seg, abc, zin, n_dir,
             dir, gain);
  else
    launch_k(K_BATCH_PATTERN, batch_pattern_kernel<256>, dim3(S), dim3(256), 0u, stream, sol, seg, abc, zin, n_dir,
             dir, gain);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace necgpu
