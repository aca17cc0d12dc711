// current.hip -- basis amplitudes -> segment current coefficients.
//
// c_geometry::get_current_coefficients (src/c_geometry.cpp:3299-3386)
// scatters, for every basis i in ascending order, the tbf(i) coefficients
// times the amplitude x_i into the segments of its support.  Here each
// segment gathers the same terms from a list sorted by basis index (tbf order
// within one basis), so every sum is formed in the reference's order and no
// atomics are needed.
#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

__global__ void __launch_bounds__(256) currents_kernel(int n, const double* __restrict__ x,
                                                       const int* __restrict__ segptr,
                                                       const int* __restrict__ segbasis,
                                                       const double* __restrict__ segabc, double* __restrict__ abc,
                                                       double* __restrict__ curx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double air = 0.0, aii = 0.0, bir = 0.0, bii = 0.0, cir = 0.0, cii = 0.0;
  for (int q = segptr[j]; q < segptr[j + 1]; ++q) {
    const int i = segbasis[q];
    const double ar = x[2 * i], ai = x[2 * i + 1];
    const double a = segabc[3 * q], b = segabc[3 * q + 1], c = segabc[3 * q + 2];
    air += a * ar;
    aii += a * ai;
    bir += b * ar;
    bii += b * ai;
    cir += c * ar;
    cii += c * ai;
  }
  abc[j] = air;
  abc[n + j] = aii;
  abc[2 * n + j] = bir;
  abc[3 * n + j] = bii;
  abc[4 * n + j] = cir;
  abc[5 * n + j] = cii;
  curx[2 * j] = air + cir;
  curx[2 * j + 1] = aii + cii;
}

// netwk_compute_inputs (src/nec_context.cpp:4937-5009): for every voltage
// source, I = curx[seg] * lambda, Z = V / I, Y = I / V, P = Re(V conj(I)) / 2.
// out[8*s ..]: Z (2), Y (2), I (2), P, 0.
__global__ void zin_kernel(int nsrc, const int* __restrict__ seg, const double* __restrict__ volt,
                           const double* __restrict__ curx, double wavelength, double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsrc) return;
  const int k = seg[s];
  const zc v = zmk(volt[2 * s], volt[2 * s + 1]);
  const zc cur = zmulr(zmk(curx[2 * k], curx[2 * k + 1]), wavelength);
  const zc z = zdiv(v, cur);
  const zc y = zdiv(cur, v);
  double* o = out + 8 * s;
  o[0] = z.re;
  o[1] = z.im;
  o[2] = y.re;
  o[3] = y.im;
  o[4] = cur.re;
  o[5] = cur.im;
  o[6] = 0.5 * zmul(v, zconj(cur)).re;
  o[7] = 0.0;
}

}  // namespace

int launch_zin(int nsrc, const int* seg, const double* volt, const double* curx, double wavelength, double* out,
               hipStream_t stream) {
  if (nsrc <= 0) return 0;
  launch_k(K_ZIN, zin_kernel, dim3((nsrc + 63) / 64), dim3(64), 0, stream, nsrc, seg, volt, curx, wavelength, out);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_currents(int n, const double* x, const int* segptr, const int* segbasis, const double* segabc, double* abc,
                    double* curx, hipStream_t stream) {
  launch_k(K_CURRENTS, currents_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, n, x, segptr, segbasis, segabc, abc,
                     curx);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace necgpu
