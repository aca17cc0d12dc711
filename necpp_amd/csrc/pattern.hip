// pattern.hip -- far-field radiation pattern (free space or over a ground).
//
// nec_context::ffld (src/nec_context.cpp:6639-6887) for every direction of
// an RP card, then the per-direction part of
// nec_radiation_pattern::analyze (src/nec_radiation_pattern.cpp:209-450):
// polarisation ellipse, dB gains, field scaling.  Directions are
// d = kph * n_theta + kth (the reference's column-major pattern index).
//
// pattern_partial  grid (direction blocks, segment slices); every thread sums
//                  the segment terms of its slice for one direction, with
//                  the slice's segment data staged through LDS; over a
//                  ground also the sum of the structure image (roz -> -roz).
// pattern_finish   one thread per direction: adds the slice partials in
//                  slice order, adds the image with the reflection
//                  coefficients (ffld :6666-6700, :6817-6822), forms
//                  E_theta / E_phi and the gains.
#include "kernels.h"
#include "nec_math.h"

namespace necgpu {

namespace {

constexpr int SEG_CHUNK = 64;
constexpr int SEG_FIELDS = 13;

// The segment term of ffld (src/nec_context.cpp:6706-6742) for one segment
// (staged data sg) and direction cosines (rox, roy, roz).
__device__ __forceinline__ zc ffld_term(const double* sg, double rox, double roy, double roz) {
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
  return zmul(zmk(ca, sa), zmk(rr, ri));
}

// GROUND: also the image sum (roz -> -roz); part holds 6 doubles per
// (slice, direction), 12 with the image (direct then image).
template <bool GROUND>
__global__ void __launch_bounds__(256) pattern_partial_kernel(DevSegments g, int n, const double* __restrict__ abc,
                                                              int n_theta, int n_dir, const double* __restrict__ trig,
                                                              int seg_per_slice, double* __restrict__ part) {
  __shared__ double sseg[SEG_CHUNK * SEG_FIELDS];
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  const int slice = blockIdx.y;
  const int s0 = slice * seg_per_slice;
  const int s1 = min(n, s0 + seg_per_slice);
  const int n_phi = n_dir / n_theta;
  double rox = 0, roy = 0, roz = 0;
  if (d < n_dir) {
    const int kth = d % n_theta, kph = d / n_theta;
    const double sth = trig[kth], cth = trig[n_theta + kth];
    const double sph = trig[3 * n_theta + kph], cph = trig[3 * n_theta + n_phi + kph];
    const double phx = -sph;
    const double phy = cph;
    roz = cth;
    const double thz = -sth;
    rox = -thz * phy;
    roy = thz * phx;
  }
  double cxr = 0, cxi = 0, cyr = 0, cyi = 0, czr = 0, czi = 0;
  double ixr = 0, ixi = 0, iyr = 0, iyi = 0, izr = 0, izi = 0;
  bool below = false;
  if (GROUND && d < n_dir) below = trig[3 * n_theta + 2 * n_phi + d % n_theta] != 0.0;
  for (int c0 = s0; c0 < s1; c0 += SEG_CHUNK) {
    const int cnt = min(SEG_CHUNK, s1 - c0);
    __syncthreads();
    for (int e = threadIdx.x; e < cnt * SEG_FIELDS; e += blockDim.x) {
      const int q = e / SEG_FIELDS, f = e % SEG_FIELDS;
      const int i = c0 + q;
      double v;
      switch (f) {
        case 0: v = g.cab[i]; break;
        case 1: v = g.sab[i]; break;
        case 2: v = g.salp[i]; break;
        case 3: v = kPi * g.s[i]; break;
        case 4: v = g.x[i]; break;
        case 5: v = g.y[i]; break;
        case 6: v = g.z[i]; break;
        default: v = abc[(size_t)(f - 7) * n + i]; break;
      }
      sseg[q * SEG_FIELDS + f] = v;
    }
    __syncthreads();
    if (d < n_dir && !below) {
      for (int q = 0; q < cnt; ++q) {
        const double* sg = sseg + q * SEG_FIELDS;
        const double cab = sg[0], sab = sg[1], salp = sg[2];
        const zc exa = ffld_term(sg, rox, roy, roz);
        cxr = cxr + exa.re * cab;
        cxi = cxi + exa.im * cab;
        cyr = cyr + exa.re * sab;
        cyi = cyi + exa.im * sab;
        czr = czr + exa.re * salp;
        czi = czi + exa.im * salp;
        if (GROUND) {
          const zc exi = ffld_term(sg, rox, roy, -roz);
          ixr = ixr + exi.re * cab;
          ixi = ixi + exi.im * cab;
          iyr = iyr + exi.re * sab;
          iyi = iyi + exi.im * sab;
          izr = izr + exi.re * salp;
          izi = izi + exi.im * salp;
        }
      }
    }
  }
  if (d < n_dir) {
    double* o = part + ((size_t)slice * n_dir + d) * (GROUND ? 12 : 6);
    o[0] = cxr;
    o[1] = cxi;
    o[2] = cyr;
    o[3] = cyi;
    o[4] = czr;
    o[5] = czi;
    if (GROUND) {
      o[6] = ixr;
      o[7] = ixi;
      o[8] = iyr;
      o[9] = iyi;
      o[10] = izr;
      o[11] = izi;
    }
  }
}

__global__ void __launch_bounds__(256) pattern_finish_kernel(int n_dir, int slices, const double* __restrict__ part,
                                                             const double* __restrict__ trig, PatternSpec sp,
                                                             PatternOut out) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= n_dir) return;
  const int n_theta = sp.n_theta, n_phi = sp.n_phi;
  const int kth = d % n_theta, kph = d / n_theta;
  if (sp.ground != 0 && trig[3 * n_theta + 2 * n_phi + kth] != 0.0) {
    // below the horizon over a ground: the reference skips the direction
    // (nec_radiation_pattern.cpp:263-264) and leaves its entries unset;
    // here they read as no field
    out.gain[d] = -999.99;
    out.pg_vert[d] = -999.99;
    out.pg_horiz[d] = -999.99;
    out.pg_tot[d] = -999.99;
    out.pg_rhcp[d] = -999.99;
    out.pg_lhcp[d] = -999.99;
    out.e_theta[2 * d] = 0.0;
    out.e_theta[2 * d + 1] = 0.0;
    out.e_phi[2 * d] = 0.0;
    out.e_phi[2 * d + 1] = 0.0;
    out.axial[d] = 0.0;
    out.tilt[d] = 0.0;
    out.sense[d] = 3;
    out.avg_scale[d] = 0.0;
    if (sp.iavp != 0) out.pint[d] = 0.0;
    return;
  }
  const int pw = sp.ground ? 12 : 6;
  zc cix = zmk(0, 0), ciy = zmk(0, 0), ciz = zmk(0, 0);
  zc iix = zmk(0, 0), iiy = zmk(0, 0), iiz = zmk(0, 0);
  for (int s = 0; s < slices; ++s) {
    const double* p = part + ((size_t)s * n_dir + d) * pw;
    cix = zadd(cix, zmk(p[0], p[1]));
    ciy = zadd(ciy, zmk(p[2], p[3]));
    ciz = zadd(ciz, zmk(p[4], p[5]));
    if (sp.ground) {
      iix = zadd(iix, zmk(p[6], p[7]));
      iiy = zadd(iiy, zmk(p[8], p[9]));
      iiz = zadd(iiz, zmk(p[10], p[11]));
    }
  }
  const double sth = trig[kth], cth = trig[n_theta + kth], tha = trig[2 * n_theta + kth];
  const double sph = trig[3 * n_theta + kph], cph = trig[3 * n_theta + n_phi + kph];
  const double phx = -sph, phy = cph, roz = cth;
  const double thx = roz * phy, thy = -roz * phx, thz = -sth;
  if (sp.ground) {
    // reflection coefficients (ffld :6666-6678), then the image combination
    // (:6817-6822)
    zc rrv, rrh;
    if (sp.ground == 1) {
      rrv = zmk(-1.0, -0.0);
      rrh = zmk(-1.0, -0.0);
    } else {
      const zc zr = zmk(sp.zr_re, sp.zr_im);
      const zc q = zmulr(zmulr(zmul(zr, zr), thz), thz);
      const zc zrsin = zsqrt(zmk(1. - q.re, -q.im));
      const zc zz = zmul(zr, zrsin);
      rrv = zdiv(zneg(zmk(roz - zz.re, -zz.im)), zmk(zz.re + roz, zz.im));
      const zc zro = zmulr(zr, roz);
      rrh = zdiv(zsub(zro, zrsin), zadd(zro, zrsin));
    }
    const zc cdp = zmul(zadd(zmulr(iix, phx), zmulr(iiy, phy)), zsub(rrh, rrv));
    cix = zadd(zadd(cix, zmul(iix, rrv)), zmulr(cdp, phx));
    ciy = zadd(zadd(ciy, zmul(iiy, rrv)), zmulr(cdp, phy));
    ciz = zsub(ciz, zmul(iiz, rrv));
  }
  const zc const3 = zmk(0.0, -em_impedance() / kFourPi);
  const zc eth = zmul(zadd(zadd(zmulr(cix, thx), zmulr(ciy, thy)), zmulr(ciz, thz)), const3);
  const zc eph = zmul(zadd(zmulr(cix, phx), zmulr(ciy, phy)), const3);
  // nec_radiation_pattern.cpp:282-430
  const double ethm2 = eth.re * eth.re + eth.im * eth.im;
  double ethm = sqrt(ethm2);
  double etha = rad_to_deg(atan2(eth.im, eth.re));
  const double ephm2 = eph.re * eph.re + eph.im * eph.im;
  double ephm = sqrt(ephm2);
  double epha = rad_to_deg(atan2(eph.im, eph.re));
  double tilta, emajr2, eminr2, axrat;
  int sense;
  if ((ethm2 <= 1.0e-20) && (ephm2 <= 1.0e-20)) {
    tilta = 0.;
    emajr2 = 0.;
    eminr2 = 0.;
    axrat = 0.;
    sense = 3;
  } else {
    double dfaz = epha - etha, dfaz2;
    if (epha >= 0.) dfaz2 = dfaz - 360.;
    else dfaz2 = dfaz + 360.;
    if (fabs(dfaz) > fabs(dfaz2)) dfaz = dfaz2;
    const double cdfaz = cos(deg_to_rad(dfaz));
    double tstor1 = ethm2 - ephm2;
    double tstor2 = 2. * ephm * ethm * cdfaz;
    tilta = .5 * atan2(tstor2, tstor1);
    const double stilta = sin(tilta);
    tstor1 = tstor1 * stilta * stilta;
    tstor2 = tstor2 * stilta * cos(tilta);
    emajr2 = -tstor1 + tstor2 + ethm2;
    eminr2 = tstor1 - tstor2 + ephm2;
    if (eminr2 < 0.) eminr2 = 0.;
    axrat = sqrt(eminr2 / emajr2);
    tilta = rad_to_deg(tilta);
    if (axrat <= 1.0e-5) sense = 0;
    else if (dfaz <= 0.) sense = 1;
    else sense = 2;
  }
  const double gcon = sp.gcon;
  const double gnmj = db10(gcon * emajr2);
  const double gnmn = db10(gcon * eminr2);
  const double gnv = db10(gcon * ethm2);
  const double gnh = db10(gcon * ephm2);
  const double gtot = db10(gcon * (ethm2 + ephm2));
  double tg;
  switch (sp.normalization) {
    case 1: tg = gnmj; break;
    case 2: tg = gnmn; break;
    case 3: tg = gnv; break;
    case 4: tg = gnh; break;
    default: tg = gtot; break;
  }
  out.gain[d] = tg;
  if (sp.iavp != 0) {
    const double ts = sp.gcop * (ethm2 + ephm2);
    double tmp3 = tha - sp.half_dtheta_rad, tmp4 = tha + sp.half_dtheta_rad;
    if (kth == 0) tmp3 = tha;
    else if (kth == n_theta - 1) tmp4 = tha;
    double da = fabs(sp.dphi_rad * (cos(tmp3) - cos(tmp4)));
    if ((kph == 0) || (kph == n_phi - 1)) da *= .5;
    out.pint[d] = ts * da;
    if (sp.iavp == 2) return;
  }
  out.pg_vert[d] = (sp.output_format != 1) ? gnmj : gnv;
  out.pg_horiz[d] = (sp.output_format != 1) ? gnmn : gnh;
  ethm *= sp.wavelength;
  ephm *= sp.wavelength;
  if (sp.rfld >= 1.0e-20) {
    ethm = ethm * sp.exrm;
    etha = etha + sp.exra;
    ephm = ephm * sp.exrm;
    epha = epha + sp.exra;
  }
  {
    double s, c;
    sincos(deg_to_rad(etha), &s, &c);
    out.e_theta[2 * d] = ethm * c;
    out.e_theta[2 * d + 1] = ethm * s;
    sincos(deg_to_rad(epha), &s, &c);
    out.e_phi[2 * d] = ephm * c;
    out.e_phi[2 * d + 1] = ephm * s;
  }
  out.pg_tot[d] = gtot;
  {
    double a = axrat;
    if (sense == 1) a = -a;
    const double gain = from_db10(gtot);
    const double lf = (1 + 2 * a + a * a) / (2 * (1 + a * a));
    const double rf = (1 - 2 * a + a * a) / (2 * (1 + a * a));
    out.pg_rhcp[d] = db10(gain * rf);
    out.pg_lhcp[d] = db10(gain * lf);
  }
  out.axial[d] = axrat;
  out.tilt[d] = tilta;
  out.sense[d] = sense;
  out.avg_scale[d] = (sin(tha) * (n_theta - 1) / n_theta) + 1.0 / n_theta;
}

}  // namespace

int launch_pattern(const DevSegments& g, int n, const double* abc, const PatternSpec& spec, const double* trig,
                   double* part, const PatternOut& out, hipStream_t stream) {
  const int n_dir = spec.n_theta * spec.n_phi;
  if (n_dir <= 0) return 0;
  const int dir_blocks = (n_dir + 255) / 256;
  const int slices = pattern_slices(n, n_dir);
  const int per = (n + slices - 1) / slices;
  if (spec.ground)
    launch_k(K_PATTERN_PARTIAL, pattern_partial_kernel<true>, dim3(dir_blocks, slices), dim3(256), 0, stream, g, n,
             abc, spec.n_theta, n_dir, trig, per, part);
  else
    launch_k(K_PATTERN_PARTIAL, pattern_partial_kernel<false>, dim3(dir_blocks, slices), dim3(256), 0, stream, g, n,
             abc, spec.n_theta, n_dir, trig, per, part);
  launch_k(K_PATTERN_FINISH, pattern_finish_kernel, dim3(dir_blocks), dim3(256), 0, stream, n_dir, slices, part, trig, spec,
                     out);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int pattern_slices(int n, int n_dir) {
  // enough segment slices to give every CU a few waves
  int slices = 1;
  const int dir_blocks = (n_dir + 255) / 256;
  while (dir_blocks * slices < 1024 && slices < 64 && n / (slices * 2) >= 32) slices *= 2;
  const int per = (n + slices - 1) / slices;
  return (n + per - 1) / per;
}

}  // namespace necgpu
