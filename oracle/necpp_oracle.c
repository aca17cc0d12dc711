/*
 * necpp_oracle.c -- CPU restatement of nec2++'s free-space thin-wire MoM
 * path.  TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the
 * bench's cpu_baseline leg as the checker; the product never calls it.
 *
 * Every function names the reference routine it restates.  The arithmetic
 * keeps the reference's evaluation order, and complex operations are
 * spelled out the way g++ evaluates std::complex<double> on x86-64 without
 * FMA: a*b = (ar*br - ai*bi, ar*bi + ai*br); a/b (both complex) = libgcc's
 * Smith division; complex*real and complex/real act per component.
 * Build with -ffp-contract=off.
 */
#include "necpp_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Complex helpers (std::complex<double> semantics under g++)          */
/* ------------------------------------------------------------------ */
typedef struct { double re, im; } cx;

static inline cx mk(double r, double i) { cx z; z.re = r; z.im = i; return z; }
static inline cx cadd(cx a, cx b) { return mk(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return mk(a.re - b.re, a.im - b.im); }
static inline cx cneg(cx a) { return mk(-a.re, -a.im); }
static inline cx cmul(cx a, cx b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx cmulr(cx a, double s) { return mk(a.re * s, a.im * s); }
static inline cx cdivr(cx a, double s) { return mk(a.re / s, a.im / s); }
/* libgcc __divdc3 (Smith's method), the path g++ takes for complex/complex. */
static inline cx cdiv(cx a, cx b) {
  double ratio, denom;
  cx r;
  if (fabs(b.re) < fabs(b.im)) {
    ratio = b.re / b.im;
    denom = (b.re * ratio) + b.im;
    r.re = ((a.re * ratio) + a.im) / denom;
    r.im = ((a.im * ratio) - a.re) / denom;
  } else {
    ratio = b.im / b.re;
    denom = (b.im * ratio) + b.re;
    r.re = ((a.im * ratio) + a.re) / denom;
    r.im = (a.im - (a.re * ratio)) / denom;
  }
  return r;
}

/* ------------------------------------------------------------------ */
/* Constants (src/math_util.h:72-107, src/electromag.cpp:22-47,         */
/* src/common.h:57, src/c_geometry.h:26)                               */
/* ------------------------------------------------------------------ */
static const double PI_T = 3.1415926536;
#define TWO_PI (2.0 * PI_T)
#define FOUR_PI (4.0 * PI_T)
#define SMIN 1.e-3
#define PCHCON 100000

static double eps0(void) { return 8.854e-12; }
static double mu0(void) { return FOUR_PI * 1.0e-7; }
double ora_speed_of_light(void) { return 1.0 / sqrt(mu0() * eps0()); }
double ora_impedance(void) { return sqrt(mu0() / eps0()); }
double ora_wavelength(double freq_mhz) { return ora_speed_of_light() / (1.0e6 * freq_mhz); }
static double deg_to_rad(double d) { return d * ((2 * PI_T) / 360.0); }
static double rad_to_deg(double r) { return r * (360.0 / (2 * PI_T)); }
static double db10(double x) { return (x < 1.0e-20) ? -999.99 : 10.0 * log10(x); }
static double from_db10(double x) { return (x < -99.9) ? 0.0 : pow(10, x / 10.0); }

/* ------------------------------------------------------------------ */
/* Errors                                                              */
/* ------------------------------------------------------------------ */
static char g_err[512];
static void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
const char* ora_last_error(void) { return g_err; }

/* ------------------------------------------------------------------ */
/* Geometry state                                                      */
/* ------------------------------------------------------------------ */
struct ora_geom {
  int64_t n, cap;
  double *x, *y, *z, *x2, *y2, *z2, *rad;
  int32_t* tag;
  int32_t *icon1, *icon2;
  double *len, *cab, *sab, *salp;   /* after geometry_complete */
  double *xu, *yu, *zu, *su, *bu;   /* unscaled copies */
  double *xs, *ys, *zs, *ss, *bs;   /* frequency-scaled */
  int complete;
  /* ground (nec_ground): 0 free space, 1 perfect (GN 1), 2 finite ground by
   * reflection coefficients (GN 0); zrati and frati are set by
   * ora_frequency_scale */
  int gtype;
  double epsr, sig;
  double zr_re, zr_im, fr_re, fr_im;
  /* loads (LD cards) and, per frequency, lambda * impedance per unit length
   * of every segment (2 doubles each; set by ora_frequency_scale) */
  int nload, load_err;
  int32_t *ldtyp, *ldtag, *ldtagf, *ldtagt;
  double *zlr, *zli, *zlc;
  double* zarr;
  double wl;
  /* networks and transmission lines (NT / TL cards): type 1 network, 2
   * straight line, 3 crossed line; the two port segments (1-based); x: y11
   * y12 y22 (re, im) for a network, Z0, length (m; <= 0: the distance of the
   * port segments' centres, set at the first run), then the shunt
   * admittances of ends 1 and 2 for a line */
  int nnet;
  int32_t* ntyp;
  int64_t *nseg1, *nseg2;
  double* nx;
};

static void* xrealloc(void* p, size_t n) {
  void* q = realloc(p, n ? n : 1);
  if (!q) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return q;
}

ora_geom* ora_geom_new(void) { return (ora_geom*)calloc(1, sizeof(ora_geom)); }

void ora_geom_free(ora_geom* g) {
  if (!g) return;
  double** d[] = {&g->x, &g->y, &g->z, &g->x2, &g->y2, &g->z2, &g->rad, &g->len, &g->cab, &g->sab, &g->salp,
                  &g->xu, &g->yu, &g->zu, &g->su, &g->bu, &g->xs, &g->ys, &g->zs, &g->ss, &g->bs};
  for (size_t i = 0; i < sizeof d / sizeof d[0]; i++) free(*d[i]);
  free(g->tag); free(g->icon1); free(g->icon2);
  free(g->ldtyp); free(g->ldtag); free(g->ldtagf); free(g->ldtagt);
  free(g->zlr); free(g->zli); free(g->zlc); free(g->zarr);
  free(g->ntyp); free(g->nseg1); free(g->nseg2); free(g->nx);
  free(g);
}

int64_t ora_nseg(const ora_geom* g) { return g->n; }

static void grow(ora_geom* g, int64_t n) {
  if (n <= g->cap) return;
  int64_t c = g->cap ? g->cap : 16;
  while (c < n) c *= 2;
  size_t b = (size_t)c * sizeof(double);
  g->x = xrealloc(g->x, b); g->y = xrealloc(g->y, b); g->z = xrealloc(g->z, b);
  g->x2 = xrealloc(g->x2, b); g->y2 = xrealloc(g->y2, b); g->z2 = xrealloc(g->z2, b);
  g->rad = xrealloc(g->rad, b);
  g->tag = xrealloc(g->tag, (size_t)c * sizeof(int32_t));
  g->cap = c;
}

/* c_geometry::wire, src/c_geometry.cpp:614-683. */
int ora_wire(ora_geom* g, int tag, int ns, double xw1, double yw1, double zw1,
             double xw2, double yw2, double zw2, double rad, double rdel, double rrad) {
  if (ns < 1) return 0;
  int64_t istart = g->n;
  grow(g, g->n + ns);
  g->n += ns;
  g->complete = 0;
  double dx = xw2 - xw1, dy = yw2 - yw1, dz = zw2 - zw1, delz, rd;
  if (fabs(rdel - 1.0) >= 1.0e-6) {
    delz = sqrt((dx * dx + dy * dy) + dz * dz);
    dx /= delz; dy /= delz; dz /= delz;
    delz = delz * (1. - rdel) / (1. - pow(rdel, ns));
    rd = rdel;
  } else {
    double fns = ns;
    dx /= fns; dy /= fns; dz /= fns;
    delz = 1.0;
    rd = 1.0;
  }
  double radz = rad, xs1 = xw1, ys1 = yw1, zs1 = zw1;
  for (int64_t i = istart; i < g->n; i++) {
    g->tag[i] = tag;
    double xs2 = xs1 + dx * delz, ys2 = ys1 + dy * delz, zs2 = zs1 + dz * delz;
    g->x[i] = xs1; g->y[i] = ys1; g->z[i] = zs1;
    g->x2[i] = xs2; g->y2[i] = ys2; g->z2[i] = zs2;
    g->rad[i] = radz;
    delz = delz * rd;
    radz = radz * rrad;
    xs1 = xs2; ys1 = ys2; zs1 = zs2;
  }
  g->x2[g->n - 1] = xw2; g->y2[g->n - 1] = yw2; g->z2[g->n - 1] = zw2;
  return 0;
}

/* c_geometry::move, src/c_geometry.cpp:830-900 (wires only): rotate the
 * segments from the first one tagged its (all for its = 0) about x, y and z
 * by rox, roy, roz radians, then shift them by (xs, ys, zs); with nrpt > 0
 * append nrpt such copies, each of the one before, with non-zero tags
 * raised by itgi. */
int ora_move(ora_geom* g, double rox, double roy, double roz, double xs, double ys, double zs, int its, int nrpt,
             int itgi) {
  double sps = sin(rox), cps = cos(rox), sth = sin(roy), cth = cos(roy), sph = sin(roz), cph = cos(roz);
  double xx = cph * cth, xy = cph * sth * sps - sph * cps, xz = cph * sth * cps + sph * sps;
  double yx = sph * cth, yy = sph * sth * sps + cph * cps, yz = sph * sth * cps - cph * sps;
  double zx = -sth, zy = cth * sps, zz = cth * cps;
  int64_t nrp = nrpt == 0 ? 1 : nrpt;
  if (g->n <= 0) return 0;
  int64_t i1 = ora_segment_number(g, its, 1);
  if (i1 < 0) return 1;
  if (i1 < 1) i1 = 1;
  int64_t k = (nrpt == 0) ? i1 - 1 : g->n;
  if (nrpt != 0) grow(g, g->n + (g->n + 1 - i1) * nrpt);
  g->complete = 0;
  for (int64_t ir = 0; ir < nrp; ir++) {
    for (int64_t i = i1 - 1; i < g->n; i++, k++) {
      double xi = g->x[i], yi = g->y[i], zi = g->z[i];
      g->x[k] = xi * xx + yi * xy + zi * xz + xs;
      g->y[k] = xi * yx + yi * yy + zi * yz + ys;
      g->z[k] = xi * zx + yi * zy + zi * zz + zs;
      xi = g->x2[i]; yi = g->y2[i]; zi = g->z2[i];
      g->x2[k] = xi * xx + yi * xy + zi * xz + xs;
      g->y2[k] = xi * yx + yi * yy + zi * yz + ys;
      g->z2[k] = xi * zx + yi * zy + zi * zz + zs;
      g->rad[k] = g->rad[i];
      g->tag[k] = g->tag[i] != 0 ? g->tag[i] + itgi : g->tag[i];
    }
    i1 = g->n + 1;
    g->n = k;
  }
  return 0;
}

/* c_geometry::scale, src/c_geometry.cpp:1308-1318 (GS): end points and
 * radii of every segment times f. */
void ora_scale(ora_geom* g, double f) {
  for (int64_t i = 0; i < g->n; i++) {
    g->x[i] *= f; g->y[i] *= f; g->z[i] *= f;
    g->x2[i] *= f; g->y2[i] *= f; g->z2[i] *= f;
    g->rad[i] *= f;
  }
  g->complete = 0;
}

static double l1(double ax, double ay, double az, double bx, double by, double bz) {
  return (fabs(ax - bx) + fabs(ay - by)) + fabs(az - bz);
}

/* build_connections, src/c_geometry.cpp:1705-1915 (wires only). */
static int build_connections(ora_geom* g, int ignd) {
  int64_t n = g->n;
  for (int64_t i = 0; i < n; i++) {
    int64_t iz = i + 1;
    double zi1 = g->z[i], zi2 = g->z2[i];
    double ex = g->x2[i] - g->x[i], ey = g->y2[i] - g->y[i], ez = g->z2[i] - g->z[i];
    double slen = sqrt((ex * ex + ey * ey) + ez * ez) * SMIN;
    int on_ground = 0;
    if (ignd > 0) {
      if (zi1 <= -slen) { set_err("GEOMETRY DATA ERROR--SEGMENT %lld EXTENDS BELOW GROUND", (long long)iz); return 1; }
      if (zi1 <= slen) { g->icon1[i] = (int32_t)iz; g->z[i] = 0.; on_ground = 1; }
    }
    if (!on_ground) {
      double vx = g->x[i], vy = g->y[i], vz = g->z[i];
      int64_t ic = i;
      int found = 0;
      for (int64_t j = 1; j < n; j++) {
        ic++;
        if (ic >= n) ic = 0;
        if (l1(vx, vy, vz, g->x[ic], g->y[ic], g->z[ic]) <= slen) { g->icon1[i] = (int32_t)(-(ic + 1)); found = 1; break; }
        if (l1(vx, vy, vz, g->x2[ic], g->y2[ic], g->z2[ic]) <= slen) { g->icon1[i] = (int32_t)(ic + 1); found = 1; break; }
      }
      if (!found) g->icon1[i] = 0;
    }
    if ((ignd > 0) || on_ground) {
      if (zi2 <= -slen) { set_err("GEOMETRY DATA ERROR--SEGMENT %lld EXTENDS BELOW GROUND", (long long)iz); return 1; }
      if (zi2 <= slen) {
        if (g->icon1[i] == iz) { set_err("GEOMETRY DATA ERROR--SEGMENT %lld LIES IN GROUND PLANE", (long long)iz); return 1; }
        g->icon2[i] = (int32_t)iz;
        g->z2[i] = 0.;
        continue;
      }
    }
    double vx = g->x2[i], vy = g->y2[i], vz = g->z2[i];
    int64_t ic = i;
    int found = 0;
    for (int64_t j = 1; j < n; j++) {
      ic++;
      if (ic >= n) ic = 0;
      if (l1(vx, vy, vz, g->x[ic], g->y[ic], g->z[ic]) <= slen) { g->icon2[i] = (int32_t)(ic + 1); found = 1; break; }
      if (l1(vx, vy, vz, g->x2[ic], g->y2[ic], g->z2[ic]) <= slen) { g->icon2[i] = (int32_t)(-(ic + 1)); found = 1; break; }
    }
    if (!found) g->icon2[i] = 0;
  }
  return 0;
}

/* resolve_junctions, src/c_geometry.cpp:1918-2074: average the coordinates
 * of all segment ends meeting at a node, in connection-chain order. */
static int resolve_junctions(ora_geom* g) {
  int64_t n = g->n;
  int64_t jcap = 8;
  int64_t* jco = (int64_t*)malloc((size_t)jcap * sizeof(int64_t));
  for (int64_t j = 0; j < n; j++) {
    int64_t jx = j + 1;
    int iend = -1, jend = -1;
    int64_t ix = g->icon1[j];
    int64_t ic = 1;
    jco[0] = -jx;
    double xa = g->x[j], ya = g->y[j], za = g->z[j];
    for (;;) {
      if ((ix != 0) && (ix != jx) && (ix <= n)) {
        int jump = 0;
        do {
          if (ix < 0) ix = -ix; else jend = -jend;
          jump = 0;
          if (ix == jx) break;
          if (ix < jx) { jump = 1; break; }
          ic++;
          if (ic > jcap) { jcap *= 2; jco = (int64_t*)xrealloc(jco, (size_t)jcap * sizeof(int64_t)); }
          jco[ic - 1] = ix * jend;
          int64_t ixx = ix - 1;
          if (ixx < 0 || ixx >= n) { free(jco); set_err("CONNECT - SEGMENT CONNECTION ERROR (bad index): %lld", (long long)ix); return 1; }
          if (jend != 1) {
            xa = xa + g->x[ixx]; ya = ya + g->y[ixx]; za = za + g->z[ixx];
            ix = g->icon1[ixx];
          } else {
            xa = xa + g->x2[ixx]; ya = ya + g->y2[ixx]; za = za + g->z2[ixx];
            ix = g->icon2[ixx];
          }
        } while (ix != 0);
        if (jump && (iend == 1)) break;
        if (jump) {
          iend = 1; jend = 1; ix = g->icon2[j]; ic = 1; jco[0] = jx;
          xa = g->x2[j]; ya = g->y2[j]; za = g->z2[j];
          continue;
        }
        double sep = (double)ic;
        xa = xa / sep; ya = ya / sep; za = za / sep;
        for (int64_t i = 0; i < ic; i++) {
          int64_t e = jco[i];
          if (e <= 0) {
            int64_t ixx = -e - 1;
            g->x[ixx] = xa; g->y[ixx] = ya; g->z[ixx] = za;
          } else {
            int64_t ixx = e - 1;
            g->x2[ixx] = xa; g->y2[ixx] = ya; g->z2[ixx] = za;
          }
        }
      }
      if (iend == 1) break;
      iend = 1; jend = 1; ix = g->icon2[j]; ic = 1; jco[0] = jx;
      xa = g->x2[j]; ya = g->y2[j]; za = g->z2[j];
    }
  }
  free(jco);
  return 0;
}

/* c_geometry::geometry_complete, src/c_geometry.cpp:454-604 (wires). */
int ora_geometry_complete(ora_geom* g, int gpflag) {
  int64_t n = g->n;
  if (n == 0) { set_err("Geometry has no wires or patches."); return 1; }
  if (n <= 1) { set_err("GEOMETRY HAS ONE OR FEWER SEGMENTS and no patches."); return 1; }
  size_t b = (size_t)n * sizeof(double), bi = (size_t)n * sizeof(int32_t);
  g->icon1 = xrealloc(g->icon1, bi); g->icon2 = xrealloc(g->icon2, bi);
  if (build_connections(g, gpflag)) return 1;
  if (resolve_junctions(g)) return 1;
  g->len = xrealloc(g->len, b); g->cab = xrealloc(g->cab, b);
  g->sab = xrealloc(g->sab, b); g->salp = xrealloc(g->salp, b);
  for (int64_t i = 0; i < n; i++) {
    double xw1 = g->x2[i] - g->x[i], yw1 = g->y2[i] - g->y[i], zw1 = g->z2[i] - g->z[i];
    g->x[i] = (g->x[i] + g->x2[i]) * .5;
    g->y[i] = (g->y[i] + g->y2[i]) * .5;
    g->z[i] = (g->z[i] + g->z2[i]) * .5;
    double xw2 = xw1 * xw1 + yw1 * yw1 + zw1 * zw1;
    double yw2 = sqrt(xw2);
    yw2 = (xw2 / yw2 + yw2) * .5;
    g->len[i] = yw2;
    g->cab[i] = xw1 / yw2;
    g->sab[i] = yw1 / yw2;
    xw2 = zw1 / yw2;
    if (xw2 > 1.) xw2 = 1.;
    if (xw2 < -1.) xw2 = -1.;
    g->salp[i] = xw2;
    if ((g->len[i] <= 1.e-20) || (g->rad[i] <= 0.)) { set_err("SEGMENT DATA ERROR"); return 1; }
  }
  g->xu = xrealloc(g->xu, b); g->yu = xrealloc(g->yu, b); g->zu = xrealloc(g->zu, b);
  g->su = xrealloc(g->su, b); g->bu = xrealloc(g->bu, b);
  g->xs = xrealloc(g->xs, b); g->ys = xrealloc(g->ys, b); g->zs = xrealloc(g->zs, b);
  g->ss = xrealloc(g->ss, b); g->bs = xrealloc(g->bs, b);
  for (int64_t i = 0; i < n; i++) {
    g->xu[i] = g->x[i]; g->yu[i] = g->y[i]; g->zu[i] = g->z[i];
    g->su[i] = g->len[i]; g->bu[i] = g->rad[i];
    g->xs[i] = g->x[i]; g->ys[i] = g->y[i]; g->zs[i] = g->z[i];
    g->ss[i] = g->len[i]; g->bs[i] = g->rad[i];
  }
  g->complete = 1;
  return 0;
}

void ora_segments(const ora_geom* g, double* out, int32_t* icon1, int32_t* icon2, int32_t* tags) {
  for (int64_t i = 0; i < g->n; i++) {
    if (out) {
      double* o = out + 8 * i;
      o[0] = g->xu[i]; o[1] = g->yu[i]; o[2] = g->zu[i]; o[3] = g->su[i];
      o[4] = g->bu[i]; o[5] = g->cab[i]; o[6] = g->sab[i]; o[7] = g->salp[i];
    }
    if (icon1) icon1[i] = g->icon1[i];
    if (icon2) icon2[i] = g->icon2[i];
    if (tags) tags[i] = g->tag[i];
  }
}

/* c_geometry::get_segment_number, src/c_geometry.cpp:76-99. */
int64_t ora_segment_number(const ora_geom* g, int tag, int m) {
  if (m <= 0) { set_err("CHECK DATA, PARAMETER SPECIFYING SEGMENT POSITION IN A GROUP OF EQUAL TAGS MUST NOT BE ZERO"); return -1; }
  if (tag == 0) return m;
  int64_t count = 0;
  for (int64_t i = 0; i < g->n; i++)
    if (g->tag[i] == tag && ++count == m) return i + 1;
  set_err("NO SEGMENT HAS AN ITAG OF %d", tag);
  return -1;
}

/* nec_context::zint, src/nec_context.cpp:6321-6406: internal impedance
 * per unit length of a round wire of conductivity * wavelength sigl and
 * radius rolam (wavelengths): the Kelvin-function ratio by its polynomial
 * approximations for x <= 8, the asymptotic expansions with the theta / phi
 * series for x > 8, and the skin-effect limit for x > 110. */
static double complex zint(double sigl, double rolam) {
  const double complex cc1 = CMPLX(6.0e-7, 1.9e-6), cc2 = CMPLX(-3.4e-6, 5.1e-6), cc3 = CMPLX(-2.52e-5, 0.0);
  const double complex cc4 = CMPLX(-9.06e-5, -9.01e-5), cc5 = CMPLX(0., -9.765e-4), cc6 = CMPLX(.0110486, -.0110485);
  const double complex cc7 = CMPLX(0., -.3926991), cc8 = CMPLX(1.6e-6, -3.2e-6), cc9 = CMPLX(1.17e-5, -2.4e-6);
  const double complex cc10 = CMPLX(3.46e-5, 3.38e-5), cc11 = CMPLX(5.0e-7, 2.452e-4);
  const double complex cc12 = CMPLX(-1.3813e-3, 1.3811e-3), cc13 = CMPLX(-6.25001e-2, -1.0e-7);
  const double complex cc14 = CMPLX(.7071068, .7071068), cn = cc14;
  const double tpcmu = 2.368705e+3, cmotp = 60.00;
  const double complex j1 = CMPLX(0.0, 1.0);
  double x = sqrt(tpcmu * sigl) * rolam;
#define ORA_TH(d) ((((((cc1 * (d) + cc2) * (d) + cc3) * (d) + cc4) * (d) + cc5) * (d) + cc6) * (d) + cc7)
#define ORA_PH(d) ((((((cc8 * (d) + cc9) * (d) + cc10) * (d) + cc11) * (d) + cc12) * (d) + cc13) * (d) + cc14)
  if (x > 110.0) {
    double complex br1 = CMPLX(0.70710678, -0.70710678);
    return j1 * sqrt(cmotp / sigl) * br1 / rolam;
  }
  if (x > 8.0) {
    double complex fx = sqrt((PI_T / 2.0) / x) * cexp(-cn * x + ORA_TH(-8. / x));
    double complex gx = cexp(cn * x + ORA_TH(8. / x)) / sqrt(TWO_PI * x);
    double complex br2 = j1 * fx / PI_T;
    double complex br1 = gx + br2;
    br2 = gx * ORA_PH(8.0 / x) - br2 * ORA_PH(-8.0 / x);
    br1 = br1 / br2;
    return j1 * sqrt(cmotp / sigl) * br1 / rolam;
  }
#undef ORA_TH
#undef ORA_PH
  double x8 = x / 8.0;
  double y = x8 * x8;
  double s = y * y;
  double ber = ((((((-9.01e-6 * s + 1.22552e-3) * s - .08349609) * s + 2.6419140) * s - 32.363456) * s + 113.77778) * s - 64.) * s + 1.;
  double bei = ((((((1.1346e-4 * s - .01103667) * s + .52185615) * s - 10.567658) * s + 72.817777) * s - 113.77778) * s + 16.) * y;
  double complex br1 = CMPLX(ber, bei);
  ber = (((((((-3.94e-6 * s + 4.5957e-4) * s - 0.02609253) * s + 0.66047849) * s - 6.0681481) * s + 14.222222) * s - 4.0) * y) * x;
  bei = ((((((4.609e-5 * s - 3.79386e-3) * s + .14677204) * s - 2.3116751) * s + 11.377778) * s - 10.666667) * s + .5) * x;
  double complex br2 = CMPLX(ber, bei);
  br1 = br1 / br2;
  return j1 * sqrt(cmotp / sigl) * br1 / rolam;
}

/* nec_context::load, src/nec_context.cpp:1861-2055 (nop = 1): lambda *
 * impedance per unit length of every loaded segment at the current
 * scaling (lengths and radii in wavelengths), summed over LD cards. */
static void compute_loads(ora_geom* g) {
  g->load_err = 0;
  if (g->nload == 0) return;
  int64_t n = g->n;
  g->zarr = (double*)xrealloc(g->zarr, (size_t)(2 * n) * sizeof(double));
  double complex* za = (double complex*)g->zarr;
  for (int64_t i = 0; i < n; i++) za[i] = 0.0;
  const double complex tpcj = CMPLX(0.0, 1.883698955e+9);
  const double wl = g->wl;
  for (int k = 0; k < g->nload; k++) {
    if (g->ldtyp[k] > 5) { set_err("IMPROPER LOAD TYPE CHOSEN, REQUESTED TYPE IS %d", g->ldtyp[k]); g->load_err = 1; return; }
    int ldtags = g->ldtag[k], jump = g->ldtyp[k] + 1, ichk = 0;
    int64_t l1 = 1, l2 = n;
    if (ldtags == 0 && (g->ldtagf[k] != 0 || g->ldtagt[k] != 0)) { l1 = g->ldtagf[k]; l2 = g->ldtagt[k]; }
    for (int64_t i = l1 - 1; i < l2; i++) {
      if (ldtags != 0) {
        if (ldtags != g->tag[i]) continue;
        if (g->ldtagf[k] != 0) {
          ichk++;
          if (ichk < g->ldtagf[k] || ichk > g->ldtagt[k]) continue;
        } else {
          ichk = 1;
        }
      } else {
        ichk = 1;
      }
      double sl = g->ss[i];
      double zlr = g->zlr[k], zli = g->zli[k], zlc = g->zlc[k];
      double complex zt = 0.0;
      switch (jump) {
        case 1:
          zt = zlr / sl + tpcj * zli / (sl * wl);
          if (fabs(zlc) > 1.0e-20) zt += wl / (tpcj * sl * zlc);
          break;
        case 2:
          zt = tpcj * sl * zlc / wl;
          if (fabs(zli) > 1.0e-20) zt += sl * wl / (tpcj * zli);
          if (fabs(zlr) > 1.0e-20) zt += sl / zlr;
          zt = 1. / zt;
          break;
        case 3:
          zt = zlr * wl + tpcj * zli;
          if (fabs(zlc) > 1.0e-20) zt += 1. / (tpcj * sl * sl * zlc);
          break;
        case 4:
          zt = tpcj * sl * sl * zlc;
          if (fabs(zli) > 1.0e-20) zt += 1. / (tpcj * zli);
          if (fabs(zlr) > 1.0e-20) zt += 1. / (zlr * wl);
          zt = 1. / zt;
          break;
        case 5:
          zt = CMPLX(zlr, zli) / sl;
          break;
        case 6:
          zt = zint(zlr * wl, g->bs[i]);
          break;
      }
      za[i] += zt;
    }
    if (ichk == 0) {
      set_err("LOADING DATA CARD ERROR, NO SEGMENT HAS AN ITAG = %d", ldtags);
      g->load_err = 1;
      return;
    }
  }
}

/* Test hook: zint (out[2] = re, im). */
void ora_zint(double sigl, double rolam, double* out) {
  double complex z = zint(sigl, rolam);
  out[0] = creal(z);
  out[1] = cimag(z);
}

/* nec_context::ld_card, src/nec_context.cpp:424-474: count 0 clears. */
int ora_set_loads(ora_geom* g, int count, const int32_t* type, const int32_t* tag, const int32_t* tagf,
                  const int32_t* tagt, const double* zlr, const double* zli, const double* zlc) {
  g->nload = 0;
  if (count <= 0) return 0;
  for (int k = 0; k < count; k++) {
    int f = tagf[k], t = tagt[k] == 0 ? tagf[k] : tagt[k];
    if (t < f) { set_err("DATA FAULT ON LOADING CARD No: %d: ITAG STEP1: %d IS GREATER THAN ITAG STEP2: %d", k + 1, f, t); return 1; }
  }
  g->ldtyp = (int32_t*)xrealloc(g->ldtyp, (size_t)count * sizeof(int32_t));
  g->ldtag = (int32_t*)xrealloc(g->ldtag, (size_t)count * sizeof(int32_t));
  g->ldtagf = (int32_t*)xrealloc(g->ldtagf, (size_t)count * sizeof(int32_t));
  g->ldtagt = (int32_t*)xrealloc(g->ldtagt, (size_t)count * sizeof(int32_t));
  g->zlr = (double*)xrealloc(g->zlr, (size_t)count * sizeof(double));
  g->zli = (double*)xrealloc(g->zli, (size_t)count * sizeof(double));
  g->zlc = (double*)xrealloc(g->zlc, (size_t)count * sizeof(double));
  for (int k = 0; k < count; k++) {
    g->ldtyp[k] = type[k];
    g->ldtag[k] = tag[k];
    g->ldtagf[k] = tagf[k];
    g->ldtagt[k] = tagt[k] == 0 ? tagf[k] : tagt[k];
    g->zlr[k] = zlr[k];
    g->zli[k] = zli[k];
    g->zlc[k] = zlc[k];
  }
  g->nload = count;
  return 0;
}

/* c_geometry::frequency_scale, src/c_geometry.cpp:3389-3418. */
void ora_frequency_scale(ora_geom* g, double freq_mhz) {
  double fr = (1.0e6 * freq_mhz) / ora_speed_of_light();
  for (int64_t i = 0; i < g->n; i++) {
    g->xs[i] = g->xu[i] * fr; g->ys[i] = g->yu[i] * fr; g->zs[i] = g->zu[i] * fr;
    g->ss[i] = g->su[i] * fr; g->bs[i] = g->bu[i] * fr;
  }
  /* nec_ground::calculate_antenna_environment, src/nec_ground.cpp:176-197:
   * frati = 1; for the finite ground zrati = 1 / sqrt(epsc),
   * epsc = epsr - j sig lambda eta / (2 pi); a negative sig is replaced by
   * -sig / (eta lambda / (2 pi)) and kept. */
  g->wl = ora_speed_of_light() / (freq_mhz * 1.0e6);
  compute_loads(g);
  g->fr_re = 1.0; g->fr_im = 0.0;
  if (g->gtype == 2) {
    double wavelength = ora_speed_of_light() / (freq_mhz * 1.0e6);
    double eta2pi = ora_impedance() / TWO_PI;
    if (g->sig < 0.) g->sig = -g->sig / (eta2pi * wavelength);
    double complex epsc = CMPLX(g->epsr, -g->sig * wavelength * eta2pi);
    double complex zr = 1.0 / csqrt(epsc);
    g->zr_re = creal(zr); g->zr_im = cimag(zr);
  }
}

/* nec_ground::parse_gn, src/nec_ground.cpp:75-121, for iperf -1 (free
 * space), 0 (finite ground, reflection-coefficient approximation) and 1
 * (perfect ground), no radial wires.  0 = ok. */
int ora_set_ground(ora_geom* g, int iperf, double epsr, double sig) {
  if (iperf == -1) { g->gtype = 0; return 0; }
  if (iperf == 1) { g->gtype = 1; g->epsr = epsr; g->sig = sig; return 0; }
  if (iperf == 0) {
    if (epsr == 0.0) { set_err("GROUND DIELECTRIC CONSTANT IS ZERO"); return 1; }
    if (sig == 0.0) { set_err("GROUND CONDUCTIVITY IS ZERO"); return 1; }
    g->gtype = 2; g->epsr = epsr; g->sig = sig;
    return 0;
  }
  set_err("ground type %d is not restated", iperf);
  return 1;
}

/* ------------------------------------------------------------------ */
/* Basis functions: sbf, trio, tbf                                      */
/* ------------------------------------------------------------------ */
static double omc_of(double d, double sdh, double cdh, int use_cd, double cd) {
  double omc;
  if (d <= 0.015) {
    omc = 4. * d * d;
    omc = ((1.3888889e-3 * omc - 4.1666666667e-2) * omc + .5) * omc;
  } else {
    omc = use_cd ? (1. - cd) : (1. - cdh * cdh + sdh * sdh);
  }
  return omc;
}

static double log_weight(double b) { return 1. / (log(1. / (PI_T * b)) - .577215664); }

/* c_geometry::sbf, src/c_geometry.cpp:3102-3289: coefficient of the basis
 * function centred on segment i, evaluated on segment is (both 1-based). */
static int sbf(const ora_geom* g, int64_t i, int64_t is, double* aa, double* bb, double* cc) {
  const int32_t* icon1 = g->icon1;
  const int32_t* icon2 = g->icon2;
  const double* sl = g->ss;
  const double* sr = g->bs;
  int64_t jcox, jcoxx, ix, local = 0, njun1 = 0, njun2;
  int jend, iend, june = 0;
  double d, sig, pp = 0., sdh, cdh, sd, omc, aj, pm = 0, cd, ap, qp, qm, xxi;
  *aa = 0.; *bb = 0.; *cc = 0.;
  ix = i - 1;
  jcox = icon1[ix];
  if (jcox > PCHCON) jcox = i;
  jend = -1; iend = -1; sig = -1.;
  do {
    if (jcox != 0) {
      if (jcox < 0) jcox = -jcox;
      else { sig = -sig; jend = -jend; }
      jcoxx = jcox - 1;
      local++;
      d = PI_T * sl[jcoxx];
      sdh = sin(d);
      cdh = cos(d);
      sd = 2. * sdh * cdh;
      omc = omc_of(d, sdh, cdh, 0, 0.);
      aj = log_weight(sr[jcoxx]);
      pp -= omc / sd * aj;
      if (jcox == is) {
        *aa = aj / sd * sig;
        *bb = aj / (2. * cdh);
        *cc = -aj / (2. * sdh) * sig;
        june = iend;
      }
      if (jcox != i) {
        if (jend != 1) jcox = icon1[jcoxx]; else jcox = icon2[jcoxx];
        if (llabs(jcox) != i) {
          if (jcox == 0) { set_err("SBF - SEGMENT CONNECTION ERROR FOR SEGMENT %lld", (long long)i); return 1; }
          continue;
        }
      } else if (jcox == is) {
        *bb = -*bb;
      }
      if (iend == 1) break;
    }
    pm = -pp;
    pp = 0.;
    njun1 = local;
    jcox = icon2[ix];
    if (jcox > PCHCON) jcox = i;
    jend = 1; iend = 1; sig = -1.;
  } while (jcox != 0);
  njun2 = local - njun1;
  d = PI_T * sl[ix];
  sdh = sin(d);
  cdh = cos(d);
  sd = 2. * sdh * cdh;
  cd = cdh * cdh - sdh * sdh;
  omc = omc_of(d, sdh, cdh, 1, cd);
  ap = log_weight(sr[ix]);
  aj = ap;
  if (njun1 == 0) {
    if (njun2 == 0) {
      *aa = -1.;
      qp = PI_T * sr[ix];
      xxi = qp * qp;
      xxi = qp * (1. - .5 * xxi) / (1. - xxi);
      *cc = 1. / (cdh - xxi * sdh);
      return 0;
    }
    qp = PI_T * sr[ix];
    xxi = qp * qp;
    xxi = qp * (1. - .5 * xxi) / (1. - xxi);
    qp = -(omc + xxi * sd) / (sd * (ap + xxi * pp) + cd * (xxi * ap - pp));
    if (june == 1) {
      *aa = -*aa * qp;
      *bb = *bb * qp;
      *cc = -*cc * qp;
      if (i != is) return 0;
    }
    *aa -= 1.;
    d = cd - xxi * sd;
    *bb += (sdh + ap * qp * (cdh - xxi * sdh)) / d;
    *cc += (cdh + ap * qp * (sdh + xxi * cdh)) / d;
    return 0;
  }
  if (njun2 == 0) {
    qm = PI_T * sr[ix];
    xxi = qm * qm;
    xxi = qm * (1. - .5 * xxi) / (1. - xxi);
    qm = (omc + xxi * sd) / (sd * (aj - xxi * pm) + cd * (pm + xxi * aj));
    if (june == -1) {
      *aa = *aa * qm;
      *bb = *bb * qm;
      *cc = *cc * qm;
      if (i != is) return 0;
    }
    *aa -= 1.;
    d = cd - xxi * sd;
    *bb += (aj * qm * (cdh - xxi * sdh) - sdh) / d;
    *cc += (cdh - aj * qm * (sdh + xxi * cdh)) / d;
    return 0;
  }
  qp = sd * (pm * pp + aj * ap) + cd * (pm * ap - pp * aj);
  qm = (ap * omc - pp * sd) / qp;
  qp = -(aj * omc + pm * sd) / qp;
  if (june != 0) {
    if (june < 0) {
      *aa = *aa * qm; *bb = *bb * qm; *cc = *cc * qm;
    } else {
      *aa = -*aa * qp; *bb = *bb * qp; *cc = -*cc * qp;
    }
    if (i != is) return 0;
  }
  *aa -= 1.;
  *bb += (aj * qm + ap * qp) * sdh / sd;
  *cc += (aj * qm - ap * qp) * cdh / sd;
  return 0;
}

/* c_geometry::trio, src/c_geometry.cpp:2982-3087. */
int ora_trio(const ora_geom* g, int j, int32_t* jco, double* ax, double* bx, double* cx, int cap) {
  int jsno = 0, jend = 0, iend = 0;
  int64_t jx = j - 1;
  int64_t jcox = g->icon1[jx], jcoxx;
  if (jcox <= PCHCON) { jend = -1; iend = -1; }
  if ((jcox == 0) || (jcox > PCHCON)) {
    jcox = g->icon2[jx];
    if (jcox <= PCHCON) { jend = 1; iend = 1; }
    if (jcox == 0 || (jcox > PCHCON)) {
      if (cap < 1) return -1;
      if (sbf(g, j, j, &ax[0], &bx[0], &cx[0])) return -1;
      jco[0] = j;
      return 1;
    }
  }
  do {
    if (jcox < 0) jcox = -jcox; else jend = -jend;
    jcoxx = jcox - 1;
    if (jcox != j) {
      if (jsno >= cap) { set_err("trio capacity"); return -1; }
      if (sbf(g, jcox, j, &ax[jsno], &bx[jsno], &cx[jsno])) return -1;
      jco[jsno] = (int32_t)jcox;
      jsno++;
      if (jend != 1) jcox = g->icon1[jcoxx]; else jcox = g->icon2[jcoxx];
      if (jcox == 0) { set_err("TRIO - SEGMENT CONNENTION ERROR FOR SEGMENT %d", j); return -1; }
      continue;
    }
    if (iend == 1) break;
    jcox = g->icon2[jx];
    if (jcox > PCHCON) break;
    jend = 1; iend = 1;
  } while (jcox != 0);
  if (jsno >= cap) { set_err("trio capacity"); return -1; }
  if (sbf(g, j, j, &ax[jsno], &bx[jsno], &cx[jsno])) return -1;
  jco[jsno] = j;
  return jsno + 1;
}

/* c_geometry::tbf, src/c_geometry.cpp:2763-2977: all coefficients of the
 * basis function centred on segment i (1-based). */
int ora_tbf(const ora_geom* g, int i, int icap, int32_t* jco, double* ax, double* bx, double* cx, int cap) {
  const int32_t* icon1 = g->icon1;
  const int32_t* icon2 = g->icon2;
  const double* sl = g->ss;
  const double* sr = g->bs;
  int64_t jcoxx, njun1 = 0, njun2, jsnop, jsnox;
  double sdh, cdh, sd, omc, aj, pm = 0, cd, ap, qp, qm, xxi, d;
  int64_t jsno = 0;
  double pp = 0.0;
  int64_t ix = i - 1;
  int64_t jcox = icon1[ix];
  if (jcox > PCHCON) jcox = i;
  int jend = -1, iend = -1;
  double sig = -1.0;
  do {
    if (jcox != 0) {
      if (jcox < 0) jcox = -jcox;
      else { sig = -sig; jend = -jend; }
      jcoxx = jcox - 1;
      jsno++;
      jsnox = jsno - 1;
      if (jsnox >= cap) { set_err("tbf capacity"); return -1; }
      jco[jsnox] = (int32_t)jcox;
      d = PI_T * sl[jcoxx];
      sdh = sin(d);
      cdh = cos(d);
      sd = 2.0 * sdh * cdh;
      omc = omc_of(d, sdh, cdh, 0, 0.);
      aj = log_weight(sr[jcoxx]);
      pp = pp - omc / sd * aj;
      ax[jsnox] = aj / sd * sig;
      bx[jsnox] = aj / (2. * cdh);
      cx[jsnox] = -aj / (2. * sdh) * sig;
      if (jcox != i) {
        if (jend == 1) jcox = icon2[jcoxx]; else jcox = icon1[jcoxx];
        if (llabs(jcox) != i) {
          if (jcox != 0) continue;
          set_err("TBF - SEGMENT CONNECTION ERROR FOR SEGMENT %d", i);
          return -1;
        }
      } else {
        bx[jsnox] = -bx[jsnox];
      }
      if (iend == 1) break;
    }
    pm = -pp;
    pp = 0.;
    njun1 = jsno;
    jcox = icon2[ix];
    if (jcox > PCHCON) jcox = i;
    jend = 1; iend = 1; sig = -1.;
  } while (jcox != 0);
  njun2 = jsno - njun1;
  jsnop = jsno;
  if (jsnop >= cap) { set_err("tbf capacity"); return -1; }
  jco[jsnop] = i;
  d = PI_T * sl[ix];
  sdh = sin(d);
  cdh = cos(d);
  sd = 2.0 * sdh * cdh;
  cd = cdh * cdh - sdh * sdh;
  omc = omc_of(d, sdh, cdh, 1, cd);
  ap = log_weight(sr[ix]);
  aj = ap;
  if (njun1 == 0) {
    if (njun2 == 0) {
      bx[jsnop] = 0.;
      if (icap == 0) xxi = 0.;
      else { qp = PI_T * sr[ix]; xxi = qp * qp; xxi = qp * (1. - .5 * xxi) / (1. - xxi); }
      cx[jsnop] = 1. / (cdh - xxi * sdh);
      ax[jsnop] = -1.;
      return (int)(jsnop + 1);
    }
    if (icap == 0) xxi = 0.;
    else { qp = PI_T * sr[ix]; xxi = qp * qp; xxi = qp * (1. - .5 * xxi) / (1. - xxi); }
    qp = -(omc + xxi * sd) / (sd * (ap + xxi * pp) + cd * (xxi * ap - pp));
    d = cd - xxi * sd;
    bx[jsnop] = (sdh + ap * qp * (cdh - xxi * sdh)) / d;
    cx[jsnop] = (cdh + ap * qp * (sdh + xxi * cdh)) / d;
    for (int64_t e = 0; e < njun2; e++) {
      ax[e] = -ax[e] * qp;
      bx[e] = bx[e] * qp;
      cx[e] = -cx[e] * qp;
    }
    ax[jsnop] = -1.;
    return (int)(jsnop + 1);
  }
  if (njun2 == 0) {
    if (icap == 0) xxi = 0.;
    else { qm = PI_T * sr[ix]; xxi = qm * qm; xxi = qm * (1. - .5 * xxi) / (1. - xxi); }
    qm = (omc + xxi * sd) / (sd * (aj - xxi * pm) + cd * (pm + xxi * aj));
    d = cd - xxi * sd;
    bx[jsnop] = (aj * qm * (cdh - xxi * sdh) - sdh) / d;
    cx[jsnop] = (cdh - aj * qm * (sdh + xxi * cdh)) / d;
    for (int64_t e = 0; e < njun1; e++) {
      ax[e] = ax[e] * qm;
      bx[e] = bx[e] * qm;
      cx[e] = cx[e] * qm;
    }
    ax[jsnop] = -1.;
    return (int)(jsnop + 1);
  }
  qp = sd * (pm * pp + aj * ap) + cd * (pm * ap - pp * aj);
  qm = (ap * omc - pp * sd) / qp;
  qp = -(aj * omc + pm * sd) / qp;
  bx[jsnop] = (aj * qm + ap * qp) * sdh / sd;
  cx[jsnop] = (aj * qm - ap * qp) * cdh / sd;
  for (int64_t e = 0; e < njun1; e++) {
    ax[e] = ax[e] * qm;
    bx[e] = bx[e] * qm;
    cx[e] = cx[e] * qm;
  }
  for (int64_t e = njun1; e < jsno; e++) {
    ax[e] = -ax[e] * qp;
    bx[e] = bx[e] * qp;
    cx[e] = -cx[e] * qp;
  }
  ax[jsnop] = -1.;
  return (int)(jsnop + 1);
}

/* ------------------------------------------------------------------ */
/* Decision tape (finite differences with frozen decisions)            */
/* ------------------------------------------------------------------ */
/* The fill takes discrete decisions that depend continuously on the
 * geometry: the near/far branch of efld_compute (r >= rkh), every accept /
 * refine choice of intx's adaptive Romberg steps, and gf's small-argument
 * branch of the self term.  Recording them at a point p and replaying them
 * at p +- h gives central differences of the function with those decisions
 * held fixed: the derivative that an adjoint of the same code computes.
 * One tape per thread; a taped fill must run on one thread (its decisions
 * are recorded and replayed in call order). */
static _Thread_local int tape_mode;           /* 0 off, 1 record, 2 replay */
static _Thread_local unsigned char* tape_buf;
static _Thread_local int64_t tape_len, tape_cap, tape_pos;
static _Thread_local int tape_bad;            /* replay read past the tape */

static int tape_decide(int actual) {
  if (tape_mode == 1) {
    if (tape_len == tape_cap) {
      tape_cap = tape_cap ? 2 * tape_cap : 1 << 20;
      tape_buf = (unsigned char*)xrealloc(tape_buf, (size_t)tape_cap);
    }
    tape_buf[tape_len++] = (unsigned char)(actual != 0);
    return actual;
  }
  if (tape_mode == 2) {
    if (tape_pos >= tape_len) { tape_bad = 1; return actual; }
    return tape_buf[tape_pos++];
  }
  return actual;
}

void ora_tape_record(void) { tape_mode = 1; tape_len = 0; tape_pos = 0; tape_bad = 0; }
void ora_tape_replay(void) { tape_mode = 2; tape_pos = 0; tape_bad = 0; }
void ora_tape_off(void) { tape_mode = 0; }
/* decisions recorded, decisions replayed so far, replay ran past the tape */
void ora_tape_stats(int64_t* out) { out[0] = tape_len; out[1] = tape_pos; out[2] = tape_bad; }

/* ------------------------------------------------------------------ */
/* Fill kernels                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
  /* source segment (nec_context members m_s, m_b, xj..salpj, ind1, ind2) */
  double s, b, xj, yj, zj, cabj, sabj, salpj;
  int ind1, ind2;
  double rkh;
  int use_ek;
  /* gf state (ija, zpk, rkb2) */
  int ija;
  double zpk, rkb2;
  ora_fill_counts* cnt;
} fill_state;

/* Operation tally of one basic block (the counters' comment in the header
 * gives the convention): f flops, of which d divisions and q square roots,
 * and sp specials. */
#define TALLY(c, f, d, q, sp) \
  do { (c)->flops += (f); (c)->divs += (d); (c)->sqrts += (q); (c)->specials += (sp); } while (0)

/* gf, src/nec_context.cpp:3553-3576. */
static void gf(fill_state* st, double zk, double* co, double* si) {
  static const double c0 = -1.38888889e-3, c1 = 4.16666667e-2;
  st->cnt->gf_calls++;
  double zdk = zk - st->zpk;
  double rk = sqrt(st->rkb2 + zdk * zdk);
  *si = sin(rk) / rk;
  TALLY(st->cnt, 5, 1, 1, 1);
  if (st->ija != 0) { *co = cos(rk) / rk; TALLY(st->cnt, 1, 1, 0, 1); return; }
  if (tape_decide(rk >= .2)) { *co = (cos(rk) - 1.0) / rk; TALLY(st->cnt, 2, 1, 0, 1); return; }
  double rks = rk * rk;
  *co = ((c0 * rks + c1) * rks - 0.5) * rk;
  TALLY(st->cnt, 6, 0, 0, 0);
}

/* test, src/c_evlcom.cpp:845-868 (cnt: the tally, or NULL for none). */
static void conv_test(ora_fill_counts* cnt, double f1r, double f2r, double* tr, double f1i, double f2i, double* ti,
                      double dmin) {
  double den = fabs(f2r), t = fabs(f2i);
  if (den < t) den = t;
  if (den < dmin) den = dmin;
  if (den < 1.0e-37) { *tr = 0.0; *ti = 0.0; return; }
  *tr = fabs((f1r - f2r) / den);
  *ti = fabs((f1i - f2i) / den);
  if (cnt) TALLY(cnt, 4, 2, 0, 0);
}

/* intx, src/nec_context.cpp:4135-4297: adaptive Romberg integration of the
 * gf integrand over the source segment. */
static void intx(fill_state* st, double el1, double el2, double b, int ij, double* sgr, double* sgi) {
  int ns, nt, nx = 1, nma = 65536, nts = 4, flag = 1;
  double z, s, ze, fnm, ep, zend, fns, dz = 0., zp, dzot = 0.;
  double t00r, g1r, g5r, t00i, g1i, g5i, t01r, g3r, t01i, g3i, t10r, t10i, te1i, te1r, t02r;
  double g2r, g4r, t02i, g2i, g4i, t11r, t11i, t20r, t20i, te2i, te2r;
  const double rx = 1.0e-4;
  st->cnt->intx_calls++;
  if (ij == 0) st->cnt->intx_self++;
  z = el1;
  ze = el2;
  if (ij == 0) ze = 0.;
  s = ze - z;
  fnm = nma;
  ep = s / (10. * fnm);
  zend = ze - ep;
  TALLY(st->cnt, 4, 1, 0, 0);
  *sgr = 0.;
  *sgi = 0.;
  ns = nx;
  nt = 0;
  gf(st, z, &g1r, &g1i);
  for (;;) {
    if (flag) {
      fns = ns;
      dz = s / fns;
      zp = z + dz;
      TALLY(st->cnt, 2, 1, 0, 0);
      if (zp > ze) {
        dz = ze - z;
        TALLY(st->cnt, 1, 0, 0, 0);
        if (fabs(dz) <= ep) goto done;
      }
      dzot = dz * .5;
      zp = z + dzot;
      gf(st, zp, &g3r, &g3i);
      zp = z + dz;
      gf(st, zp, &g5r, &g5i);
      TALLY(st->cnt, 3, 0, 0, 0);
    }
    t00r = (g1r + g5r) * dzot;
    t00i = (g1i + g5i) * dzot;
    t01r = (t00r + dz * g3r) * 0.5;
    t01i = (t00i + dz * g3i) * 0.5;
    t10r = (4.0 * t01r - t00r) / 3.0;
    t10i = (4.0 * t01i - t00i) / 3.0;
    TALLY(st->cnt, 16, 2, 0, 0);
    conv_test(st->cnt, t01r, t10r, &te1r, t01i, t10i, &te1i, 0.);
    st->cnt->intx_steps++;
    if (tape_decide((te1i <= rx) && (te1r <= rx))) {
      *sgr = *sgr + t10r;
      *sgi = *sgi + t10i;
      nt += 2;
      z += dz;
      TALLY(st->cnt, 3, 0, 0, 0);
      if (z >= zend) goto done;
      g1r = g5r;
      g1i = g5i;
      if (nt >= nts && ns > nx) { ns = ns / 2; nt = 1; }
      flag = 1;
      continue;
    }
    zp = z + dz * 0.25;
    gf(st, zp, &g2r, &g2i);
    zp = z + dz * 0.75;
    gf(st, zp, &g4r, &g4i);
    t02r = (t01r + dzot * (g2r + g4r)) * 0.5;
    t02i = (t01i + dzot * (g2i + g4i)) * 0.5;
    t11r = (4.0 * t02r - t01r) / 3.0;
    t11i = (4.0 * t02i - t01i) / 3.0;
    t20r = (16.0 * t11r - t10r) / 15.0;
    t20i = (16.0 * t11i - t10i) / 15.0;
    TALLY(st->cnt, 24, 4, 0, 0);
    conv_test(st->cnt, t11r, t20r, &te2r, t11i, t20i, &te2i, 0.);
    st->cnt->intx_refines++;
    const int bad5 = (te2i > rx) || (te2r > rx);
    if (tape_decide(bad5 && ns < nma)) {
      nt = 0;
      ns = ns * 2;
      fns = ns;
      dz = s / fns;
      dzot = dz * 0.5;
      TALLY(st->cnt, 2, 1, 0, 0);
      g5r = g3r; g5i = g3i;
      g3r = g2r; g3i = g2i;
      flag = 0;
      continue;
    }
    if (bad5) nt = 0; /* step size limited: accept the 5-point result */
    *sgr = *sgr + t20r;
    *sgi = *sgi + t20i;
    nt++;
    z += dz;
    TALLY(st->cnt, 3, 0, 0, 0);
    if (z >= zend) goto done;
    g1r = g5r;
    g1i = g5i;
    if (nt >= nts && ns > nx) { ns = ns / 2; nt = 1; }
    flag = 1;
  }
done:
  if (ij == 0) {
    *sgr = 2. * (*sgr + log((sqrt(b * b + s * s) + s) / b));
    *sgi = 2. * *sgi;
    TALLY(st->cnt, 9, 1, 1, 1);
  }
}

/* gx, src/nec_context.cpp:3605-3616. */
static void gx(fill_state* st, double zz, double rh, double xk, cx* gz, cx* gzp) {
  st->cnt->gx_calls++;
  double r2 = zz * zz + rh * rh;
  double r = sqrt(r2);
  double rkz = xk * r;
  cx temp = mk(cos(rkz), -sin(rkz));
  temp = cdivr(temp, r);
  *gz = temp;
  *gzp = cdivr(cmul(cneg(mk(1.0, rkz)), temp), r2);
  TALLY(st->cnt, 15, 4, 1, 2);
}

/* gxx, src/nec_context.cpp:3623-3673. */
static void gxx(fill_state* st, double zz, double rh, double a, double a2, double xk, int ira,
                cx* g1, cx* g1p, cx* g2, cx* g2p, cx* g3, cx* gzp) {
  st->cnt->gxx_calls++;
  double r2 = zz * zz + rh * rh;
  double r = sqrt(r2);
  double r4 = r2 * r2;
  double rk = xk * r;
  double rk2 = rk * rk;
  double rh2 = rh * rh;
  double t1 = .25 * a2 * rh2 / r4;
  double t2 = .5 * a2 / r2;
  cx c1 = mk(1.0, rk);
  cx c2 = cmulr(c1, 3.);
  c2.re = c2.re - rk2;
  cx c3 = csub(cmulr(mk(6.0, rk), rk2), cmulr(c1, 15.));
  cx gz = cdivr(mk(cos(rk), -sin(rk)), r);
  cx tc = cmulr(c2, t1);
  *g2 = cmul(gz, mk(1. + tc.re, tc.im));
  *g1 = csub(*g2, cmul(cmulr(c1, t2), gz));
  gz = cdivr(gz, r2);
  *g2p = cmul(gz, csub(cmulr(c3, t1), c1));
  *gzp = cmul(cmulr(c2, t2), gz);
  *g3 = cadd(*g2p, *gzp);
  *g1p = cmulr(*g3, zz);
  TALLY(st->cnt, 67, 6, 1, 2);
  if (ira != 1) {
    *g3 = cmulr(cadd(*g3, *gzp), rh);
    *gzp = cmul(cmulr(c1, -zz), gz);
    TALLY(st->cnt, 12, 0, 0, 0);
    if (rh <= 1.0e-10) { *g2 = mk(0., 0.); *g2p = mk(0., 0.); return; }
    *g2 = cdivr(*g2, rh);
    *g2p = cdivr(cmulr(*g2p, zz), rh);
    TALLY(st->cnt, 6, 4, 0, 0);
    return;
  }
  t2 = .5 * a;
  *g2 = cmul(cmulr(c1, -t2), gz);
  *g2p = cdivr(cmul(cmulr(gz, t2), c2), r2);
  *g3 = csub(cmulr(*g2p, rh2), cmul(cmulr(gz, a), c1));
  *g2p = cmulr(*g2p, zz);
  *gzp = cmul(cmulr(c1, -zz), gz);
  TALLY(st->cnt, 41, 2, 0, 0);
}

static const cx CONST1 = {0.0, 4.771341189};

/* eksc, src/nec_context.cpp:3051-3100: thin-wire fields of the sine, cosine
 * and constant filaments. */
static void eksc(fill_state* st, double s, double z, double rh, double xk, int ij,
                 cx* ezs, cx* ers, cx* ezc, cx* erc, cx* ezk, cx* erk) {
  st->ija = ij;
  st->zpk = xk * z;
  double rhk = xk * rh;
  st->rkb2 = rhk * rhk;
  double sh = .5 * s;
  double shk = xk * sh;
  double ss = sin(shk);
  double cs = cos(shk);
  double z2a = sh - z;
  double z1a = -(sh + z);
  TALLY(st->cnt, 7, 0, 0, 2);
  cx gz1, gz2, gp1, gp2;
  gx(st, z1a, rh, xk, &gz1, &gp1);
  gx(st, z2a, rh, xk, &gz2, &gp2);
  cx gzp1 = cmulr(gp1, z1a);
  cx gzp2 = cmulr(gp2, z2a);
  cx mc1 = cneg(CONST1);
  *ezs = cmul(CONST1, csub(cmulr(cmulr(csub(gz2, gz1), cs), xk), cmulr(cadd(gzp2, gzp1), ss)));
  *ezc = cmul(mc1, cadd(cmulr(cmulr(cadd(gz2, gz1), ss), xk), cmulr(csub(gzp2, gzp1), cs)));
  *erk = cmulr(cmul(CONST1, csub(gp2, gp1)), rh);
  TALLY(st->cnt, 50, 0, 0, 0);
  double cint, sint;
  intx(st, -shk, shk, rhk, ij, &cint, &sint);
  *ezk = cmul(mc1, cadd(csub(gzp2, gzp1), cmulr(mk(cint, -sint), xk * xk)));
  TALLY(st->cnt, 13, 0, 0, 0);
  if (rh >= 1.0e-10) {
    gzp1 = cmulr(gzp1, z1a);
    gzp2 = cmulr(gzp2, z2a);
    cx t1 = cmulr(cadd(cadd(cadd(gzp2, gzp1), gz2), gz1), ss);
    cx t2 = cmulr(cmulr(csub(cmulr(gz2, z2a), cmulr(gz1, z1a)), cs), xk);
    *ers = cdivr(cmul(mc1, csub(t1, t2)), rh);
    cx u1 = cmulr(csub(cadd(csub(gzp2, gzp1), gz2), gz1), cs);
    cx u2 = cmulr(cmulr(cadd(cmulr(gz2, z2a), cmulr(gz1, z1a)), ss), xk);
    *erc = cdivr(cmul(mc1, cadd(u1, u2)), rh);
    TALLY(st->cnt, 60, 4, 0, 0);
    return;
  }
  *ers = mk(0., 0.);
  *erc = mk(0., 0.);
}

/* ekscx, src/nec_context.cpp:3108-3188: extended thin-wire kernel. */
static void ekscx(fill_state* st, double bx, double s, double z, double rhx, double xk, int ij, int inx1, int inx2,
                  cx* ezs, cx* ers, cx* ezc, cx* erc, cx* ezk, cx* erk) {
  int ira;
  double b, rh;
  if (rhx >= bx) { rh = rhx; b = bx; ira = 0; }
  else { rh = bx; b = rhx; ira = 1; }
  double sh = .5 * s;
  st->ija = ij;
  st->zpk = xk * z;
  double rhk = xk * rh;
  st->rkb2 = rhk * rhk;
  double shk = xk * sh;
  double ss = sin(shk);
  double cs = cos(shk);
  double z2a = sh - z;
  double z1a = -(sh + z);
  double a2 = b * b;
  TALLY(st->cnt, 8, 0, 0, 2);
  cx gz1, gz2, gzp1, gzp2, gr1, gr2, grp1, grp2, grk1, grk2, gzz1, gzz2;
  if (inx1 != 2) gxx(st, z1a, rh, b, a2, xk, ira, &gz1, &gzp1, &gr1, &grp1, &grk1, &gzz1);
  else {
    gx(st, z1a, rhx, xk, &gz1, &grk1);
    gzp1 = cmulr(grk1, z1a);
    gr1 = cdivr(gz1, rhx);
    grp1 = cdivr(gzp1, rhx);
    grk1 = cmulr(grk1, rhx);
    gzz1 = mk(0., 0.);
    TALLY(st->cnt, 8, 4, 0, 0);
  }
  if (inx2 != 2) gxx(st, z2a, rh, b, a2, xk, ira, &gz2, &gzp2, &gr2, &grp2, &grk2, &gzz2);
  else {
    gx(st, z2a, rhx, xk, &gz2, &grk2);
    gzp2 = cmulr(grk2, z2a);
    gr2 = cdivr(gz2, rhx);
    grp2 = cdivr(gzp2, rhx);
    grk2 = cmulr(grk2, rhx);
    gzz2 = mk(0., 0.);
    TALLY(st->cnt, 8, 4, 0, 0);
  }
  cx mc1 = cneg(CONST1);
  *ezs = cmul(CONST1, csub(cmulr(cmulr(csub(gz2, gz1), cs), xk), cmulr(cadd(gzp2, gzp1), ss)));
  *ezc = cmul(mc1, cadd(cmulr(cmulr(cadd(gz2, gz1), ss), xk), cmulr(csub(gzp2, gzp1), cs)));
  {
    cx t1 = cmulr(cadd(cadd(cadd(cmulr(grp2, z2a), cmulr(grp1, z1a)), gr2), gr1), ss);
    cx t2 = cmulr(cmulr(csub(cmulr(gr2, z2a), cmulr(gr1, z1a)), cs), xk);
    *ers = cmul(mc1, csub(t1, t2));
    cx u1 = cmulr(csub(cadd(csub(cmulr(grp2, z2a), cmulr(grp1, z1a)), gr2), gr1), cs);
    cx u2 = cmulr(cmulr(cadd(cmulr(gr2, z2a), cmulr(gr1, z1a)), ss), xk);
    *erc = cmul(mc1, cadd(u1, u2));
  }
  *erk = cmul(CONST1, csub(grk2, grk1));
  TALLY(st->cnt, 104, 0, 0, 0);
  double cint, sint;
  intx(st, -shk, shk, rhk, ij, &cint, &sint);
  double bk = b * xk;
  double bk2 = bk * bk * .25;
  cx term = cmulr(mk(cint, -sint), (xk * xk) * (1. - bk2));
  *ezk = cmul(mc1, csub(cadd(csub(gzp2, gzp1), term), cmulr(csub(gzz2, gzz1), bk2)));
  TALLY(st->cnt, 24, 0, 0, 0);
}

/* efld_compute, src/nec_context.cpp:2777-2837 (direct field only). */
static double efld_compute(fill_state* st, double xij, double yij, double ai, double salpr, double zij, int ijx, cx egnd[9]) {
  cx tezs = mk(0, 0), ters = mk(0, 0), tezc, terc, tezk, terk;
  double cabj = st->cabj, sabj = st->sabj;
  double zp = xij * cabj + yij * sabj + zij * salpr;
  double rhox = xij - cabj * zp;
  double rhoy = yij - sabj * zp;
  double rhoz = zij - salpr * zp;
  double rh = sqrt(rhox * rhox + rhoy * rhoy + rhoz * rhoz + ai * ai);
  if (rh <= 1.e-10) { rhox = 0.0; rhoy = 0.0; rhoz = 0.0; }
  else { rhox /= rh; rhoy /= rh; rhoz /= rh; TALLY(st->cnt, 3, 3, 0, 0); }
  double r = sqrt(zp * zp + rh * rh);
  TALLY(st->cnt, 23, 0, 2, 0);
  if (tape_decide(r >= st->rkh)) {
    st->cnt->far_pairs++;
    double eta = ora_impedance();
    double rmag = TWO_PI * r;
    double cth = zp / r;
    double px = rh / r;
    cx e0 = mk(cos(rmag), -sin(rmag));
    double py = TWO_PI * r * r;
    cx e1 = cdivr(cmul(cmulr(e0, eta * cth), mk(1.0, -1.0 / rmag)), py);
    cx e2 = cdivr(cmul(cmulr(e0, eta * px), mk(1.0, rmag - 1.0 / rmag)), 2. * py);
    tezk = csub(cmulr(e1, cth), cmulr(e2, px));
    terk = cadd(cmulr(e1, px), cmulr(e2, cth));
    rmag = sin(PI_T * st->s) / PI_T;
    tezc = cmulr(tezk, rmag);
    terc = cmulr(terk, rmag);
    tezk = cmulr(tezk, st->s);
    terk = cmulr(terk, st->s);
    TALLY(st->cnt, 53, 9, 0, 3);
    egnd[3] = mk(0, 0); egnd[4] = mk(0, 0); egnd[5] = mk(0, 0);
  } else {
    st->cnt->near_pairs++;
    if (!st->use_ek)
      eksc(st, st->s, zp, rh, TWO_PI, ijx, &tezs, &ters, &tezc, &terc, &tezk, &terk);
    else
      ekscx(st, st->b, st->s, zp, rh, TWO_PI, ijx, st->ind1, st->ind2, &tezs, &ters, &tezc, &terc, &tezk, &terk);
    egnd[3] = cadd(cmulr(tezs, cabj), cmulr(ters, rhox));
    egnd[4] = cadd(cmulr(tezs, sabj), cmulr(ters, rhoy));
    egnd[5] = cadd(cmulr(tezs, salpr), cmulr(ters, rhoz));
    TALLY(st->cnt, 18, 0, 0, 0);
  }
  egnd[0] = cadd(cmulr(tezk, cabj), cmulr(terk, rhox));
  egnd[1] = cadd(cmulr(tezk, sabj), cmulr(terk, rhoy));
  egnd[2] = cadd(cmulr(tezk, salpr), cmulr(terk, rhoz));
  egnd[6] = cadd(cmulr(tezc, cabj), cmulr(terc, rhox));
  egnd[7] = cadd(cmulr(tezc, sabj), cmulr(terc, rhoy));
  egnd[8] = cadd(cmulr(tezc, salpr), cmulr(terc, rhoz));
  TALLY(st->cnt, 36, 0, 0, 0);
  return r;
}

/* test_ek_approximation, src/c_geometry.cpp:3449-3458. */
static int test_ek(const ora_geom* g, int64_t s1, int64_t s2) {
  double ratio = g->bs[s2] / g->bs[s1];
  double xi = fabs(g->cab[s1] * g->cab[s2] + g->sab[s1] * g->sab[s2] + g->salp[s1] * g->salp[s2]);
  return ((xi < 0.999999) || (fabs(ratio - 1.0) > 1.e-6)) ? 2 : 0;
}

/* End-condition indices of cmww for the extended kernel,
 * src/nec_context.cpp:2535-2601 (j 0-based). */
static void ek_ends(const ora_geom* g, int64_t j, int* ind1, int* ind2) {
  int64_t jx = j + 1;
  double cabj = g->cab[j], sabj = g->sab[j];
  int64_t ipr = g->icon1[j];
  int64_t iprx = llabs(ipr) - 1;
  *ind1 = 2;
  if (ipr > PCHCON) *ind1 = 0;
  else if (ipr == 0) *ind1 = 1;
  else if ((ipr == jx) && (cabj * cabj + sabj * sabj <= 1.e-8)) *ind1 = 0;
  else if ((ipr > 0) && (g->icon2[iprx] == jx)) *ind1 = test_ek(g, j, ipr - 1);
  else if ((ipr < 0) && (g->icon1[iprx] == -jx)) *ind1 = test_ek(g, j, -ipr - 1);
  ipr = g->icon2[j];
  if (ipr > PCHCON) *ind2 = 2;
  else if (ipr == 0) *ind2 = 1;
  else if (ipr < 0) {
    iprx = -ipr - 1;
    if (-g->icon2[iprx] != jx) *ind2 = 2; else *ind2 = test_ek(g, j, iprx);
  } else {
    if (ipr != jx) {
      iprx = ipr - 1;
      if (g->icon1[iprx] != jx) *ind2 = 2; else *ind2 = test_ek(g, j, iprx);
    } else if ((cabj * cabj + sabj * sabj) > 1.e-8) *ind2 = 2;
    else *ind2 = 0;
  }
}

/* Test hooks (see necpp_oracle.h). */
static void put_cx(double* out, int k, cx v) { out[2 * k] = v.re; out[2 * k + 1] = v.im; }

void ora_eksc(double s, double z, double rh, int ij, double* out) {
  ora_fill_counts cnt;
  fill_state st;
  memset(&st, 0, sizeof st);
  st.cnt = &cnt;
  cx ezs, ers, ezc, erc, ezk, erk;
  eksc(&st, s, z, rh, TWO_PI, ij, &ezs, &ers, &ezc, &erc, &ezk, &erk);
  put_cx(out, 0, ezs); put_cx(out, 1, ers); put_cx(out, 2, ezc);
  put_cx(out, 3, erc); put_cx(out, 4, ezk); put_cx(out, 5, erk);
}

void ora_ekscx(double b, double s, double z, double rh, int ij, int inx1, int inx2, double* out) {
  ora_fill_counts cnt;
  fill_state st;
  memset(&st, 0, sizeof st);
  st.cnt = &cnt;
  cx ezs, ers, ezc, erc, ezk, erk;
  ekscx(&st, b, s, z, rh, TWO_PI, ij, inx1, inx2, &ezs, &ers, &ezc, &erc, &ezk, &erk);
  put_cx(out, 0, ezs); put_cx(out, 1, ers); put_cx(out, 2, ezc);
  put_cx(out, 3, erc); put_cx(out, 4, ezk); put_cx(out, 5, erk);
}

void ora_intx(double zpk, double rkb2, int ija, double el1, double el2, double b, int ij, double* out) {
  ora_fill_counts cnt;
  fill_state st;
  memset(&st, 0, sizeof st);
  st.cnt = &cnt;
  st.zpk = zpk;
  st.rkb2 = rkb2;
  st.ija = ija;
  intx(&st, el1, el2, b, ij, &out[0], &out[1]);
}

double ora_efld(double xj, double yj, double zj, double cabj, double sabj, double salpj, double s, double b,
                double xi, double yi, double zi, double ai, int not_self, double rkh, int use_ek, int ind1, int ind2,
                double* out) {
  ora_fill_counts cnt;
  fill_state st;
  memset(&st, 0, sizeof st);
  st.cnt = &cnt;
  st.s = s; st.b = b; st.xj = xj; st.yj = yj; st.zj = zj;
  st.cabj = cabj; st.sabj = sabj; st.salpj = salpj;
  st.rkh = rkh; st.use_ek = use_ek; st.ind1 = ind1; st.ind2 = ind2;
  cx e[9];
  double r = efld_compute(&st, xi - xj, yi - yj, ai, salpj, zi - zj, not_self, e);
  for (int k = 0; k < 9; k++) put_cx(out, k, e[k]);
  return r;
}

static double complex to_c(cx a) { return CMPLX(a.re, a.im); }
static cx from_c(double complex a) { return mk(creal(a), cimag(a)); }
static void ground_image(const ora_geom* g, double xij, double yij, double zij, double r, cx t[9], cx e[9],
                         ora_fill_counts* cnt);

/* nefld, src/nec_context.cpp:4304-4450, for wires (no patches): the
 * electric field at npts points (xyz, in wavelengths) of the current
 * coefficients abc (planes air aii bir bii cir cii).  ax is the radius of
 * the first segment the point lies on (|zp| <= 0.5001 s, rho^2 <= 0.9 b^2),
 * else 0; every segment then contributes efld's nine components with
 * ij = true, plus over a ground the reflected field of efld, e = ek A +
 * es B + ec C per axis.  out: 6 doubles per point (ex, ey, ez; re, im). */
int ora_near_e(const ora_geom* g, const double* abc, int64_t npts, const double* xyz, double rkh, int use_ek,
               double* out) {
  const int64_t n = g->n;
  const double *air = abc, *aii = abc + n, *bir = abc + 2 * n, *bii = abc + 3 * n, *cir = abc + 4 * n,
               *cii = abc + 5 * n;
  for (int64_t p = 0; p < npts; ++p) {
    const double xob = xyz[3 * p], yob = xyz[3 * p + 1], zob = xyz[3 * p + 2];
    double ax = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double xj = xob - g->xs[i], yj = yob - g->ys[i], zj = zob - g->zs[i];
      double zp = g->cab[i] * xj + g->sab[i] * yj + g->salp[i] * zj;
      if (fabs(zp) > 0.5001 * g->ss[i]) continue;
      zp = xj * xj + yj * yj + zj * zj - zp * zp;
      const double b = g->bs[i];
      if (zp > 0.9 * b * b) continue;
      ax = b;
      break;
    }
    cx ex = mk(0.0, 0.0), ey = mk(0.0, 0.0), ez = mk(0.0, 0.0);
    for (int64_t i = 0; i < n; ++i) {
      int ind1 = 2, ind2 = 2;
      if (use_ek) ek_ends(g, i, &ind1, &ind2);
      ora_fill_counts cnt;
      fill_state st;
      memset(&st, 0, sizeof st);
      st.cnt = &cnt;
      st.s = g->ss[i]; st.b = g->bs[i]; st.xj = g->xs[i]; st.yj = g->ys[i]; st.zj = g->zs[i];
      st.cabj = g->cab[i]; st.sabj = g->sab[i]; st.salpj = g->salp[i];
      st.rkh = rkh; st.use_ek = use_ek; st.ind1 = ind1; st.ind2 = ind2;
      cx e[9];
      efld_compute(&st, xob - st.xj, yob - st.yj, ax, st.salpj, zob - st.zj, 1, e);
      if (g->gtype != 0) {
        /* the reflected field, as in the fill (efld, src/nec_context.cpp:2880-2956) */
        cx t[9];
        const double zij = zob + st.zj;
        const double r = efld_compute(&st, xob - st.xj, yob - st.yj, ax, -st.salpj, zij, 1, t);
        ground_image(g, xob - st.xj, yob - st.yj, zij, r, t, e, NULL);
      }
      const cx acx = mk(air[i], aii[i]), bcx = mk(bir[i], bii[i]), ccx = mk(cir[i], cii[i]);
      ex = cadd(ex, cadd(cadd(cmul(e[0], acx), cmul(e[3], bcx)), cmul(e[6], ccx)));
      ey = cadd(ey, cadd(cadd(cmul(e[1], acx), cmul(e[4], bcx)), cmul(e[7], ccx)));
      ez = cadd(ez, cadd(cadd(cmul(e[2], acx), cmul(e[5], bcx)), cmul(e[8], ccx)));
    }
    out[6 * p] = ex.re; out[6 * p + 1] = ex.im;
    out[6 * p + 2] = ey.re; out[6 * p + 3] = ey.im;
    out[6 * p + 4] = ez.re; out[6 * p + 5] = ez.im;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Near magnetic field (nhfld)                                         */
/* ------------------------------------------------------------------ */

/* gh, src/nec_context.cpp:3582-3596: the integrand of hfk at z, for the
 * point at (rho^2 = rhks, z = zpka), distances in radians (2 pi x
 * wavelengths): hr + j hi = (sin r / r^2 + cos r / r^3) + j (cos r / r^2 -
 * sin r / r^3). */
static void gh_integrand(double zk, double zpka, double rhks, double* hr, double* hi) {
  double dz = zk - zpka;
  const double rs = rhks + dz * dz;
  const double r = sqrt(rs);
  const double ckr = cos(r), skr = sin(r);
  const double rr2 = 1. / rs;
  const double rr3 = rr2 / r;
  *hr = skr * rr2 + ckr * rr3;
  *hi = ckr * rr2 - skr * rr3;
}

/* hfk, src/nec_context.cpp:3681-3822: the same adaptive Romberg scheme as
 * intx (rx = 1e-4, steps halved after nts accepted, doubled up to 65536) of
 * gh over [el1, el2], times rhk / 2. */
static void hfk(double el1, double el2, double rhk, double zpkx, double* sgr, double* sgi) {
  const int nx = 1, nma = 65536, nts = 4;
  const double rx = 1.0e-4;
  const double zpka = zpkx, rhks = rhk * rhk;
  double z = el1;
  const double ze = el2;
  const double s = ze - z;
  const double ep = s / (10. * nma);
  const double zend = ze - ep;
  double accr = 0.0, acci = 0.0;
  int ns = nx, nt = 0, fresh = 1;
  double dz = 0., dzot = 0., g1r, g1i, g2r, g2i, g3r = 0., g3i = 0., g4r, g4i, g5r = 0., g5i = 0.;
  gh_integrand(z, zpka, rhks, &g1r, &g1i);
  for (;;) {
    if (fresh) {
      dz = s / ns;
      if (z + dz > ze) {
        dz = ze - z;
        if (fabs(dz) <= ep) break;
      }
      dzot = dz * .5;
      gh_integrand(z + dzot, zpka, rhks, &g3r, &g3i);
      gh_integrand(z + dz, zpka, rhks, &g5r, &g5i);
    }
    const double t00r = (g1r + g5r) * dzot, t00i = (g1i + g5i) * dzot;
    const double t01r = (t00r + dz * g3r) * 0.5, t01i = (t00i + dz * g3i) * 0.5;
    const double t10r = (4.0 * t01r - t00r) / 3.0, t10i = (4.0 * t01i - t00i) / 3.0;
    double te1r, te1i;
    conv_test(NULL, t01r, t10r, &te1r, t01i, t10i, &te1i, 0.);
    double addr, addi;
    if ((te1i <= rx) && (te1r <= rx)) {
      addr = t10r;
      addi = t10i;
      nt += 2;
    } else {
      gh_integrand(z + dz * 0.25, zpka, rhks, &g2r, &g2i);
      gh_integrand(z + dz * 0.75, zpka, rhks, &g4r, &g4i);
      const double t02r = (t01r + dzot * (g2r + g4r)) * 0.5, t02i = (t01i + dzot * (g2i + g4i)) * 0.5;
      const double t11r = (4.0 * t02r - t01r) / 3.0, t11i = (4.0 * t02i - t01i) / 3.0;
      const double t20r = (16.0 * t11r - t10r) / 15.0, t20i = (16.0 * t11i - t10i) / 15.0;
      double te2r, te2i;
      conv_test(NULL, t11r, t20r, &te2r, t11i, t20i, &te2i, 0.);
      const int bad = (te2i > rx) || (te2r > rx);
      if (bad && ns < nma) {
        /* halve the step and try again from z with the points in hand */
        nt = 0;
        ns *= 2;
        dz = s / ns;
        dzot = dz * 0.5;
        g5r = g3r; g5i = g3i;
        g3r = g2r; g3i = g2i;
        fresh = 0;
        continue;
      }
      if (bad) nt = 0; /* step size limited: accept the 5-point result */
      addr = t20r;
      addi = t20i;
      nt++;
    }
    accr += addr;
    acci += addi;
    z += dz;
    if (z >= zend) break;
    g1r = g5r;
    g1i = g5i;
    if (nt >= nts && ns > nx) { ns /= 2; nt = 1; }
    fresh = 1;
  }
  *sgr = accr * rhk * .5;
  *sgi = acci * rhk * .5;
}

/* hsflx, src/nec_context.cpp:4059-4131: the azimuthal H of the constant
 * (hpk), sine (hps) and cosine (hpc) currents of a segment of length s at
 * radial distance rh and axial offset zpx from its centre (wavelengths). */
static void hsflx(double s, double rh, double zpx, cx* hpk, cx* hps, cx* hpc) {
  if (rh < 1.0e-10) {
    *hpk = mk(0, 0); *hps = mk(0, 0); *hpc = mk(0, 0);
    return;
  }
  const double hss = zpx >= 0.0 ? 1.0 : -1.0;
  const double zp = fabs(zpx);
  const double dh = 0.5 * s;
  const double z1 = zp + dh, z2a = zp - dh;
  const double rhz = z2a >= 1.0e-7 ? rh / z2a : 1.0;
  const double dk = TWO_PI * dh;
  const double cdk = cos(dk), sdk = sin(dk);
  double hkr, hki;
  hfk(-dk, dk, rh * TWO_PI, zp * TWO_PI, &hkr, &hki);
  *hpk = mk(hkr, hki);
  if (rhz >= 1.0e-3) {
    const double rh2 = rh * rh;
    const double r1 = sqrt(rh2 + z1 * z1), r2 = sqrt(rh2 + z2a * z2a);
    /* exp(-j 2 pi r) */
    const cx ekr1 = mk(cos(TWO_PI * r1), -sin(TWO_PI * r1));
    const cx ekr2 = mk(cos(TWO_PI * r2), -sin(TWO_PI * r2));
    const cx t1 = cdivr(cmulr(ekr1, z1), r1);
    const cx t2 = cdivr(cmulr(ekr2, z2a), r2);
    /* hps = (cdk (ekr2 - ekr1) - j sdk (t2 + t1)) hss; hpc = -sdk (ekr2 +
     * ekr1) - j cdk (t2 - t1); both times cons = -j / (4 pi rh) */
    const cx d = csub(ekr2, ekr1), sm = cadd(t2, t1);
    cx ps = mk(cdk * d.re + sdk * sm.im, cdk * d.im - sdk * sm.re);
    ps = cmulr(ps, hss);
    const cx a2 = cadd(ekr2, ekr1), df = csub(t2, t1);
    const cx pc = mk(-sdk * a2.re + cdk * df.im, -sdk * a2.im - cdk * df.re);
    const double cons = -1.0 / (2.0 * TWO_PI * rh); /* times j */
    *hps = mk(-cons * ps.im, cons * ps.re);
    *hpc = mk(-cons * pc.im, cons * pc.re);
    return;
  }
  /* close to the segment's axis beyond an end */
  const cx ekr1 = cdivr(mk(cdk, sdk), z2a * z2a);
  const cx ekr2 = cdivr(mk(cdk, -sdk), z1 * z1);
  const double t1 = TWO_PI * (1.0 / z1 - 1.0 / z2a);
  const cx t2 = cdivr(cmulr(mk(cos(TWO_PI * zp), -sin(TWO_PI * zp)), rh), PI_T * 8.0);
  const cx sps = cadd(mk(t1, 0.0), cmulr(cadd(ekr1, ekr2), sdk));
  *hps = cmulr(cmul(t2, sps), hss);
  const cx dpc = cmulr(csub(ekr1, ekr2), cdk);
  *hpc = cmul(t2, mk(dpc.re, dpc.im - t1));
}

/* hsfld, src/nec_context.cpp:3927-4055: the H field at (xi, yi, zi) of the
 * constant, sine and cosine currents of source segment st (h[0..2] k,
 * h[3..5] s, h[6..8] c; x, y, z each), ai the observation radius.  Over a
 * ground the image segment (z and the z cosine negated) is subtracted,
 * weighted for the finite ground by the reflection coefficients of the
 * specular direction.  An evaluation point on a segment's axis (rho <=
 * 1e-10) zeroes all nine components, the image's too. */
static void hsfld(const ora_geom* g, const fill_state* st, double xi, double yi, double zi, double ai, cx h[9]) {
  const double xij = xi - st->xj, yij = yi - st->yj;
  const int loops = g->gtype != 0 ? 2 : 1;
  for (int q = 0; q < 9; ++q) h[q] = mk(0, 0);
  for (int loop = 0; loop < loops; ++loop) {
    const double rfl = loop == 0 ? 1.0 : -1.0;
    const double salpr = st->salpj * rfl;
    const double zij = zi - rfl * st->zj;
    const double zp = xij * st->cabj + yij * st->sabj + zij * salpr;
    double rhox = xij - st->cabj * zp, rhoy = yij - st->sabj * zp, rhoz = zij - salpr * zp;
    const double rh = sqrt(rhox * rhox + rhoy * rhoy + rhoz * rhoz + ai * ai);
    if (rh <= 1.0e-10) {
      for (int q = 0; q < 9; ++q) h[q] = mk(0, 0);
      continue;
    }
    rhox /= rh; rhoy /= rh; rhoz /= rh;
    /* phi-hat = source direction x rho-hat */
    const double phx = st->sabj * rhoz - salpr * rhoy;
    const double phy = salpr * rhox - st->cabj * rhoz;
    const double phz = st->cabj * rhoy - st->sabj * rhox;
    cx hp[3];
    hsflx(st->s, rh, zp, &hp[0], &hp[1], &hp[2]);
    if (loop == 0) {
      for (int k = 0; k < 3; ++k) {
        h[3 * k] = cmulr(hp[k], phx);
        h[3 * k + 1] = cmulr(hp[k], phy);
        h[3 * k + 2] = cmulr(hp[k], phz);
      }
      continue;
    }
    if (g->gtype == 1) {
      for (int k = 0; k < 3; ++k) {
        h[3 * k] = csub(h[3 * k], cmulr(hp[k], phx));
        h[3 * k + 1] = csub(h[3 * k + 1], cmulr(hp[k], phy));
        h[3 * k + 2] = csub(h[3 * k + 2], cmulr(hp[k], phz));
      }
      continue;
    }
    /* finite ground: reflection coefficients (vertical rrv, horizontal rrh) */
    const double complex zratx = CMPLX(g->zr_re, g->zr_im);
    const double rmag = sqrt(zp * zp + rh * rh);
    const double xymag = sqrt(xij * xij + yij * yij);
    double px, py, cth;
    double complex rrv;
    if (xymag <= 1.0e-6) {
      px = 0.; py = 0.; cth = 1.; rrv = 1.0;
    } else {
      px = -yij / xymag;
      py = xij / xymag;
      cth = zij / rmag;
      rrv = csqrt(1.0 - zratx * zratx * (1.0 - cth * cth));
    }
    double complex rrh = zratx * cth;
    rrh = -(rrh - rrv) / (rrh + rrv);
    rrv = zratx * rrv;
    rrv = (cth - rrv) / (cth + rrv);
    double complex qy = (phx * px + phy * py) * (rrv - rrh);
    const double complex qx = qy * px + phx * rrh;
    qy = qy * py + phy * rrh;
    const double complex qz = phz * rrh;
    for (int k = 0; k < 3; ++k) {
      const double complex hk = CMPLX(hp[k].re, hp[k].im);
      h[3 * k] = from_c(to_c(h[3 * k]) - hk * qx);
      h[3 * k + 1] = from_c(to_c(h[3 * k + 1]) - hk * qy);
      h[3 * k + 2] = from_c(to_c(h[3 * k + 2]) - hk * qz);
    }
  }
}

/* nhfld, src/nec_context.cpp:5189-5262, for wires (no patches): the
 * magnetic field at npts points (xyz in wavelengths), the observation radius
 * found as nefld does, H = hk A + hs B + hc C per axis.  out: 6 doubles per
 * point (hx, hy, hz; re, im). */
int ora_near_h(const ora_geom* g, const double* abc, int64_t npts, const double* xyz, double* out) {
  const int64_t n = g->n;
  const double *air = abc, *aii = abc + n, *bir = abc + 2 * n, *bii = abc + 3 * n, *cir = abc + 4 * n,
               *cii = abc + 5 * n;
  for (int64_t p = 0; p < npts; ++p) {
    const double xob = xyz[3 * p], yob = xyz[3 * p + 1], zob = xyz[3 * p + 2];
    double ax = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double xj = xob - g->xs[i], yj = yob - g->ys[i], zj = zob - g->zs[i];
      double zp = g->cab[i] * xj + g->sab[i] * yj + g->salp[i] * zj;
      if (fabs(zp) > 0.5001 * g->ss[i]) continue;
      zp = xj * xj + yj * yj + zj * zj - zp * zp;
      const double b = g->bs[i];
      if (zp > 0.9 * b * b) continue;
      ax = b;
      break;
    }
    cx hx = mk(0.0, 0.0), hy = mk(0.0, 0.0), hz = mk(0.0, 0.0);
    for (int64_t i = 0; i < n; ++i) {
      fill_state st;
      memset(&st, 0, sizeof st);
      st.s = g->ss[i]; st.b = g->bs[i]; st.xj = g->xs[i]; st.yj = g->ys[i]; st.zj = g->zs[i];
      st.cabj = g->cab[i]; st.sabj = g->sab[i]; st.salpj = g->salp[i];
      cx h[9];
      hsfld(g, &st, xob, yob, zob, ax, h);
      const cx acx = mk(air[i], aii[i]), bcx = mk(bir[i], bii[i]), ccx = mk(cir[i], cii[i]);
      hx = cadd(hx, cadd(cadd(cmul(h[0], acx), cmul(h[3], bcx)), cmul(h[6], ccx)));
      hy = cadd(hy, cadd(cadd(cmul(h[1], acx), cmul(h[4], bcx)), cmul(h[7], ccx)));
      hz = cadd(hz, cadd(cadd(cmul(h[2], acx), cmul(h[5], bcx)), cmul(h[8], ccx)));
    }
    out[6 * p] = hx.re; out[6 * p + 1] = hx.im;
    out[6 * p + 2] = hy.re; out[6 * p + 3] = hy.im;
    out[6 * p + 4] = hz.re; out[6 * p + 5] = hz.im;
  }
  return 0;
}

#define TRIO_CAP 64

/* cmset + cmww, src/nec_context.cpp:2077-2203 and :2514-2654, transposed
 * fill (itrp = 1): Z[i][k] += etk*ax + ets*bx + etc*cx for each basis k in
 * trio(j), sources j in ascending order. */

/* The reflected field of efld, src/nec_context.cpp:2880-2956: t holds the
 * image fields (efld_compute with zij = zi + zj and -salpj); for the finite
 * ground they are weighted by the reflection coefficients of the specular
 * point (no radial-wire screen), then subtracted from e times frati. */
static void ground_image(const ora_geom* g, double xij, double yij, double zij, double r, cx t[9], cx e[9],
                         ora_fill_counts* cnt) {
  if (g->gtype == 2) {
    double complex zratx = CMPLX(g->zr_re, g->zr_im);
    double rmag = r;
    double xymag = sqrt(xij * xij + yij * yij);
    double px, py, cth;
    double complex zrsin;
    if (xymag <= 1.0e-6) {
      px = 0.; py = 0.; cth = 1.; zrsin = CMPLX(1.0, 0.0);
    } else {
      px = -yij / xymag;
      py = xij / xymag;
      cth = zij / rmag;
      zrsin = csqrt(1.0 - zratx * zratx * (1.0 - cth * cth));
      /* a complex square root tallied as 9 flops (2 square roots, 1 division)
       * and a complex division as 9 (3 divisions) */
      if (cnt) TALLY(cnt, 23, 4, 2, 0);
    }
    if (cnt) TALLY(cnt, 4 + 23 + 17 + 2 + 132, 6, 1, 0);
    double complex refs = (cth - zratx * zrsin) / (cth + zratx * zrsin);
    double complex refps = -(zratx * cth - zrsin) / (zratx * cth + zrsin);
    refps = refps - refs;
    for (int q = 0; q < 9; q += 3) {
      double complex tx = to_c(t[q]), ty = to_c(t[q + 1]), tz = to_c(t[q + 2]);
      double complex epy = px * tx + py * ty;
      double complex epx = px * epy;
      epy = py * epy;
      t[q] = from_c(refs * tx + refps * epx);
      t[q + 1] = from_c(refs * ty + refps * epy);
      t[q + 2] = from_c(refs * tz);
    }
  }
  cx fr = mk(g->fr_re, g->fr_im);
  for (int q = 0; q < 9; q++) e[q] = csub(e[q], cmul(t[q], fr));
  if (cnt) TALLY(cnt, 72, 0, 0, 0);
}

int ora_fill(const ora_geom* g, double rkh, int use_ek, double* Z, int nthreads, ora_fill_counts* counts) {
  int64_t n = g->n;
  memset(Z, 0, (size_t)n * (size_t)n * 2 * sizeof(double));
  /* trio tables of every source segment */
  int32_t* tj = (int32_t*)malloc((size_t)n * TRIO_CAP * sizeof(int32_t));
  double* ta = (double*)malloc((size_t)n * TRIO_CAP * sizeof(double));
  double* tb = (double*)malloc((size_t)n * TRIO_CAP * sizeof(double));
  double* tc = (double*)malloc((size_t)n * TRIO_CAP * sizeof(double));
  int* tn = (int*)malloc((size_t)n * sizeof(int));
  int bad = 0;
  for (int64_t j = 0; j < n; j++) {
    tn[j] = ora_trio(g, (int)(j + 1), tj + j * TRIO_CAP, ta + j * TRIO_CAP, tb + j * TRIO_CAP, tc + j * TRIO_CAP, TRIO_CAP);
    if (tn[j] < 0) { bad = 1; break; }
  }
  ora_fill_counts total;
  memset(&total, 0, sizeof total);
  if (g->nload > 0 && (g->load_err || !g->zarr)) bad = 1;
  if (!bad) {
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
      ora_fill_counts local;
      memset(&local, 0, sizeof local);
      fill_state st;
      memset(&st, 0, sizeof st);
      st.cnt = &local;
      st.rkh = rkh;
      st.use_ek = use_ek;
#pragma omp for schedule(dynamic, 4)
      for (int64_t i = 0; i < n; i++) {
        double xi = g->xs[i], yi = g->ys[i], zi = g->zs[i], ai = g->bs[i];
        double cabi = g->cab[i], sabi = g->sab[i], salpi = g->salp[i];
        double* row = Z + 2 * (size_t)i * (size_t)n;
        for (int64_t j = 0; j < n; j++) {
          st.s = g->ss[j]; st.b = g->bs[j];
          st.xj = g->xs[j]; st.yj = g->ys[j]; st.zj = g->zs[j];
          st.cabj = g->cab[j]; st.sabj = g->sab[j]; st.salpj = g->salp[j];
          if (use_ek) ek_ends(g, j, &st.ind1, &st.ind2);
          cx e[9];
          local.pairs++;
          efld_compute(&st, xi - st.xj, yi - st.yj, ai, st.salpj, zi - st.zj, i != j, e);
          TALLY(&local, 3, 0, 0, 0);
          if (g->gtype != 0) {
            cx t[9];
            double zij = zi + st.zj;
            double r = efld_compute(&st, xi - st.xj, yi - st.yj, ai, -st.salpj, zij, 1, t);
            ground_image(g, xi - st.xj, yi - st.yj, zij, r, t, e, &local);
            TALLY(&local, 5, 0, 0, 0);
          }
          cx etk = cadd(cadd(cmulr(e[0], cabi), cmulr(e[1], sabi)), cmulr(e[2], salpi));
          cx ets = cadd(cadd(cmulr(e[3], cabi), cmulr(e[4], sabi)), cmulr(e[5], salpi));
          cx etc = cadd(cadd(cmulr(e[6], cabi), cmulr(e[7], sabi)), cmulr(e[8], salpi));
          TALLY(&local, 30, 0, 0, 0);
          const int32_t* jj = tj + j * TRIO_CAP;
          for (int q = 0; q < tn[j]; q++) {
            int64_t k = jj[q] - 1;
            cx t = cadd(cadd(cmulr(etk, ta[j * TRIO_CAP + q]), cmulr(ets, tb[j * TRIO_CAP + q])),
                        cmulr(etc, tc[j * TRIO_CAP + q]));
            row[2 * k] += t.re;
            row[2 * k + 1] += t.im;
            local.scatter_terms++;
            TALLY(&local, 12, 0, 0, 0);
          }
          /* loading of segment j, right after its own scatter
           * (cmset, src/nec_context.cpp:2126-2142): row j, bases of trio(j) */
          if (g->nload > 0 && j == i) {
            const double zr = g->zarr[2 * i], zi = g->zarr[2 * i + 1];
            for (int q = 0; q < tn[j]; q++) {
              int64_t k = jj[q] - 1;
              double sab = ta[j * TRIO_CAP + q] + tc[j * TRIO_CAP + q];
              row[2 * k] -= sab * zr;
              row[2 * k + 1] -= sab * zi;
              TALLY(&local, 5, 0, 0, 0);
            }
          }
        }
      }
#pragma omp critical
      {
        total.pairs += local.pairs; total.far_pairs += local.far_pairs; total.near_pairs += local.near_pairs;
        total.gx_calls += local.gx_calls; total.gxx_calls += local.gxx_calls; total.intx_calls += local.intx_calls;
        total.gf_calls += local.gf_calls; total.scatter_terms += local.scatter_terms;
        total.intx_steps += local.intx_steps; total.intx_refines += local.intx_refines;
        total.intx_self += local.intx_self;
        total.flops += local.flops; total.divs += local.divs; total.sqrts += local.sqrts;
        total.specials += local.specials;
      }
    }
  }
  if (counts) *counts = total;
  free(tj); free(ta); free(tb); free(tc); free(tn);
  return bad;
}

/* ------------------------------------------------------------------ */
/* LU and solve                                                        */
/* ------------------------------------------------------------------ */
#define LU_NB 64

/* Right-looking partial-pivot LU on row-major complex Z (row pivoting; the
 * pivot is the first maximum of |z| = hypot(re, im) as in Eigen's
 * PartialPivLU, src/eigen/Eigen/src/LU/PartialPivLU.h:338-385).  Blocked for
 * speed, but each element still receives its rank-1 updates in pivot order,
 * so the result equals the unblocked algorithm bit for bit. */
int64_t ora_lu(int64_t n, double* Z, int32_t* piv, int nthreads) {
  int64_t first_zero = -1;
  if (nthreads < 1) nthreads = 1;
  cx* A = (cx*)Z;
  for (int64_t k0 = 0; k0 < n; k0 += LU_NB) {
    int64_t nb = (n - k0 < LU_NB) ? (n - k0) : LU_NB;
    int64_t kend = k0 + nb;
    /* panel: columns k0..kend-1 over rows k0..n-1 */
    for (int64_t k = k0; k < kend; k++) {
      int64_t p = k;
      double best = hypot(A[k * n + k].re, A[k * n + k].im);
      for (int64_t i = k + 1; i < n; i++) {
        double v = hypot(A[i * n + k].re, A[i * n + k].im);
        if (v > best) { best = v; p = i; }
      }
      piv[k] = (int32_t)p;
      if (best != 0.0) {
        if (p != k) {
          cx* rk = A + k * n;
          cx* rp = A + p * n;
          for (int64_t c = 0; c < n; c++) { cx t = rk[c]; rk[c] = rp[c]; rp[c] = t; }
        }
        cx d = A[k * n + k];
        for (int64_t i = k + 1; i < n; i++) A[i * n + k] = cdiv(A[i * n + k], d);
      } else if (first_zero < 0) {
        first_zero = k;
      }
      /* rank-1 update restricted to the panel columns */
#pragma omp parallel for num_threads(nthreads) schedule(static) if (n - k > 256)
      for (int64_t i = k + 1; i < n; i++) {
        cx l = A[i * n + k];
        cx* ri = A + i * n;
        const cx* rk = A + k * n;
        for (int64_t c = k + 1; c < kend; c++) ri[c] = csub(ri[c], cmul(l, rk[c]));
      }
    }
    if (kend >= n) break;
    /* U12 = L11^-1 A12 (rows k0..kend-1, columns kend..n-1) */
    for (int64_t r = k0 + 1; r < kend; r++) {
      cx* rr = A + r * n;
      for (int64_t q = k0; q < r; q++) {
        cx l = rr[q];
        const cx* rq = A + q * n;
        for (int64_t c = kend; c < n; c++) rr[c] = csub(rr[c], cmul(l, rq[c]));
      }
    }
    /* A22 -= L21 U12 */
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = kend; i < n; i++) {
      cx* ri = A + i * n;
      for (int64_t q = k0; q < kend; q++) {
        cx l = ri[q];
        const cx* rq = A + q * n;
        for (int64_t c = kend; c < n; c++) ri[c] = csub(ri[c], cmul(l, rq[c]));
      }
    }
  }
  return first_zero;
}

/* matrix_algebra.cpp solve (:271-307): permute, unit-lower forward
 * (column-oriented), upper backward (row-oriented dot, then divide). */
void ora_lu_solve(int64_t n, const double* LU, const int32_t* piv, double* b) {
  const cx* A = (const cx*)LU;
  cx* x = (cx*)b;
  for (int64_t k = 0; k < n; k++) {
    int64_t p = piv[k];
    if (p != k) { cx t = x[k]; x[k] = x[p]; x[p] = t; }
  }
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = i + 1; j < n; j++) x[j] = csub(x[j], cmul(A[j * n + i], x[i]));
  for (int64_t k = 0; k < n; k++) {
    int64_t i = n - k - 1;
    cx sum = mk(0.0, 0.0);
    for (int64_t j = i + 1; j < n; j++) sum = cadd(sum, cmul(A[i * n + j], x[j]));
    x[i] = cdiv(csub(x[i], sum), A[i * n + i]);
  }
}

/* c_geometry::get_current_coefficients, src/c_geometry.cpp:3299-3386. */
int ora_current_coefficients(const ora_geom* g, double* cur, double* abc) {
  int64_t n = g->n;
  double *air = abc, *aii = abc + n, *bir = abc + 2 * n, *bii = abc + 3 * n, *cir = abc + 4 * n, *cii = abc + 5 * n;
  memset(abc, 0, (size_t)6 * n * sizeof(double));
  int32_t jco[TRIO_CAP];
  double ax[TRIO_CAP], bx[TRIO_CAP], cxx[TRIO_CAP];
  for (int64_t i = 0; i < n; i++) {
    double ar = cur[2 * i], ai = cur[2 * i + 1];
    int jsno = ora_tbf(g, (int)(i + 1), 1, jco, ax, bx, cxx, TRIO_CAP);
    if (jsno < 0) return 1;
    for (int jx = 0; jx < jsno; jx++) {
      int64_t j = jco[jx] - 1;
      air[j] += ax[jx] * ar;
      aii[j] += ax[jx] * ai;
      bir[j] += bx[jx] * ar;
      bii[j] += bx[jx] * ai;
      cir[j] += cxx[jx] * ar;
      cii[j] += cxx[jx] * ai;
    }
  }
  for (int64_t i = 0; i < n; i++) {
    cur[2 * i] = air[i] + cir[i];
    cur[2 * i + 1] = aii[i] + cii[i];
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Far field                                                           */
/* ------------------------------------------------------------------ */

/* When set, ffld_sum tallies its operations here (ora_pattern_tally). */
static _Thread_local ora_fill_counts* ffld_tally;

/* The segment sum of ffld, src/nec_context.cpp:6706-6742, for the
 * direction cosines (rox, roy, roz). */
static void ffld_sum(const ora_geom* g, const double* abc, double rox, double roy, double roz, cx* cix, cx* ciy,
                     cx* ciz) {
  int64_t n = g->n;
  const double *air = abc, *aii = abc + n, *bir = abc + 2 * n, *bii = abc + 3 * n, *cir = abc + 4 * n, *cii = abc + 5 * n;
  *cix = mk(0, 0); *ciy = mk(0, 0); *ciz = mk(0, 0);
  for (int64_t i = 0; i < n; i++) {
    double omega = -(rox * g->cab[i] + roy * g->sab[i] + roz * g->salp[i]);
    double el = PI_T * g->ss[i];
    double sill = omega * el;
    double top = el + sill;
    double bot = el - sill;
    double a, too, boo;
    if (fabs(omega) >= 1.0e-7) a = 2. * sin(sill) / omega;
    else a = (2. - omega * omega * el * el / 3.) * el;
    if (fabs(top) >= 1.0e-7) too = sin(top) / top;
    else too = 1. - top * top / 6.;
    if (fabs(bot) >= 1.0e-7) boo = sin(bot) / bot;
    else boo = 1. - bot * bot / 6.;
    if (ffld_tally) {
      TALLY(ffld_tally, 9 + 38, 0, 0, 2);
      if (fabs(omega) >= 1.0e-7) TALLY(ffld_tally, 2, 1, 0, 1); else TALLY(ffld_tally, 6, 1, 0, 0);
      if (fabs(top) >= 1.0e-7) TALLY(ffld_tally, 1, 1, 0, 1); else TALLY(ffld_tally, 3, 1, 0, 0);
      if (fabs(bot) >= 1.0e-7) TALLY(ffld_tally, 1, 1, 0, 1); else TALLY(ffld_tally, 3, 1, 0, 0);
      ffld_tally->pairs++;
    }
    double b = el * (boo - too);
    double c = el * (boo + too);
    double rr = a * air[i] + b * bii[i] + c * cir[i];
    double ri = a * aii[i] - b * bir[i] + c * cii[i];
    double arg = TWO_PI * (g->xs[i] * rox + g->ys[i] * roy + g->zs[i] * roz);
    cx exa = cmul(mk(cos(arg), sin(arg)), mk(rr, ri));
    *cix = cadd(*cix, cmulr(exa, g->cab[i]));
    *ciy = cadd(*ciy, cmulr(exa, g->sab[i]));
    *ciz = cadd(*ciz, cmulr(exa, g->salp[i]));
  }
}

/* nec_context::ffld, src/nec_context.cpp:6639-6887, no patches; with a
 * ground (ifar 0 or 1) the structure image (roz -> -roz) is added with the
 * reflection coefficients rrv, rrh (-1 for the perfect ground,
 * :6666-6700, :6817-6822). */
void ora_ffld(const ora_geom* g, const double* abc, double sth, double cth, double sph, double cph,
              double* eth, double* eph) {
  cx const3 = mk(0.0, -ora_impedance() / FOUR_PI);
  double phx = -sph;
  double phy = cph;
  double roz = cth;
  double thx = roz * phy;
  double thy = -roz * phx;
  double thz = -sth;
  double rox = -thz * phy;
  double roy = thz * phx;
  cx cix, ciy, ciz;
  ffld_sum(g, abc, rox, roy, roz, &cix, &ciy, &ciz);
  if (g->gtype != 0) {
    double complex rrv, rrh;
    if (g->gtype == 1) {
      rrv = -CMPLX(1.0, 0.0);
      rrh = -CMPLX(1.0, 0.0);
    } else {
      double complex zr = CMPLX(g->zr_re, g->zr_im);
      double complex zrsin = csqrt(1. - zr * zr * thz * thz);
      rrv = -(roz - zr * zrsin) / (roz + zr * zrsin);
      rrh = (zr * roz - zrsin) / (zr * roz + zrsin);
    }
    double complex ccx = to_c(cix), ccy = to_c(ciy), ccz = to_c(ciz);
    cx ix, iy, iz;
    ffld_sum(g, abc, rox, roy, -roz, &ix, &iy, &iz);
    double complex ex = to_c(ix), ey = to_c(iy), ez = to_c(iz);
    double complex cdp = (ex * phx + ey * phy) * (rrh - rrv);
    cix = from_c(ccx + ex * rrv + cdp * phx);
    ciy = from_c(ccy + ey * rrv + cdp * phy);
    ciz = from_c(ccz - ez * rrv);
  }
  cx et = cmul(cadd(cadd(cmulr(cix, thx), cmulr(ciy, thy)), cmulr(ciz, thz)), const3);
  cx ep = cmul(cadd(cmulr(cix, phx), cmulr(ciy, phy)), const3);
  eth[0] = et.re; eth[1] = et.im;
  eph[0] = ep.re; eph[1] = ep.im;
}

/* nec_radiation_pattern::analyze (src/nec_radiation_pattern.cpp:209-450),
 * voltage excitation, ifar = 0, free space. */
int ora_pattern(const ora_geom* g, const double* abc, double wavelength, double pinr, double pnlr,
                const ora_rp_spec* rp, ora_rp_out* out, int nthreads) {
  int nth = rp->n_theta, nph = rp->n_phi;
  double eta = ora_impedance();
  double gcop = wavelength * wavelength * 2.0 * PI_T / (eta * pinr);
  double prad = pinr - 0.0 - pnlr;
  double gcon = gcop;
  if (rp->ipd != 0) gcon = gcon * pinr / prad;
  double exrm = 0., exra = 0.;
  if (rp->rfld >= 1.0e-20) {
    exrm = 1. / rp->rfld;
    exra = rp->rfld / wavelength;
    exra = -360. * (exra - floor(exra));
  }
  /* angle sequences accumulate exactly as the reference's loops do */
  double* thv = (double*)malloc((size_t)nth * sizeof(double));
  double* phv = (double*)malloc((size_t)nph * sizeof(double));
  double phi = rp->phi0 - rp->dphi;
  for (int k = 0; k < nph; k++) { phi += rp->dphi; phv[k] = phi; }
  double thet = rp->theta0 - rp->dtheta;
  for (int k = 0; k < nth; k++) { thet += rp->dtheta; thv[k] = thet; }
  double delta_phi_rad = deg_to_rad(rp->dphi);
  double tmp2 = 0.5 * deg_to_rad(rp->dtheta);
  double pint = 0.0;
  if (nthreads < 1) nthreads = 1;
  double* gainbuf = out->gain ? out->gain : (double*)malloc((size_t)nth * nph * sizeof(double));
  double* totbuf = out->pg_tot ? out->pg_tot : (double*)malloc((size_t)nth * nph * sizeof(double));
  double* rhbuf = out->pg_rhcp ? out->pg_rhcp : (double*)malloc((size_t)nth * nph * sizeof(double));
  double* lhbuf = out->pg_lhcp ? out->pg_lhcp : (double*)malloc((size_t)nth * nph * sizeof(double));
  double* scbuf = out->avg_scale ? out->avg_scale : (double*)malloc((size_t)nth * nph * sizeof(double));
  double* pintv = (double*)calloc((size_t)nth * nph, sizeof(double));
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
  for (int64_t idx = 0; idx < (int64_t)nth * nph; idx++) {
    int kph = (int)(idx / nth), kth = (int)(idx % nth);
    double pha = deg_to_rad(phv[kph]);
    double th = thv[kth];
    double tha = deg_to_rad(th);
    if (g->gtype != 0 && th > 90.01) {
      /* below the horizon over a ground the reference skips the direction
       * (nec_radiation_pattern.cpp:263-264) and its arrays keep unset
       * values; here they read as no field: -999.99 dB, zero E, sense 3,
       * averaging weight 0 */
      gainbuf[idx] = -999.99;
      if (out->pg_vert) out->pg_vert[idx] = -999.99;
      if (out->pg_horiz) out->pg_horiz[idx] = -999.99;
      totbuf[idx] = -999.99; rhbuf[idx] = -999.99; lhbuf[idx] = -999.99;
      if (out->e_theta) { out->e_theta[2 * idx] = 0.0; out->e_theta[2 * idx + 1] = 0.0; }
      if (out->e_phi) { out->e_phi[2 * idx] = 0.0; out->e_phi[2 * idx + 1] = 0.0; }
      if (out->axial_ratio) out->axial_ratio[idx] = 0.0;
      if (out->tilt) out->tilt[idx] = 0.0;
      if (out->sense) out->sense[idx] = 3;
      scbuf[idx] = 0.0;
      continue;
    }
    double e_t[2], e_p[2];
    ora_ffld(g, abc, sin(tha), cos(tha), sin(pha), cos(pha), e_t, e_p);
    double ethm2 = e_t[0] * e_t[0] + e_t[1] * e_t[1];
    double ethm = sqrt(ethm2);
    double etha = rad_to_deg(atan2(e_t[1], e_t[0]));
    double ephm2 = e_p[0] * e_p[0] + e_p[1] * e_p[1];
    double ephm = sqrt(ephm2);
    double epha = rad_to_deg(atan2(e_p[1], e_p[0]));
    double tilta, emajr2, eminr2, axrat;
    int sense;
    if ((ethm2 <= 1.0e-20) && (ephm2 <= 1.0e-20)) {
      tilta = 0.; emajr2 = 0.; eminr2 = 0.; axrat = 0.; sense = 3;
    } else {
      double dfaz = epha - etha, dfaz2;
      if (epha >= 0.) dfaz2 = dfaz - 360.; else dfaz2 = dfaz + 360.;
      if (fabs(dfaz) > fabs(dfaz2)) dfaz = dfaz2;
      double cdfaz = cos(deg_to_rad(dfaz));
      double tstor1 = ethm2 - ephm2;
      double tstor2 = 2. * ephm * ethm * cdfaz;
      tilta = .5 * atan2(tstor2, tstor1);
      double stilta = sin(tilta);
      tstor1 = tstor1 * stilta * stilta;
      tstor2 = tstor2 * stilta * cos(tilta);
      emajr2 = -tstor1 + tstor2 + ethm2;
      eminr2 = tstor1 - tstor2 + ephm2;
      if (eminr2 < 0.) eminr2 = 0.;
      axrat = sqrt(eminr2 / emajr2);
      tilta = rad_to_deg(tilta);
      if (axrat <= 1.0e-5) sense = 0;       /* POL_LINEAR */
      else if (dfaz <= 0.) sense = 1;       /* POL_RIGHT */
      else sense = 2;                       /* POL_LEFT */
    }
    double gnmj = db10(gcon * emajr2);
    double gnmn = db10(gcon * eminr2);
    double gnv = db10(gcon * ethm2);
    double gnh = db10(gcon * ephm2);
    double gtot = db10(gcon * (ethm2 + ephm2));
    double tg;
    switch (rp->normalization) {
      case 1: tg = gnmj; break;
      case 2: tg = gnmn; break;
      case 3: tg = gnv; break;
      case 4: tg = gnh; break;
      default: tg = gtot; break;
    }
    gainbuf[idx] = tg;
    if (rp->iavp != 0) {
      double ts = gcop * (ethm2 + ephm2);
      double tmp3 = tha - tmp2, tmp4 = tha + tmp2;
      if (kth == 0) tmp3 = tha;
      else if (kth == nth - 1) tmp4 = tha;
      double da = fabs(delta_phi_rad * (cos(tmp3) - cos(tmp4)));
      if ((kph == 0) || (kph == nph - 1)) da *= .5;
      pintv[idx] = ts * da;
      if (rp->iavp == 2) continue;
    }
    if (out->pg_vert) out->pg_vert[idx] = (rp->output_format != 1) ? gnmj : gnv;
    if (out->pg_horiz) out->pg_horiz[idx] = (rp->output_format != 1) ? gnmn : gnh;
    ethm *= wavelength;
    ephm *= wavelength;
    if (rp->rfld >= 1.0e-20) {
      ethm = ethm * exrm; etha = etha + exra;
      ephm = ephm * exrm; epha = epha + exra;
    }
    if (out->e_theta) {
      double r = deg_to_rad(etha);
      out->e_theta[2 * idx] = ethm * cos(r);
      out->e_theta[2 * idx + 1] = ethm * sin(r);
    }
    if (out->e_phi) {
      double r = deg_to_rad(epha);
      out->e_phi[2 * idx] = ephm * cos(r);
      out->e_phi[2 * idx + 1] = ephm * sin(r);
    }
    totbuf[idx] = gtot;
    {
      double a = axrat;
      if (sense == 1) a = -a;
      double gain = from_db10(gtot);
      double lf = (1 + 2 * a + a * a) / (2 * (1 + a * a));
      double rf = (1 - 2 * a + a * a) / (2 * (1 + a * a));
      rhbuf[idx] = db10(gain * rf);
      lhbuf[idx] = db10(gain * lf);
    }
    if (out->axial_ratio) out->axial_ratio[idx] = axrat;
    if (out->tilt) out->tilt[idx] = tilta;
    if (out->sense) out->sense[idx] = sense;
    scbuf[idx] = (sin(tha) * (nth - 1) / nth) + 1.0 / nth;
  }
  /* pint accumulates in the reference's loop order (phi outer, theta inner) */
  for (int64_t idx = 0; idx < (int64_t)nth * nph; idx++) pint += pintv[idx];
  if (rp->iavp != 0) {
    double ts = deg_to_rad(rp->theta0);
    double tr = deg_to_rad(rp->dtheta) * (double)(nth - 1);
    double pr = deg_to_rad(rp->dphi) * (double)(nph - 1);
    double tt = ts + tr;
    double sa = fabs(pr * (cos(ts) - cos(tt)));
    out->average_power_gain = pint / sa;
    out->average_power_solid_angle = sa / PI_T;
  } else {
    out->average_power_gain = 0.0;
    out->average_power_solid_angle = 0.0;
  }
  double mx = gainbuf[0];
  for (int64_t idx = 1; idx < (int64_t)nth * nph; idx++) if (gainbuf[idx] > mx) mx = gainbuf[idx];
  out->max_gain = mx;
  /* statistics, nec_radiation_pattern.h:252-333 (rows = theta, cols = phi) */
  const double* arrs[3] = {totbuf, rhbuf, lhbuf};
  for (int a = 0; a < 3; a++) {
    const double* p = arrs[a];
    double mxv = p[0], mnv = p[0];
    for (int64_t idx = 1; idx < (int64_t)nth * nph; idx++) {
      if (p[idx] > mxv) mxv = p[idx];
      if (p[idx] < mnv) mnv = p[idx];
    }
    double sum = 0.0;
    for (int i = 0; i < nth; i++)
      for (int j = 0; j < nph; j++) sum += p[(int64_t)j * nth + i] * scbuf[(int64_t)j * nth + i];
    int64_t len = (int64_t)nth * nph;
    double mean = sum / ((double)(len * 2) / PI_T);
    double s2 = 0.0;
    for (int i = 0; i < nth; i++)
      for (int j = 0; j < nph; j++) {
        double diff = p[(int64_t)j * nth + i] - mean;
        s2 += diff * diff * scbuf[(int64_t)j * nth + i];
      }
    double sdv = sqrt(s2 / ((double)(len * 2) / PI_T));
    out->stats[4 * a + 0] = mxv;
    out->stats[4 * a + 1] = mnv;
    out->stats[4 * a + 2] = mean;
    out->stats[4 * a + 3] = sdv;
  }
  if (!out->gain) free(gainbuf);
  if (!out->pg_tot) free(totbuf);
  if (!out->pg_rhcp) free(rhbuf);
  if (!out->pg_lhcp) free(lhbuf);
  if (!out->avg_scale) free(scbuf);
  free(pintv); free(thv); free(phv);
  return 0;
}

/* The operation tally of the pattern's segment sums (ffld_sum) over the
 * directions of rp, in the convention of ora_fill_counts: pairs counts the
 * (direction, segment) terms; the per-direction work around the sums is not
 * tallied.  Directions below the horizon over a ground are skipped as in
 * ora_pattern. */
void ora_pattern_tally(const ora_geom* g, const double* abc, const ora_rp_spec* rp, int nthreads,
                       ora_fill_counts* out) {
  int nth = rp->n_theta, nph = rp->n_phi;
  double* thv = (double*)malloc((size_t)nth * sizeof(double));
  double* phv = (double*)malloc((size_t)nph * sizeof(double));
  double phi = rp->phi0 - rp->dphi;
  for (int k = 0; k < nph; k++) { phi += rp->dphi; phv[k] = phi; }
  double thet = rp->theta0 - rp->dtheta;
  for (int k = 0; k < nth; k++) { thet += rp->dtheta; thv[k] = thet; }
  ora_fill_counts total;
  memset(&total, 0, sizeof total);
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
  {
    ora_fill_counts local;
    memset(&local, 0, sizeof local);
    ffld_tally = &local;
#pragma omp for schedule(dynamic, 16)
    for (int64_t idx = 0; idx < (int64_t)nth * nph; idx++) {
      int kph = (int)(idx / nth), kth = (int)(idx % nth);
      if (g->gtype != 0 && thv[kth] > 90.01) continue;
      double pha = deg_to_rad(phv[kph]), tha = deg_to_rad(thv[kth]);
      double e_t[2], e_p[2];
      ora_ffld(g, abc, sin(tha), cos(tha), sin(pha), cos(pha), e_t, e_p);
    }
    ffld_tally = NULL;
#pragma omp critical
    {
      total.pairs += local.pairs; total.flops += local.flops; total.divs += local.divs;
      total.sqrts += local.sqrts; total.specials += local.specials;
    }
  }
  free(thv); free(phv);
  *out = total;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------ */
/* Networks and transmission lines                                     */
/* ------------------------------------------------------------------ */

/* tl_card / nt_card, src/nec_context.cpp:630-726: replaces the networks
 * (count 0 clears them); a line with Z0 <= 0 is crossed (type 3, |Z0|). */
int ora_set_networks(ora_geom* g, int count, const int32_t* type, const int64_t* seg1, const int64_t* seg2,
                     const double* x) {
  g->nnet = count;
  g->ntyp = (int32_t*)xrealloc(g->ntyp, (size_t)count * sizeof(int32_t));
  g->nseg1 = (int64_t*)xrealloc(g->nseg1, (size_t)count * sizeof(int64_t));
  g->nseg2 = (int64_t*)xrealloc(g->nseg2, (size_t)count * sizeof(int64_t));
  g->nx = (double*)xrealloc(g->nx, (size_t)count * 6 * sizeof(double));
  for (int k = 0; k < count; ++k) {
    if (seg1[k] < 1 || seg1[k] > g->n || seg2[k] < 1 || seg2[k] > g->n) {
      set_err("network %d: port segment out of range", k + 1);
      g->nnet = 0;
      return 1;
    }
    g->ntyp[k] = type[k];
    g->nseg1[k] = seg1[k];
    g->nseg2[k] = seg2[k];
    for (int q = 0; q < 6; ++q) g->nx[6 * k + q] = x[6 * k + q];
    if (type[k] == 2 && x[6 * k] <= 0.0) {
      g->ntyp[k] = 3;
      g->nx[6 * k] = -x[6 * k];
    }
  }
  return 0;
}

/* The short-circuit admittance parameters of network k at the current
 * wavelength, netwk_solve_and_check src/nec_context.cpp:4660-4688: a line
 * of Z0 and length l (beta l = 2 pi l / lambda) has y12 = j / (Z0 sin beta
 * l), y11 = y22 = -j cot(beta l) / Z0 plus its end admittances; crossed, y12
 * changes sign.  Lengths <= 0 were replaced by the distance of the port
 * centres (calculate_network_data, :1210-1241). */
static void network_y(const ora_geom* g, int k, double wl, cx* y11, cx* y12, cx* y22) {
  const double* x = g->nx + 6 * k;
  if (g->ntyp[k] <= 1) {
    *y11 = mk(x[0], x[1]);
    *y12 = mk(x[2], x[3]);
    *y22 = mk(x[4], x[5]);
    return;
  }
  const double bl = TWO_PI * x[1] / wl;
  const double y12i = 1. / (x[0] * sin(bl));
  const double yc = -y12i * cos(bl);
  *y11 = mk(x[2], yc + x[3]);
  *y12 = g->ntyp[k] == 2 ? mk(0., y12i) : mk(-0., -y12i);
  *y22 = mk(x[4], yc + x[5]);
}

/* netwk, src/nec_context.cpp:4496-5045 (ntsol = 0 each run): the port
 * segments that are not sources get network equations, numbered in order of
 * appearance; ports on sources are kept as source rows (the voltage known).
 * The network matrix M (nteq x nteq) is the structure's short-circuit
 * admittance between the equation segments, M(j, i) = current at segment j
 * of a unit field at segment i (one structure solve each), less y s of the
 * networks; M v = rhnx + structure currents of the incident field, then the
 * port fields v are subtracted from the incident field and the structure
 * solved again.  e: the incident field on entry, the segment currents (not
 * scaled by the wavelength) on return; abc: their coefficients.  A source
 * on a port gets the network's current added (cur_src); netloss: the
 * network's power loss (network_power_loss). */
static int solve_with_networks(ora_geom* g, const double* LU, const int32_t* piv, double wl, int nsrc,
                               const int64_t* src_seg, const double* v, double* e, double* abc, cx* cur_src,
                               double* netloss) {
  const int64_t n = g->n;
  const int nn = g->nnet;
  /* TL lengths from the port centres, once (the reference stores them) */
  for (int k = 0; k < nn; ++k)
    if (g->ntyp[k] >= 2 && g->nx[6 * k + 1] <= 0.0) {
      const int64_t a = g->nseg1[k] - 1, b = g->nseg2[k] - 1;
      const double xx = g->xs[b] - g->xs[a], yy = g->ys[b] - g->ys[a], zz = g->zs[b] - g->zs[a];
      g->nx[6 * k + 1] = wl * sqrt(xx * xx + yy * yy + zz * zz);
    }
  const int ndimn = 2 * nn + nsrc;
  int64_t* nteqa = (int64_t*)calloc((size_t)ndimn, sizeof(int64_t));
  int64_t* ntsca = (int64_t*)calloc((size_t)ndimn, sizeof(int64_t));
  cx* vsrc = (cx*)calloc((size_t)ndimn, sizeof(cx));
  cx* cmn = (cx*)calloc((size_t)ndimn * ndimn, sizeof(cx)); /* fill(i, j) at i + j ndimn */
  cx* rhnx = (cx*)calloc((size_t)ndimn, sizeof(cx));
  double* rhs = (double*)malloc((size_t)2 * n * sizeof(double));
  double* tmp_abc = (double*)malloc((size_t)6 * n * sizeof(double));
  int nteq = 0, ntsc = 0, rc = 0;
  /* the row of a port segment: a source row from the end, or an equation */
  for (int k = 0; k < nn; ++k) {
    cx y11, y12, y22;
    network_y(g, k, wl, &y11, &y12, &y22);
    int isc[2], row[2];
    const int64_t seg[2] = {g->nseg1[k], g->nseg2[k]};
    for (int p = 0; p < 2; ++p) {
      isc[p] = -1;
      for (int i = 0; i < nsrc; ++i)
        if (seg[p] == src_seg[i]) { isc[p] = i; break; }
      int found = -1;
      if (isc[p] < 0) {
        for (int i = 0; i < nteq; ++i)
          if (nteqa[i] == seg[p]) { found = i; break; }
        if (found < 0) { found = nteq; nteqa[nteq++] = seg[p]; }
        row[p] = found;
      } else {
        for (int i = 0; i < ntsc; ++i)
          if (ntsca[i] == seg[p]) { found = ndimn - (i + 1); break; }
        if (found < 0) {
          found = ndimn - (ntsc + 1);
          ntsca[ntsc] = seg[p];
          vsrc[ntsc] = mk(v[2 * isc[p]], v[2 * isc[p] + 1]);
          ntsc++;
        }
        row[p] = found;
      }
    }
    const double s1 = g->ss[seg[0] - 1], s2 = g->ss[seg[1] - 1];
    if (isc[0] < 0) {
      cmn[row[0] + row[0] * ndimn] = csub(cmn[row[0] + row[0] * ndimn], cmulr(y11, s1));
      cmn[row[0] + row[1] * ndimn] = csub(cmn[row[0] + row[1] * ndimn], cmulr(y12, s1));
    } else {
      const cx vs = mk(v[2 * isc[0]], v[2 * isc[0] + 1]);
      rhnx[row[0]] = cadd(rhnx[row[0]], cdivr(cmul(y11, vs), wl));
      rhnx[row[1]] = cadd(rhnx[row[1]], cdivr(cmul(y12, vs), wl));
    }
    if (isc[1] < 0) {
      cmn[row[1] + row[1] * ndimn] = csub(cmn[row[1] + row[1] * ndimn], cmulr(y22, s2));
      cmn[row[1] + row[0] * ndimn] = csub(cmn[row[1] + row[0] * ndimn], cmulr(y12, s2));
    } else {
      const cx vs = mk(v[2 * isc[1]], v[2 * isc[1] + 1]);
      rhnx[row[0]] = cadd(rhnx[row[0]], cdivr(cmul(y12, vs), wl));
      rhnx[row[1]] = cadd(rhnx[row[1]], cdivr(cmul(y22, vs), wl));
    }
  }
  /* structure admittances between the equation segments */
  for (int i = 0; i < nteq; ++i) {
    memset(rhs, 0, (size_t)2 * n * sizeof(double));
    rhs[2 * (nteqa[i] - 1)] = 1.0;
    ora_lu_solve(n, LU, piv, rhs);
    if (ora_current_coefficients(g, rhs, tmp_abc)) { rc = 1; goto out; }
    for (int j = 0; j < nteq; ++j) {
      const int64_t r = nteqa[j] - 1;
      cmn[i + j * ndimn] = cadd(cmn[i + j * ndimn], mk(rhs[2 * r], rhs[2 * r + 1]));
    }
  }
  {
    /* M(r, c) = fill(c, r), factored with partial pivoting */
    double* M = (double*)malloc((size_t)2 * (nteq > 0 ? nteq : 1) * (nteq > 0 ? nteq : 1) * sizeof(double));
    int32_t* mpiv = (int32_t*)malloc((size_t)(nteq > 0 ? nteq : 1) * sizeof(int32_t));
    for (int r = 0; r < nteq; ++r)
      for (int c = 0; c < nteq; ++c) {
        M[2 * (r * nteq + c)] = cmn[c + r * ndimn].re;
        M[2 * (r * nteq + c) + 1] = cmn[c + r * ndimn].im;
      }
    if (nteq > 0) ora_lu(nteq, M, mpiv, 1);
    /* the incident field's structure currents at the equation segments */
    memcpy(rhs, e, (size_t)2 * n * sizeof(double));
    ora_lu_solve(n, LU, piv, rhs);
    if (ora_current_coefficients(g, rhs, tmp_abc)) { free(M); free(mpiv); rc = 1; goto out; }
    double* rhnt = (double*)calloc((size_t)2 * (nteq > 0 ? nteq : 1), sizeof(double));
    for (int i = 0; i < nteq; ++i) {
      const int64_t r = nteqa[i] - 1;
      rhnt[2 * i] = rhnx[i].re + rhs[2 * r];
      rhnt[2 * i + 1] = rhnx[i].im + rhs[2 * r + 1];
    }
    if (nteq > 0) ora_lu_solve(nteq, M, mpiv, rhnt);
    for (int i = 0; i < nteq; ++i) {
      const int64_t r = nteqa[i] - 1;
      e[2 * r] -= rhnt[2 * i];
      e[2 * r + 1] -= rhnt[2 * i + 1];
    }
    ora_lu_solve(n, LU, piv, e);
    if (ora_current_coefficients(g, e, abc)) { free(M); free(mpiv); free(rhnt); rc = 1; goto out; }
    /* power of the network ports (netwk_compute_currents, :4891-4918) */
    double loss = 0.0;
    for (int i = 0; i < nteq; ++i) {
      const int64_t r = nteqa[i] - 1;
      const cx volt = cmulr(cmulr(mk(rhnt[2 * i], rhnt[2 * i + 1]), g->ss[r]), wl);
      const cx cur = cmulr(mk(e[2 * r], e[2 * r + 1]), wl);
      loss -= 0.5 * cmul(volt, mk(cur.re, -cur.im)).re;
    }
    for (int i = 0; i < ntsc; ++i) {
      const int64_t r = ntsca[i] - 1;
      const cx cur = cmulr(mk(e[2 * r], e[2 * r + 1]), wl);
      loss -= 0.5 * cmul(vsrc[i], mk(cur.re, -cur.im)).re;
    }
    /* sources on ports: the network's current joins the structure's
     * (netwk_compute_inputs, :4954-4985) */
    for (int s = 0; s < nsrc; ++s) {
      const int64_t r = src_seg[s] - 1;
      cx cur = mk(e[2 * r], e[2 * r + 1]);
      if (ntsc != 0) {
        int j = -1;
        for (int q = 0; q < ntsc; ++q)
          if (ntsca[q] == src_seg[s]) { j = q; break; }
        if (j < 0) {
          set_err("NETWORK DATA ERROR: SOURCE SEGMENT [%lld] NOT FOUND IN NETWORK TABLE", (long long)src_seg[s]);
          free(M); free(mpiv); free(rhnt);
          rc = 1;
          goto out;
        }
        const int rowi = ndimn - (j + 1);
        cx t = rhnx[rowi];
        for (int k = 0; k < nteq; ++k) t = csub(t, cmul(cmn[k + rowi * ndimn], mk(rhnt[2 * k], rhnt[2 * k + 1])));
        cur = cadd(cur, t);
        const cx c = cmulr(cur, wl);
        const cx vs = mk(v[2 * s], v[2 * s + 1]);
        loss += 0.5 * cmul(vs, mk(c.re, -c.im)).re;
      }
      cur_src[s] = cur;
    }
    *netloss = loss;
    free(M); free(mpiv); free(rhnt);
  }
out:
  free(nteqa); free(ntsca); free(vsrc); free(cmn); free(rhnx); free(rhs); free(tmp_abc);
  return rc;
}

/* One voltage-source run at one frequency: simulate() states
 * MEMORY_ALLOC -> STRUCTURE_LOADING -> EXCITATION (src/nec_context.cpp:1031-1159,
 * :1523-1558, :1622-1791, :3193-3210, :4850-5047). */
int ora_run_voltage(ora_geom* g, double freq_mhz, double rkh, int use_ek,
                    int nsrc, const int64_t* src_seg, const double* v,
                    const ora_rp_spec* rp, ora_rp_out* rp_out,
                    double* zin, double* current, double* input_power, double* abc_out,
                    int nthreads, ora_timing* timing) {
  int64_t n = g->n;
  double wl = ora_wavelength(freq_mhz);
  ora_frequency_scale(g, freq_mhz);
  double t0 = now_s();
  double* Z = (double*)malloc((size_t)n * n * 2 * sizeof(double));
  if (!Z) { set_err("oracle: cannot allocate Z"); return 1; }
  if (ora_fill(g, rkh, use_ek, Z, nthreads, NULL)) { free(Z); return 1; }
  double t1 = now_s();
  int32_t* piv = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  ora_lu(n, Z, piv, nthreads);
  double t2 = now_s();
  double* e = (double*)calloc((size_t)2 * n, sizeof(double));
  for (int s = 0; s < nsrc; s++) {
    int64_t k = src_seg[s] - 1;
    cx val = cdivr(cneg(mk(v[2 * s], v[2 * s + 1])), g->ss[k] * wl);
    e[2 * k] = val.re;
    e[2 * k + 1] = val.im;
  }
  double* abc = abc_out ? abc_out : (double*)malloc((size_t)6 * n * sizeof(double));
  cx* cur_src = (cx*)calloc((size_t)(nsrc > 0 ? nsrc : 1), sizeof(cx));
  double netloss = 0.0;
  if (g->nnet > 0) {
    if (solve_with_networks(g, Z, piv, wl, nsrc, src_seg, v, e, abc, cur_src, &netloss)) {
      free(Z); free(piv); free(e); free(cur_src);
      if (!abc_out) free(abc);
      return 1;
    }
  } else {
    ora_lu_solve(n, Z, piv, e);
    if (ora_current_coefficients(g, e, abc)) { free(Z); free(piv); free(e); free(cur_src); return 1; }
    for (int s = 0; s < nsrc; s++) cur_src[s] = mk(e[2 * (src_seg[s] - 1)], e[2 * (src_seg[s] - 1) + 1]);
  }
  double pin = 0.0;
  for (int s = 0; s < nsrc; s++) {
    cx volt = mk(v[2 * s], v[2 * s + 1]);
    cx cur = cmulr(cur_src[s], wl);
    cx z = cdiv(volt, cur);
    if (zin) { zin[2 * s] = z.re; zin[2 * s + 1] = z.im; }
    double p = 0.5 * cmul(volt, mk(cur.re, -cur.im)).re;
    pin += p;
  }
  if (input_power) *input_power = pin;
  if (current)
    for (int64_t i = 0; i < n; i++) { current[2 * i] = e[2 * i] * wl; current[2 * i + 1] = e[2 * i + 1] * wl; }
  double t3 = now_s();
  /* ohmic loss in the loads, nec_context::compute_structure_power_loss
   * (src/nec_context.cpp:1423-1441); the pattern's radiated power is the
   * input power less it (prad, used for ipd = 1) */
  double ploss = 0.0;
  if (g->nload > 0)
    for (int64_t i = 0; i < n; i++)
      if (fabs(g->zarr[2 * i]) >= 1.e-20) {
        double cm = hypot(e[2 * i] * wl, e[2 * i + 1] * wl);
        ploss += 0.5 * cm * cm * g->zarr[2 * i] * g->ss[i];
      }
  if (rp && rp_out) ora_pattern(g, abc, wl, pin, ploss + netloss, rp, rp_out, nthreads);
  free(cur_src);
  double t4 = now_s();
  if (timing) { timing->fill_s = t1 - t0; timing->factor_s = t2 - t1; timing->solve_s = t3 - t2; timing->pattern_s = t4 - t3; }
  if (!abc_out) free(abc);
  free(Z); free(piv); free(e);
  return 0;
}
