/*
 * necpp_oracle.h -- CPU restatement of nec2++'s free-space thin-wire
 * method-of-moments path (geometry -> basis -> fill -> LU -> solve ->
 * currents / Z_in -> far-field pattern).
 *
 * THIS IS TEST INFRASTRUCTURE. It exists only so that tests/, the smoke
 * check and the bench's cpu_baseline leg can check and time the HIP product
 * path against an independent CPU statement of the reference algorithm.
 * Nothing in the product library links, loads or calls it.
 *
 * Each function cites the reference routine it follows (paths relative to
 * the nec2++ source tree).  Arithmetic follows the reference's operation
 * order with IEEE binary64 and no FMA contraction (build with
 * -ffp-contract=off), so geometry, basis coefficients and matrix entries are
 * bit-identical to nec2++ compiled with g++ on x86-64.  The LU is a plain
 * right-looking partial-pivot factorisation (same pivot rule as Eigen's
 * PartialPivLU: first maximum of |z|), so its rounding differs from Eigen's
 * blocked kernel at the 1e-15 * cond level.
 */
#ifndef NECPP_ORACLE_H
#define NECPP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_geom ora_geom;

/* Last error text (static storage, overwritten by the next error). */
const char* ora_last_error(void);

ora_geom* ora_geom_new(void);
void ora_geom_free(ora_geom* g);

/* c_geometry::wire (src/c_geometry.cpp:614-683).  0 = ok. */
int ora_wire(ora_geom* g, int tag, int ns, double x1, double y1, double z1,
             double x2, double y2, double z2, double rad, double rdel, double rrad);

/* c_geometry::geometry_complete (src/c_geometry.cpp:454-604) with
 * build_connections (:1705-1915) and resolve_junctions (:1918-2074).
 * 0 = ok, 1 = geometry error (message in ora_last_error). */
int ora_geometry_complete(ora_geom* g, int gpflag);

int64_t ora_nseg(const ora_geom* g);

/* Per-segment unscaled data after geometry_complete: out[8*i + 0..7] =
 * x, y, z (centre), length, radius, cab, sab, salp.  icon1/icon2/tags may be
 * NULL. */
void ora_segments(const ora_geom* g, double* out, int32_t* icon1, int32_t* icon2, int32_t* tags);

/* nec_ground::parse_gn (src/nec_ground.cpp:75-121): iperf -1 free space,
 * 0 finite ground (reflection-coefficient approximation, epsr and sig
 * nonzero), 1 perfect ground; no radial wires, no Sommerfeld ground.  Takes
 * effect at the next ora_frequency_scale (which sets zrati, frati).  The
 * fill then adds the image term of efld (src/nec_context.cpp:2879-2956),
 * ffld the structure image and the pattern skips theta > 90.01 deg.
 * 0 = ok. */
int ora_set_ground(ora_geom* g, int iperf, double epsr, double sig);

/* nec_context::ld_card (src/nec_context.cpp:424-474) for count cards (0
 * clears): types 0-5 (series / parallel RLC, per metre or lumped, fixed
 * impedance, wire conductivity), tag and segment range as the card.  The
 * impedances are formed at the next ora_frequency_scale (load(),
 * :1861-2055); the fill subtracts them on each loaded segment's row
 * (cmset :2126-2142) and the pattern's radiated power excludes their loss.
 * 0 = ok. */
/* Test hook: nec_context::zint (src/nec_context.cpp:6321-6406), lambda *
 * internal impedance per unit length of a round wire; out[2]. */
void ora_zint(double sigl, double rolam, double* out);
int ora_set_loads(ora_geom* g, int count, const int32_t* type, const int32_t* tag, const int32_t* tagf,
                  const int32_t* tagt, const double* zlr, const double* zli, const double* zlc);

/* c_geometry::get_segment_number (src/c_geometry.cpp:76-99): 1-based segment
 * number, or -1 when no such tag (error text set). */
int64_t ora_segment_number(const ora_geom* g, int tag, int m);

/* em::get_wavelength(1e6 * f) with the reference's truncated constants
 * (src/electromag.cpp:25-35, src/math_util.h:72-74). */
double ora_wavelength(double freq_mhz);
double ora_speed_of_light(void);
double ora_impedance(void);

/* c_geometry::frequency_scale (src/c_geometry.cpp:3389-3418). */
/* Decision tape: record the fill's geometry-dependent decisions (near/far
 * branch per pair, every adaptive-quadrature accept/refine choice, the
 * self-term integrand's small-argument branch) during one run, then replay
 * them in later runs, for central differences with the decisions held
 * fixed.  Per thread; taped runs must use one thread.  stats: decisions
 * recorded, decisions replayed, replay ran past the tape (0/1). */
void ora_tape_record(void);
void ora_tape_replay(void);
void ora_tape_off(void);
void ora_tape_stats(int64_t* out);

/* GM (c_geometry::move, src/c_geometry.cpp:830-900), angles in radians;
 * GS (c_geometry::scale, :1308-1318).  Both act on the segments defined so
 * far, before ora_geometry_complete.  ora_move returns 1 (message in
 * ora_last_error) when its names a tag no segment has. */
int ora_move(ora_geom* g, double rox, double roy, double roz, double xs, double ys, double zs, int its, int nrpt,
             int itgi);
void ora_scale(ora_geom* g, double f);
void ora_frequency_scale(ora_geom* g, double freq_mhz);

/* c_geometry::trio + sbf (src/c_geometry.cpp:2982-3289) on the scaled
 * geometry: the basis functions with support on segment j (1-based).
 * Returns jsno (or -1 on a connection error); jco is 1-based. */
int ora_trio(const ora_geom* g, int j, int32_t* jco, double* ax, double* bx, double* cx, int cap);

/* c_geometry::tbf (src/c_geometry.cpp:2763-2977). */
int ora_tbf(const ora_geom* g, int i, int icap, int32_t* jco, double* ax, double* bx, double* cx, int cap);

/* Counters of the last ora_fill call (for algorithmic-work accounting). */
typedef struct {
  int64_t pairs;        /* (observation, source) pairs evaluated         */
  int64_t far_pairs;    /* r >= rkh lumped-element branch                */
  int64_t near_pairs;   /* eksc / ekscx branch                            */
  int64_t gx_calls;
  int64_t gxx_calls;
  int64_t intx_calls;
  int64_t gf_calls;
  int64_t scatter_terms;/* sum over pairs of jsno(j)                      */
  int64_t intx_steps;   /* intx: three-point estimates (one per step tried) */
  int64_t intx_refines; /* intx: five-point estimates (steps not accepted at three) */
  int64_t intx_self;    /* intx calls for the self term (ij = 0)           */
  /* Operation tally, kept per basic block: each +, -, *, / and square root
   * executed is one flop (divisions and square roots are also tallied
   * alone); each sin, cos and log is one "special" and no flop.  Negation,
   * fabs, comparisons and expressions of constants count nothing. */
  int64_t flops;
  int64_t divs;
  int64_t sqrts;
  int64_t specials;
} ora_fill_counts;

/* nec_context::cmset / cmww / efld / efld_compute / eksc / ekscx / gf / gx /
 * gxx / intx (src/nec_context.cpp:2077-4297) for free space, no loads, no
 * symmetry.  Z is N x N complex (interleaved re,im) in the reference's
 * "cm" layout: Z[i][k] at 2*(i*N + k), i = observation segment, k = basis.
 * nthreads > 1 splits observation rows over OpenMP threads (same values). */
int ora_fill(const ora_geom* g, double rkh, int use_ek, double* Z, int nthreads, ora_fill_counts* counts);

/* lu_decompose + Eigen PartialPivLU (src/matrix_algebra.cpp:217-266,
 * src/eigen/Eigen/src/LU/PartialPivLU.h:338-465) on the row-major Z above
 * (row pivoting on Z = column-major Z^T of the reference's cm): in place
 * L\U, piv[k] = row swapped with k at step k (0-based transpositions).
 * Returns -1, or the first exactly-zero pivot index. */
int64_t ora_lu(int64_t n, double* Z, int32_t* piv, int nthreads);

/* matrix_algebra.cpp solve (:271-307): b <- Z^-1 b. */
void ora_lu_solve(int64_t n, const double* LU, const int32_t* piv, double* b);

/* c_geometry::get_current_coefficients (src/c_geometry.cpp:3299-3386) for
 * nqds = 0, no patches.  cur: in = solution vector, out = curx (A+C).
 * abc: 6*N = air, aii, bir, bii, cir, cii. */
int ora_current_coefficients(const ora_geom* g, double* cur, double* abc);

/* nec_context::ffld (src/nec_context.cpp:6639-6887), free space: E_theta and
 * E_phi for direction (thet, phi) given in radians through their sines and
 * cosines. */
void ora_ffld(const ora_geom* g, const double* abc, double sth, double cth, double sph, double cph,
              double* eth, double* eph);

/* Test hooks exposing single kernels of the fill (free space, k = 2*pi):
 * ora_eksc / ora_ekscx: out[12] = ezs, ers, ezc, erc, ezk, erk (re, im) of a
 * source segment of length s (wavelengths) at axial offset z and radial
 * distance rh; ij = 0 for the self term.
 * ora_intx: out[2] = (cint, sint) of intx over [el1, el2] with the gf state
 * (zpk, rkb2, ija).
 * ora_near_e: nefld (src/nec_context.cpp:4304-4450) for wires, in free
 * space or over the ground of ora_set_ground: E at npts points (xyz in
 * wavelengths, at the last ora_frequency_scale) of the current coefficients
 * abc; out holds 6 doubles per point.
 * ora_near_h: nhfld (src/nec_context.cpp:5189-5262) for wires, likewise: H
 * at the points, 6 doubles per point. */
int ora_near_e(const ora_geom* g, const double* abc, int64_t npts, const double* xyz, double rkh, int use_ek,
               double* out);
int ora_near_h(const ora_geom* g, const double* abc, int64_t npts, const double* xyz, double* out);

/*
 * ora_efld: out[18] = the 9 complex field components egnd of efld_compute
 * for a source segment (centre xj.., cosines, length s, radius b) at the
 * observation point (xi, yi, zi) with observation radius ai. */
void ora_eksc(double s, double z, double rh, int ij, double* out);
void ora_ekscx(double b, double s, double z, double rh, int ij, int inx1, int inx2, double* out);
void ora_intx(double zpk, double rkb2, int ija, double el1, double el2, double b, int ij, double* out);
double ora_efld(double xj, double yj, double zj, double cabj, double sabj, double salpj, double s, double b,
                double xi, double yi, double zi, double ai, int not_self, double rkh, int use_ek, int ind1, int ind2,
                double* out);

/* Pattern specification as accepted by nec_rp_card (ifar must be 0). */
typedef struct {
  int n_theta, n_phi;
  double theta0, phi0, dtheta, dphi;  /* degrees */
  int output_format;                  /* XNDA X: 0 major/minor, 1 vert/horiz */
  int normalization;                  /* XNDA N */
  int ipd;                            /* XNDA D */
  int iavp;                           /* XNDA A */
  double rfld;                        /* radial distance (m) */
  double gnor;
} ora_rp_spec;

/* Pattern results; arrays are n_theta*n_phi, index phi_i*n_theta + theta_i
 * (nec_radiation_pattern.cpp:28-35). Any pointer may be NULL. */
typedef struct {
  double* gain;        /* _gain (normalisation-selected dB)        */
  double* pg_vert;     /* _power_gain_vert                         */
  double* pg_horiz;    /* _power_gain_horiz                        */
  double* pg_tot;      /* _power_gain_tot                          */
  double* pg_rhcp;
  double* pg_lhcp;
  double* axial_ratio;
  double* tilt;
  int32_t* sense;
  double* avg_scale;   /* _averaging_scales                        */
  double* e_theta;     /* 2 per point (re, im) after lambda / range scaling */
  double* e_phi;
  /* scalars */
  double max_gain;     /* _gain.maxCoeff()                         */
  double average_power_gain;
  double average_power_solid_angle;
  double stats[12];    /* tot max,min,mean,sd; rhcp max,min,mean,sd; lhcp ... */
} ora_rp_out;

/* nec_radiation_pattern::analyze (src/nec_radiation_pattern.cpp:209-450) for
 * ifar = 0 and voltage excitation, plus the statistics of
 * nec_radiation_pattern.h:252-333.  pinr = input power, pnlr = network loss.
 * nthreads > 1 splits directions over OpenMP threads. */
int ora_pattern(const ora_geom* g, const double* abc, double wavelength, double pinr, double pnlr,
                const ora_rp_spec* rp, ora_rp_out* out, int nthreads);

/* Operation tally of ora_pattern's segment sums over rp's directions (pairs
 * = direction-segment terms; flops, divs, sqrts, specials as in
 * ora_fill_counts; the other fields stay zero). */
void ora_pattern_tally(const ora_geom* g, const double* abc, const ora_rp_spec* rp, int nthreads,
                       ora_fill_counts* out);

/* One complete voltage-source run at one frequency, as XQ/RP would do it:
 * frequency_scale, fill, factor, RHS (etmns_voltage_source), solve, current
 * coefficients, Z_in per source (netwk_compute_inputs) and, if rp != NULL,
 * the pattern.  src_seg are 1-based segment numbers, v = 2*nsrc doubles.
 * Outputs (any may be NULL): zin (2*nsrc), current (2*N, curx * lambda),
 * input_power, abc (6*N). */
typedef struct {
  double fill_s, factor_s, solve_s, pattern_s;
} ora_timing;

/* tl_card / nt_card (src/nec_context.cpp:630-726): replaces the networks
 * of the structure (count 0 clears them).  type 1 network (x: y11, y12, y22
 * as re, im), 2 transmission line (x: Z0 (<= 0: crossed), length in m (<=
 * 0: the distance of the port segments' centres), shunt admittance of end 1
 * and end 2 (re, im)); seg1, seg2 the 1-based port segments.  ora_run_voltage
 * then solves with them (netwk, :4496-5045). */
int ora_set_networks(ora_geom* g, int count, const int32_t* type, const int64_t* seg1, const int64_t* seg2,
                     const double* x);

int ora_run_voltage(ora_geom* g, double freq_mhz, double rkh, int use_ek,
                    int nsrc, const int64_t* src_seg, const double* v,
                    const ora_rp_spec* rp, ora_rp_out* rp_out,
                    double* zin, double* current, double* input_power, double* abc,
                    int nthreads, ora_timing* timing);

#ifdef __cplusplus
}
#endif

#endif
