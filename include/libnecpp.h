/*
 * libnecpp.h -- C interface of the MI355X method-of-moments engine.
 *
 * Drop-in for nec2++'s libnecpp.h: same function names, argument order and
 * meaning, same return conventions (card functions return 0 on success and
 * 1 on error with the text available from nec_error_message(); result
 * getters return -999.0 when the requested result does not exist).  Each
 * declaration cites the reference declaration it replaces
 * (src/libnecpp.h:<line> in the nec2++ tree) and the implementation it
 * mirrors (src/libNEC.cpp:<line>).
 *
 * Supported by the GPU path: straight and tapered wires (GW), moves and
 * copies (GM, in decks and through nec_gm_card), voltage sources (EX 0),
 * frequency sweeps (FR), the integration-distance card (KH), the extended
 * thin-wire kernel (EK), loads (LD types 0-5), networks and transmission
 * lines (NT, TL), perfect and reflection-coefficient ground (GN 1, GN 0),
 * near electric and magnetic fields (NE, NH) in free space and over those
 * grounds, execute (XQ), far-field patterns (RP mode 0) and all gain /
 * impedance getters.  Card functions for features outside that path
 * (patches, Sommerfeld ground, radial screens, coupling, ...) exist with the
 * reference's signature and return 1 with the message "<card>: not
 * supported by the GPU path".
 */
#pragma once

#ifdef __cplusplus
namespace necgpu { class Context; }
typedef necgpu::Context nec_context;
extern "C" {
#else
struct nec_context;
typedef struct nec_context nec_context;
#endif

/* src/libnecpp.h:60, src/libNEC.cpp:33-38 */
nec_context* nec_create(void);
/* src/libnecpp.h:65, src/libNEC.cpp:40-42 */
long nec_delete(nec_context* in_context);

/* src/libnecpp.h:94-97, src/libNEC.cpp:48-52 -> c_geometry::wire */
long nec_wire(nec_context* in_context, int tag_id, int segment_count,
              double xw1, double yw1, double zw1, double xw2, double yw2, double zw2,
              double rad, double rdel, double rrad);
/* src/libnecpp.h:116 (SP card; not on the GPU path) */
long nec_sp_card(nec_context* in_context, int ns, double x1, double y1, double z1,
                 double x2, double y2, double z2);
/* src/libnecpp.h:132 (SC card; not on the GPU path) */
long nec_sc_card(nec_context* in_context, int i2, double x3, double y3, double z3,
                 double x4, double y4, double z4);
/* src/libnecpp.h:154, src/libNEC.cpp:66-74 -> c_geometry::move
 * (src/c_geometry.cpp:830-900): rotations in degrees about x, y, z, then the
 * shift, of the segments from tag its on; nrpt copies with tags + itsi */
long nec_gm_card(nec_context* in_context, int itsi, int nrpt, double rox, double roy, double roz,
                 double xs, double ys, double zs, int its);
/* src/libnecpp.h:190 (GX card; not on the GPU path) */
long nec_gx_card(nec_context* in_context, int i1, int i2);
/* src/libnecpp.h:201, src/libNEC.cpp:81-83 -> geometry_complete + calc_prepare */
long nec_geometry_complete(nec_context* in_context, int gpflag);

/* src/libnecpp.h:214 */
long nec_benchmark(void);
/* src/libnecpp.h:220, src/libNEC.cpp:86-88 */
const char* nec_error_message(void);

/* src/libnecpp.h:240 (MP card; not on the GPU path) */
long nec_medium_parameters(nec_context* in_context, double permittivity, double permeability);
/* src/libnecpp.h:264 (GN card: iperf -1 free space, 1 perfect ground, 0 finite ground by reflection
   coefficients without radial wires; 2 (Sommerfeld) and radial-wire screens return 1) */
long nec_gn_card(nec_context* in_context, int iperf, int nradl, double epse, double sig,
                 double tmp3, double tmp4, double tmp5, double tmp6);
/* src/libnecpp.h:274, src/libNEC.cpp:166-168 -> nec_context::fr_card */
long nec_fr_card(nec_context* in_context, int in_ifrq, int in_nfrq, double in_freq_mhz, double in_del_freq);
/* src/libnecpp.h:282, src/libNEC.cpp:262-265 */
long nec_ek_card(nec_context* in_context, int itmp1);
/* src/libnecpp.h:303, src/libNEC.cpp:171-174 -> nec_context::ld_card (types 0-5; the loads enter the fill and the power budget) */
long nec_ld_card(nec_context* in_context, int ldtyp, int ldtag, int ldtagf, int ldtagt,
                 double tmp1, double tmp2, double tmp3);
/* src/libnecpp.h:330, src/libNEC.cpp:182-185 -> nec_context::ex_card (type 0 only) */
long nec_ex_card(nec_context* in_context, int extype, int i2, int i3, int i4,
                 double tmp1, double tmp2, double tmp3, double tmp4, double tmp5, double tmp6);
/* src/libnecpp.h:349, src/libNEC.cpp:191-193 */
long nec_excitation_voltage(nec_context* in_context, int tag, int segment, double v_real, double v_imag);
/* src/libnecpp.h:365 (current source; not on the GPU path) */
long nec_excitation_current(nec_context* in_context, double x, double y, double z, double a,
                            double beta, double moment);
/* src/libnecpp.h:384 (plane wave; not on the GPU path) */
long nec_excitation_planewave(nec_context* in_context, int n_theta, int n_phi, double theta, double phi,
                              double eta, double dtheta, double dphi, double pol_ratio);
/* src/libnecpp.h:387-388 -> nec_context::tl_card / nt_card (src/nec_context.cpp:630-726); the network equations
 * (netwk, :4496-5045) are solved on the host around GPU structure solves */
long nec_tl_card(nec_context* in_context, int itmp1, int itmp2, int itmp3, int itmp4, double tmp1,
                 double tmp2, double tmp3, double tmp4, double tmp5, double tmp6);
long nec_nt_card(nec_context* in_context, int itmp1, int itmp2, int itmp3, int itmp4, double tmp1,
                 double tmp2, double tmp3, double tmp4, double tmp5, double tmp6);
/* src/libnecpp.h:406, src/libNEC.cpp:211-214 -> nec_context::xq_card */
long nec_xq_card(nec_context* in_context, int itmp1);
/* src/libnecpp.h:408 (GD card; not on the GPU path) */
long nec_gd_card(nec_context* in_context, double tmp1, double tmp2, double tmp3, double tmp4);
/* src/libnecpp.h:483-487, src/libNEC.cpp:222-232 -> nec_context::rp_card (calc_mode 0) */
long nec_rp_card(nec_context* in_context, int calc_mode, int n_theta, int n_phi,
                 int output_format, int normalization, int D, int A,
                 double theta0, double phi0, double delta_theta, double delta_phi,
                 double radial_distance, double gain_norm);
/* src/libnecpp.h:505-507 (print control: accepted, no printed output exists here) */
long nec_pt_card(nec_context* in_context, int itmp1, int itmp2, int itmp3, int itmp4);
long nec_pq_card(nec_context* in_context, int itmp1, int itmp2, int itmp3, int itmp4);
/* src/libnecpp.h:509, src/libNEC.cpp:250-253 -> nec_context::kh_card */
long nec_kh_card(nec_context* in_context, double tmp1);
/* src/libnecpp.h:510-511 -> nec_context::ne_card / nh_card (both run on the GPU path for wires in free space or
 * over a ground without radial wires; results through nec_gpu_near_field) */
long nec_ne_card(nec_context* in_context, int itmp1, int itmp2, int itmp3, int itmp4, double tmp1,
                 double tmp2, double tmp3, double tmp4, double tmp5, double tmp6);
long nec_nh_card(nec_context* in_context, int itmp1, int itmp2, int itmp3, int itmp4, double tmp1,
                 double tmp2, double tmp3, double tmp4, double tmp5, double tmp6);
/* src/libnecpp.h:514-515 (CP / PL; not on the GPU path) */
long nec_cp_card(nec_context* in_context, int itmp1, int itmp2, int itmp3, int itmp4);
long nec_pl_card(nec_context* in_context, char* ploutput_filename, int itmp1, int itmp2, int itmp3, int itmp4);

/* Gain getters: src/libnecpp.h:537-589, src/libNEC.cpp:90-145,
 * nec_context.h:130-214.  freq_index is the ordinal of the pattern result. */
double nec_gain(nec_context* in_context, int freq_index, int theta_index, int phi_index);
double nec_gain_max(nec_context* in_context, int freq_index);
double nec_gain_min(nec_context* in_context, int freq_index);
double nec_gain_mean(nec_context* in_context, int freq_index);
double nec_gain_sd(nec_context* in_context, int freq_index);
double nec_gain_rhcp_max(nec_context* in_context, int freq_index);
double nec_gain_rhcp_min(nec_context* in_context, int freq_index);
double nec_gain_rhcp_mean(nec_context* in_context, int freq_index);
double nec_gain_rhcp_sd(nec_context* in_context, int freq_index);
double nec_gain_lhcp_max(nec_context* in_context, int freq_index);
double nec_gain_lhcp_min(nec_context* in_context, int freq_index);
double nec_gain_lhcp_mean(nec_context* in_context, int freq_index);
double nec_gain_lhcp_sd(nec_context* in_context, int freq_index);

/* Impedance getters: src/libnecpp.h:593-606, src/libNEC.cpp:147-163,
 * nec_context.h:242-258 (feed_index 0 = first EX card at that frequency). */
double nec_impedance_real(nec_context* in_context, int freq_index);
double nec_impedance_imag(nec_context* in_context, int freq_index);
double nec_impedance_real_indexed(nec_context* in_context, int freq_index, int feed_index);
double nec_impedance_imag_indexed(nec_context* in_context, int freq_index, int feed_index);

#ifdef __cplusplus
}
#endif
