// api.cpp -- the extern "C" boundary.
//
// Same conventions as nec2++'s libNEC.cpp (src/libNEC.cpp:27-282): card
// functions catch the engine's exceptions, keep the message in one
// process-global string and return 1; getters return -999.0 when the result
// does not exist.
#include <cstring>
#include <string>

#include "../../include/libnecpp.h"
#include "../../include/necpp_gpu.h"
#include "batch.h"
#include "context.h"

using necgpu::Context;
using necgpu::NecError;

namespace {

std::string g_err_message;

template <class F>
long guarded(F&& f) {
  try {
    f();
  } catch (const NecError& e) {
    g_err_message = e.what();
    return 1;
  } catch (const std::exception& e) {
    g_err_message = e.what();
    return 1;
  }
  return 0;
}

long unsupported(const char* card) {
  g_err_message = std::string(card) + ": not supported by the GPU path";
  return 1;
}

const necgpu::PatternResult* pat(nec_context* c, int k) { return c ? c->pattern(k) : nullptr; }

}  // namespace

#pragma GCC visibility push(default)
extern "C" {

nec_context* nec_create(void) {
  try {
    return new Context();
  } catch (const std::exception& e) {
    g_err_message = e.what();
    return nullptr;
  }
}

long nec_delete(nec_context* ctx) {
  return guarded([&] { delete ctx; });
}

long nec_benchmark(void) { return 0; }

long nec_wire(nec_context* ctx, int tag_id, int segment_count, double xw1, double yw1, double zw1, double xw2,
              double yw2, double zw2, double rad, double rdel, double rrad) {
  return guarded([&] { ctx->wire(tag_id, segment_count, xw1, yw1, zw1, xw2, yw2, zw2, rad, rdel, rrad); });
}

long nec_sp_card(nec_context*, int, double, double, double, double, double, double) { return unsupported("SP"); }
long nec_sc_card(nec_context*, int, double, double, double, double, double, double) { return unsupported("SC"); }
long nec_gm_card(nec_context* ctx, int itsi, int nrpt, double rox, double roy, double roz, double xs, double ys,
                 double zs, int its) {
  // src/libNEC.cpp:66-74: the angles arrive in degrees
  return guarded([&] {
    ctx->move(necgpu::deg_to_rad(rox), necgpu::deg_to_rad(roy), necgpu::deg_to_rad(roz), xs, ys, zs, its, nrpt, itsi);
  });
}
long nec_gx_card(nec_context*, int, int) { return unsupported("GX"); }

long nec_geometry_complete(nec_context* ctx, int gpflag) {
  return guarded([&] { ctx->geometry_complete(gpflag); });
}

const char* nec_error_message(void) { return g_err_message.c_str(); }

long nec_medium_parameters(nec_context*, double, double) { return unsupported("MP"); }

long nec_gn_card(nec_context* ctx, int iperf, int nradl, double epsr, double sig, double tmp3, double tmp4,
                 double tmp5, double tmp6) {
  return guarded([&] { ctx->gn_card(iperf, nradl, epsr, sig, tmp3, tmp4, tmp5, tmp6); });
}

long nec_fr_card(nec_context* ctx, int in_ifrq, int in_nfrq, double in_freq_mhz, double in_del_freq) {
  return guarded([&] { ctx->fr_card(in_ifrq, in_nfrq, in_freq_mhz, in_del_freq); });
}

long nec_ek_card(nec_context* ctx, int itmp1) {
  return guarded([&] { ctx->ek_card(itmp1 != -1); });
}

long nec_ld_card(nec_context* ctx, int ldtyp, int ldtag, int ldtagf, int ldtagt, double tmp1, double tmp2,
                 double tmp3) {
  return guarded([&] { ctx->ld_card(ldtyp, ldtag, ldtagf, ldtagt, tmp1, tmp2, tmp3); });
}

long nec_ex_card(nec_context* ctx, int extype, int i2, int i3, int i4, double tmp1, double tmp2, double tmp3,
                 double tmp4, double tmp5, double tmp6) {
  return guarded([&] { ctx->ex_card(extype, i2, i3, i4, tmp1, tmp2, tmp3, tmp4, tmp5, tmp6); });
}

long nec_excitation_voltage(nec_context* ctx, int tag, int segment, double v_real, double v_imag) {
  return guarded([&] { ctx->ex_card(0, tag, segment, 0, v_real, v_imag, 0.0, 0.0, 0.0, 0.0); });
}

long nec_excitation_current(nec_context*, double, double, double, double, double, double) {
  return unsupported("EX 4");
}

long nec_excitation_planewave(nec_context*, int, int, double, double, double, double, double, double) {
  return unsupported("EX 1");
}

long nec_tl_card(nec_context* ctx, int itmp1, int itmp2, int itmp3, int itmp4, double tmp1, double tmp2, double tmp3,
                 double tmp4, double tmp5, double tmp6) {
  return guarded([&] { ctx->tl_card(itmp1, itmp2, itmp3, itmp4, tmp1, tmp2, tmp3, tmp4, tmp5, tmp6); });
}
long nec_nt_card(nec_context* ctx, int itmp1, int itmp2, int itmp3, int itmp4, double tmp1, double tmp2, double tmp3,
                 double tmp4, double tmp5, double tmp6) {
  return guarded([&] { ctx->nt_card(itmp1, itmp2, itmp3, itmp4, tmp1, tmp2, tmp3, tmp4, tmp5, tmp6); });
}

long nec_xq_card(nec_context* ctx, int itmp1) {
  return guarded([&] { ctx->xq_card(itmp1); });
}

long nec_gd_card(nec_context*, double, double, double, double) { return unsupported("GD"); }

long nec_rp_card(nec_context* ctx, int calc_mode, int n_theta, int n_phi, int output_format, int normalization, int D,
                 int A, double theta0, double phi0, double delta_theta, double delta_phi, double radial_distance,
                 double gain_norm) {
  return guarded([&] {
    ctx->rp_card(calc_mode, n_theta, n_phi, output_format, normalization, D, A, theta0, phi0, delta_theta, delta_phi,
                 radial_distance, gain_norm);
  });
}

long nec_pt_card(nec_context*, int, int, int, int) { return 0; }
long nec_pq_card(nec_context*, int, int, int, int) { return 0; }

long nec_kh_card(nec_context* ctx, double tmp1) {
  return guarded([&] { ctx->kh_card(tmp1); });
}

long nec_ne_card(nec_context* ctx, int itmp1, int itmp2, int itmp3, int itmp4, double tmp1, double tmp2, double tmp3,
                 double tmp4, double tmp5, double tmp6) {
  return guarded([&] { ctx->ne_card(itmp1, itmp2, itmp3, itmp4, tmp1, tmp2, tmp3, tmp4, tmp5, tmp6); });
}
long nec_nh_card(nec_context* ctx, int itmp1, int itmp2, int itmp3, int itmp4, double tmp1, double tmp2, double tmp3,
                 double tmp4, double tmp5, double tmp6) {
  return guarded([&] { ctx->nh_card(itmp1, itmp2, itmp3, itmp4, tmp1, tmp2, tmp3, tmp4, tmp5, tmp6); });
}
long nec_cp_card(nec_context*, int, int, int, int) { return unsupported("CP"); }
long nec_pl_card(nec_context*, char*, int, int, int, int) { return unsupported("PL"); }

double nec_gain(nec_context* ctx, int freq_index, int theta_index, int phi_index) {
  const auto* p = pat(ctx, freq_index);
  if (!p) return -999.0;
  if (theta_index < 0 || theta_index >= p->n_theta || phi_index < 0 || phi_index >= p->n_phi) return -999.0;
  return p->a.gain[(size_t)phi_index * p->n_theta + theta_index];
}

#define NEC_STAT_GETTER(name, idx)                      \
  double name(nec_context* ctx, int freq_index) {       \
    const auto* p = pat(ctx, freq_index);               \
    return p ? p->stats[idx] : -999.0;                  \
  }
NEC_STAT_GETTER(nec_gain_max, 0)
NEC_STAT_GETTER(nec_gain_min, 1)
NEC_STAT_GETTER(nec_gain_mean, 2)
NEC_STAT_GETTER(nec_gain_sd, 3)
NEC_STAT_GETTER(nec_gain_rhcp_max, 4)
NEC_STAT_GETTER(nec_gain_rhcp_min, 5)
NEC_STAT_GETTER(nec_gain_rhcp_mean, 6)
NEC_STAT_GETTER(nec_gain_rhcp_sd, 7)
NEC_STAT_GETTER(nec_gain_lhcp_max, 8)
NEC_STAT_GETTER(nec_gain_lhcp_min, 9)
NEC_STAT_GETTER(nec_gain_lhcp_mean, 10)
NEC_STAT_GETTER(nec_gain_lhcp_sd, 11)
#undef NEC_STAT_GETTER

double nec_impedance_real_indexed(nec_context* ctx, int freq_index, int feed_index) {
  const auto* in = ctx ? ctx->input(freq_index) : nullptr;
  if (!in || feed_index < 0 || (size_t)feed_index >= in->impedance.size()) return -999.0;
  return in->impedance[feed_index].re;
}

double nec_impedance_imag_indexed(nec_context* ctx, int freq_index, int feed_index) {
  const auto* in = ctx ? ctx->input(freq_index) : nullptr;
  if (!in || feed_index < 0 || (size_t)feed_index >= in->impedance.size()) return -999.0;
  return in->impedance[feed_index].im;
}

double nec_impedance_real(nec_context* ctx, int freq_index) { return nec_impedance_real_indexed(ctx, freq_index, 0); }
double nec_impedance_imag(nec_context* ctx, int freq_index) { return nec_impedance_imag_indexed(ctx, freq_index, 0); }

/* ---------------------------- GPU extras ---------------------------- */

long nec_gpu_info(char* out, int len) {
  return guarded([&] {
    hipDeviceProp_t prop;
    int dev = 0;
    std::string s;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      s = std::string(prop.gcnArchName) + " " + prop.name;
    else
      s = "no-device";
    s += " necpp_amd 0.1";
    if (out && len > 0) {
      std::strncpy(out, s.c_str(), (size_t)len - 1);
      out[len - 1] = 0;
    }
  });
}

long nec_gpu_segment_count(nec_context* ctx) {
  if (!ctx || !ctx->geometry().is_complete()) return -1;
  return static_cast<long>(ctx->geometry().size());
}

long nec_gpu_segments(nec_context* ctx, double* out, int* conn, long n) {
  return guarded([&] {
    const auto& g = ctx->geometry();
    if (!g.is_complete()) throw NecError("geometry is not complete");
    if (n != g.size()) throw NecError("n does not match the segment count");
    for (long i = 0; i < n; ++i) {
      if (out) {
        double* o = out + 8 * i;
        o[0] = g.x[i];
        o[1] = g.y[i];
        o[2] = g.z[i];
        o[3] = g.len[i];
        o[4] = g.rad[i];
        o[5] = g.cab[i];
        o[6] = g.sab[i];
        o[7] = g.salp[i];
      }
      if (conn) {
        conn[i] = g.icon1[i];
        conn[n + i] = g.icon2[i];
        conn[2 * n + i] = g.tags[i];
      }
    }
  });
}

long nec_gpu_batch_solve(const nec_gpu_batch* b, double* zin, double* gain, double* stats) {
  return guarded([&] {
    if (!b) throw NecError("nec_gpu_batch_solve: no batch");
    necgpu::BatchInput in;
    in.n_designs = b->n_designs;
    in.wire_ptr = b->wire_ptr;
    in.wire_tag = b->wire_tag;
    in.wire_ns = b->wire_ns;
    in.wire_geom = b->wire_geom;
    in.src_tag = b->src_tag;
    in.src_seg = b->src_seg;
    in.src_volt = b->src_volt;
    in.n_freq = b->n_freq;
    in.freq_mhz = b->freq_mhz;
    in.n_dir = b->n_dir;
    in.theta_deg = b->theta_deg;
    in.phi_deg = b->phi_deg;
    in.rkh = b->rkh;
    in.panel_width = b->panel_width;
    in.repeats = b->repeats;
    necgpu::BatchOutput out;
    necgpu::run_batch(in, &out);
    std::copy(out.zin.begin(), out.zin.end(), zin);
    if (gain && !out.gain.empty()) std::copy(out.gain.begin(), out.gain.end(), gain);
    if (stats) {
      const double v[12] = {(double)out.solves, (double)out.segments, out.host_ms, out.fill_ms, out.lu_ms,
                            out.solve_pattern_ms, out.graph_ms, (double)out.panel_width, out.autotune_ms[0],
                            out.autotune_ms[1], out.autotune_ms[2], (double)out.graph_matches_eager};
      std::copy(v, v + 12, stats);
    }
  });
}

long nec_gpu_set_refine_tolerance(nec_context* ctx, double tol) {
  return guarded([&] {
    if (!(tol > 0.0)) throw NecError("refinement tolerance must be positive");
    ctx->device().set_refine_tol(tol);
  });
}

long nec_gpu_adjoint_gradient(nec_context* ctx, double theta_deg, double phi_deg, double weight, double z0_re,
                              double z0_im, double* J_out, double* grad_out, double* dfreq_out, double* info) {
  return guarded([&] {
    if (!ctx) throw NecError("nec_gpu_adjoint_gradient: no context");
    necgpu::AdjointSpec spec;
    spec.theta_deg = theta_deg;
    spec.phi_deg = phi_deg;
    spec.weight = weight;
    spec.z0 = necgpu::zmk(z0_re, z0_im);
    necgpu::AdjointResult r;
    std::vector<double> g7;
    double df = 0.0;
    ctx->adjoint(spec, &r, &g7, &df);
    if (J_out) *J_out = r.J;
    if (grad_out) std::copy(g7.begin(), g7.end(), grad_out);
    if (dfreq_out) *dfreq_out = df;
    if (info) {
      const double v[6] = {r.gain_db, r.zin.re, r.zin.im, r.host_ms, r.device_ms, r.dlambda};
      std::copy(v, v + 6, info);
    }
  });
}

long nec_gpu_batch_adjoint(const nec_gpu_batch* b, double theta_deg, double phi_deg, double weight, double z0_re,
                           double z0_im, double* J, double* grad, double* dfreq, double* info, double* stats) {
  return guarded([&] {
    if (!b) throw NecError("nec_gpu_batch_adjoint: no batch");
    necgpu::BatchInput in;
    in.n_designs = b->n_designs;
    in.wire_ptr = b->wire_ptr;
    in.wire_tag = b->wire_tag;
    in.wire_ns = b->wire_ns;
    in.wire_geom = b->wire_geom;
    in.src_tag = b->src_tag;
    in.src_seg = b->src_seg;
    in.src_volt = b->src_volt;
    in.n_freq = b->n_freq;
    in.freq_mhz = b->freq_mhz;
    in.rkh = b->rkh;
    in.panel_width = b->panel_width;
    necgpu::AdjointSpec spec;
    spec.theta_deg = theta_deg;
    spec.phi_deg = phi_deg;
    spec.weight = weight;
    spec.z0 = necgpu::zmk(z0_re, z0_im);
    necgpu::BatchAdjointOutput out;
    necgpu::run_batch_adjoint(in, spec, &out);
    if (J) std::copy(out.J.begin(), out.J.end(), J);
    if (grad) std::copy(out.grad.begin(), out.grad.end(), grad);
    if (dfreq) std::copy(out.dfreq.begin(), out.dfreq.end(), dfreq);
    if (info)
      for (int s = 0; s < out.solves; ++s) {
        info[3 * s] = out.gain_db[s];
        info[3 * s + 1] = out.zin[2 * s];
        info[3 * s + 2] = out.zin[2 * s + 1];
      }
    if (stats) {
      const double v[10] = {(double)out.solves, (double)out.segments, out.host_prep_ms, out.fill_ms, out.lu_ms,
                            out.solve_ms, out.front_ms, out.adjoint_device_ms, out.back_ms, out.total_ms};
      std::copy(v, v + 10, stats);
    }
  });
}

long nec_gpu_refinement(nec_context* ctx, int* steps, double* rel_residual) {
  return guarded([&] {
    *steps = ctx->device().refine_steps();
    *rel_residual = ctx->device().refine_residual();
  });
}

long nec_gpu_factor_fallbacks(nec_context* ctx, long* lu, long* pivot) {
  return guarded([&] {
    if (lu) *lu = ctx->device().lu_fallbacks();
    if (pivot) *pivot = ctx->device().refine_fallbacks();
  });
}

long nec_gpu_set_precision(nec_context* ctx, int mode) {
  return guarded([&] {
    if (mode != 0 && mode != 1) throw NecError("precision mode must be 0 (FP64) or 1 (mixed)");
    ctx->set_precision(mode);
  });
}

long nec_gpu_currents(nec_context* ctx, int freq_index, double* out, long n) {
  return guarded([&] {
    const auto* in = ctx->input(freq_index);
    if (!in) throw NecError("no antenna input result " + std::to_string(freq_index));
    if (n < (long)in->segment_currents.size()) throw NecError("output too small");
    for (size_t i = 0; i < in->segment_currents.size(); ++i) {
      out[2 * i] = in->segment_currents[i].re;
      out[2 * i + 1] = in->segment_currents[i].im;
    }
  });
}

long nec_gpu_stage_times(nec_context* ctx, double* out) {
  return guarded([&] {
    const auto& t = ctx->last_times();
    out[0] = t.fill_ms;
    out[1] = t.factor_ms;
    out[2] = t.solve_ms;
    out[3] = t.pattern_ms;
  });
}

long nec_gpu_near_field(nec_context* ctx, int index, double* out, long max_points) {
  if (!ctx) return -1;
  const Context::NearFieldResult* r = ctx->near_field(index);
  if (!r) return -1;
  const long np = (long)(r->xyz.size() / 3);
  if (out)
    for (long p = 0; p < np && p < max_points; ++p) {
      double* o = out + 9 * p;
      for (int c = 0; c < 3; ++c) o[c] = r->xyz[3 * p + c];
      for (int c = 0; c < 3; ++c) {
        o[3 + 2 * c] = r->e[3 * p + c].re;
        o[4 + 2 * c] = r->e[3 * p + c].im;
      }
    }
  return np;
}

long nec_gpu_near_field_kind(nec_context* ctx, int index) {
  if (!ctx) return -1;
  const Context::NearFieldResult* r = ctx->near_field(index);
  if (!r) return -1;
  return r->magnetic ? 1 : 0;
}

long nec_gpu_deck(nec_context* ctx, const char* deck) {
  return guarded([&] { necgpu::run_deck(ctx, deck); });
}

long nec_gpu_set_kernel_timing(nec_context* ctx, int on) {
  return guarded([&] { ctx->device().set_kernel_timing(on != 0); });
}

long nec_gpu_kernel_times(nec_context* ctx, double* out, long n) {
  long rc = -1;
  const long ok = guarded([&] {
    if (n < 2L * necgpu::K_NKINDS) throw NecError("nec_gpu_kernel_times: out too short");
    const necgpu::KernelClock& c = ctx->device().kernel_clock();
    for (int k = 0; k < necgpu::K_NKINDS; ++k) {
      out[2 * k] = (double)c.launches[k];
      out[2 * k + 1] = c.total_ms[k];
    }
    rc = necgpu::K_NKINDS;
  });
  return ok == 0 ? rc : -1;
}

const char* nec_gpu_kernel_name(int kind) { return necgpu::kernel_kind_name(kind); }

long nec_gpu_benchmark_solve(nec_context* ctx, int warmup, int steps, double* out) {
  return guarded([&] {
    if (warmup < 0 || steps < 1) throw NecError("benchmark: need warmup >= 0 and steps >= 1");
    ctx->benchmark_solve(warmup, steps, out);
  });
}

long nec_gpu_fill_matrix(nec_context* ctx, double freq_mhz, double* out, long n) {
  return guarded([&] {
    if (!ctx->geometry().is_complete()) throw NecError("geometry is not complete");
    if (n != ctx->geometry().size()) throw NecError("n does not match the segment count");
    std::vector<double> Z;
    ctx->fill_matrix(freq_mhz, &Z);
    std::memcpy(out, Z.data(), Z.size() * sizeof(double));
  });
}

long nec_gpu_pattern_array(nec_context* ctx, int freq_index, int which, double* out, long len) {
  return guarded([&] {
    const auto* p = ctx->pattern(freq_index);
    if (!p) throw NecError("no pattern result " + std::to_string(freq_index));
    const auto& a = p->a;
    const std::vector<double>* src = nullptr;
    switch (which) {
      case 0: src = &a.gain; break;
      case 1: src = &a.pg_vert; break;
      case 2: src = &a.pg_horiz; break;
      case 3: src = &a.pg_tot; break;
      case 4: src = &a.pg_rhcp; break;
      case 5: src = &a.pg_lhcp; break;
      case 6: src = &a.axial; break;
      case 7: src = &a.tilt; break;
      case 8: {
        if (len < (long)a.sense.size()) throw NecError("output too small");
        for (size_t i = 0; i < a.sense.size(); ++i) out[i] = a.sense[i];
        return;
      }
      case 9: src = &a.avg_scale; break;
      case 10: src = &a.e_theta; break;
      case 11: src = &a.e_phi; break;
      default: throw NecError("unknown pattern array " + std::to_string(which));
    }
    if (len < (long)src->size()) throw NecError("output too small");
    std::memcpy(out, src->data(), src->size() * sizeof(double));
  });
}

long nec_gpu_pattern_average(nec_context* ctx, int freq_index, double* avg_gain, double* solid_angle) {
  return guarded([&] {
    const auto* p = ctx->pattern(freq_index);
    if (!p) throw NecError("no pattern result " + std::to_string(freq_index));
    *avg_gain = p->average_power_gain;
    *solid_angle = p->average_power_solid_angle;
  });
}

long long necgpu_zgetrf_workspace(int n) { return static_cast<long long>(necgpu::lu_workspace_bytes(n)); }

long necgpu_zgetrf(double* a, long long lda, int n, int* piv, int* info, void* work, void* stream) {
  return guarded([&] {
    const int rc = necgpu::launch_lu_fp64(a, lda, n, piv, info, work, static_cast<hipStream_t>(stream));
    if (rc == 2) throw NecError("necgpu_zgetrf: n too large (the panel needs one CU per 112 rows, at most 256)");
    if (rc) necgpu::hip_check(hipGetLastError(), "necgpu_zgetrf");
  });
}

long necgpu_zgetrf_noalloc(double* a, long long lda, int n, int* piv, int* info, void* stream) {
  return guarded([&] {
    const int rc = necgpu::launch_lu_fp64_noalloc(a, lda, n, piv, info, static_cast<hipStream_t>(stream));
    if (rc == 3)
      throw NecError("necgpu_zgetrf_noalloc: n = " + std::to_string(n) + " exceeds the largest order of its fixed " +
                     "workspace (" + std::to_string(necgpu::kNoAllocMaxN) + ")");
    if (rc == 4) throw NecError("necgpu_zgetrf_noalloc: the device address of the fixed workspace is unavailable");
    if (rc == 2) throw NecError("necgpu_zgetrf_noalloc: n too large for the panel");
    if (rc) necgpu::hip_check(hipGetLastError(), "necgpu_zgetrf_noalloc");
  });
}

long necgpu_cgetrf(float* a, long long lda, int n, int* piv, int* info, void* work, int pivot_mode, void* stream) {
  return guarded([&] {
    if (pivot_mode != 0 && pivot_mode != 1) throw NecError("necgpu_cgetrf: pivot_mode must be 0 or 1");
    const int rc = necgpu::launch_lu_fp32(a, lda, n, piv, info, work, static_cast<hipStream_t>(stream), pivot_mode);
    if (rc == 2) throw NecError("necgpu_cgetrf: n too large (the panel needs one CU per 224 rows, at most 256)");
    if (rc) necgpu::hip_check(hipGetLastError(), "necgpu_cgetrf");
  });
}

long necgpu_check_mfma_f64(const double* a, const double* b, double* d, void* stream) {
  return guarded([&] {
    if (necgpu::launch_mfma_probe(a, b, d, static_cast<hipStream_t>(stream)))
      necgpu::hip_check(hipGetLastError(), "necgpu_check_mfma_f64");
  });
}

long necgpu_check_mfma_f32(const float* a, const float* b, float* d, void* stream) {
  return guarded([&] {
    if (necgpu::launch_mfma_probe_f32(a, b, d, static_cast<hipStream_t>(stream)))
      necgpu::hip_check(hipGetLastError(), "necgpu_check_mfma_f32");
  });
}

long long necgpu_zgetrs_workspace(int n) { return static_cast<long long>(necgpu::solve_workspace_bytes(n)); }

long necgpu_zgetrs(const double* lu, long long lda, int n, const int* piv, double* x, void* work, unsigned epoch,
                   void* stream) {
  (void)epoch;  // kept in the signature; the sweeps clear what they poll
  return guarded([&] {
    if (necgpu::launch_lu_solve_fp64(lu, lda, n, piv, x, work, static_cast<hipStream_t>(stream)))
      necgpu::hip_check(hipGetLastError(), "necgpu_zgetrs");
  });
}

}  // extern "C"
#pragma GCC visibility pop
