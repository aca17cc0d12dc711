// batch.h -- batched solves of many small structures (see batch_host.cpp).
#pragma once

#include <vector>

#include "adjoint.h"

namespace necgpu {

struct BatchInput {
  int n_designs = 0;
  const int* wire_ptr = nullptr;      // n_designs + 1
  const int* wire_tag = nullptr;      // per wire
  const int* wire_ns = nullptr;       // per wire
  const double* wire_geom = nullptr;  // 7 per wire: x1 y1 z1 x2 y2 z2 radius (metres)
  const int* src_tag = nullptr;       // per design: EX card tag
  const int* src_seg = nullptr;       // per design: EX card segment (1-based within the tag)
  const double* src_volt = nullptr;   // 2 per design
  int n_freq = 0;
  const double* freq_mhz = nullptr;   // n_designs * n_freq
  int n_dir = 0;
  const double* theta_deg = nullptr;  // n_dir
  const double* phi_deg = nullptr;    // n_dir
  double rkh = 1.0;                   // KH card distance, wavelengths
  int panel_width = 0;                // 8, 16 or 32; anything else: autotune
  int repeats = 0;                    // graph replays to time (0: none)
};

struct BatchOutput {
  int solves = 0;
  long long segments = 0;
  std::vector<double> zin;   // 2 per solve (design-major, then frequency)
  std::vector<double> pin;   // input power per solve
  std::vector<double> gain;  // n_dir per solve, total power gain (dB)
  double host_ms = 0, fill_ms = 0, lu_ms = 0, solve_pattern_ms = 0, graph_ms = 0;
  int panel_width = 0;
  double autotune_ms[3] = {0, 0, 0};  // panel widths 8, 16, 32
  int graph_matches_eager = -1;       // 1: replayed results equal the eager pass bit for bit
};

void run_batch(const BatchInput& in, BatchOutput* out);

// Adjoint gradients of every (design, frequency) of a batch: one batched
// forward pass (fill, LU, solve), the objective and adjoint right-hand side
// of each solve on the host's cores, batched transposed solves and fill
// contraction on the device, then each solve's gradient on the host.
struct BatchAdjointOutput {
  int solves = 0;
  long long segments = 0;
  std::vector<double> J, gain_db, zin, dfreq;  // per solve (zin: 2)
  std::vector<double> grad;                    // 7 per segment, solves in order
  double host_prep_ms = 0, fill_ms = 0, lu_ms = 0, solve_ms = 0, front_ms = 0, adjoint_device_ms = 0, back_ms = 0,
         total_ms = 0;
};
void run_batch_adjoint(const BatchInput& in, const AdjointSpec& spec, BatchAdjointOutput* out);

}  // namespace necgpu
