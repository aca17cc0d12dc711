// kclock.cpp -- per-launch kernel timing with HIP events.
#include "kernels.h"

namespace necgpu {

thread_local KernelClock* g_kernel_clock = nullptr;

const char* kernel_kind_name(int kind) {
  static const char* names[K_NKINDS] = {"fill_fields",   "fill_assemble", "lu_init",    "lu_panel",
                                        "lu_laswp",      "lu_trsm",       "zgemm",      "solve_reset", "diag_inverse",
                                        "trsv_lower",    "trsv_upper",    "currents",   "zin",
                                        "pattern_partial", "pattern_finish", "refine", "residual",
                                        "batch_fill", "batch_lu", "batch_solve", "batch_pattern",
                                        "adjoint_fill", "adjoint_solve", "adjoint_reduce",
                                        "lu_next", "near_field", "batch_tables", "fill_tile"};
  return (kind >= 0 && kind < K_NKINDS) ? names[kind] : "?";
}

KernelClock::~KernelClock() {
  for (hipEvent_t e : ev_) (void)hipEventDestroy(e);
}

bool KernelClock::next(int kind, hipEvent_t* start, hipEvent_t* stop) {
  if (used_ == kind_.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return false;
    if (hipEventCreate(&b) != hipSuccess) {
      (void)hipEventDestroy(a);
      return false;
    }
    ev_.push_back(a);
    ev_.push_back(b);
    kind_.push_back(kind);
  }
  kind_[used_] = kind;
  *start = ev_[2 * used_];
  *stop = ev_[2 * used_ + 1];
  ++used_;
  return true;
}

bool KernelClock::reserve(size_t slots) {
  while (kind_.size() < slots) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return false;
    if (hipEventCreate(&b) != hipSuccess) {
      (void)hipEventDestroy(a);
      return false;
    }
    ev_.push_back(a);
    ev_.push_back(b);
    kind_.push_back(0);
  }
  return true;
}

int KernelClock::collect() {
  int rc = 0;
  auto fail = [&](size_t i, hipError_t e, const char* what) {
    if (rc == 0)
      failure = std::string(what) + " of " + kernel_kind_name(kind_[i]) + " (launch " + std::to_string(i) + " of " +
                std::to_string(used_) + "): " + hipGetErrorString(e);
    rc = 1;
  };
  for (size_t i = 0; i < used_; ++i) {
    hipError_t e = hipEventSynchronize(ev_[2 * i + 1]);
    if (e != hipSuccess) fail(i, e, "hipEventSynchronize");
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, ev_[2 * i], ev_[2 * i + 1]);
    if (e != hipSuccess) {
      fail(i, e, "hipEventElapsedTime");
      continue;
    }
    total_ms[kind_[i]] += ms;
    launches[kind_[i]] += 1;
  }
  used_ = 0;
  return rc;
}

void KernelClock::reset() {
  used_ = 0;
  for (int k = 0; k < K_NKINDS; ++k) {
    total_ms[k] = 0.0;
    launches[k] = 0;
  }
}

}  // namespace necgpu
