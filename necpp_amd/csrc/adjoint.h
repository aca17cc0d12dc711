// adjoint.h -- gradient of a design objective by the adjoint method.
#pragma once

#include <vector>

#include "engine.h"
#include "geometry.h"
#include "nec_math.h"

namespace necgpu {

// J = -G(theta, phi) + weight * |Z_in - z0|^2, with G the total power gain
// in dBi of the pattern in direction (theta, phi) and Z_in the input
// impedance of the single voltage source.
struct AdjointSpec {
  double theta_deg = 90.0, phi_deg = 0.0;
  double weight = 0.01;
  zc z0 = {50.0, 0.0};
};

struct AdjointResult {
  double J = 0.0;
  double gain_db = 0.0;
  zc zin = {0.0, 0.0};
  // dJ per segment, in wavelength units: centre x, y, z, length, radius,
  // direction cosines cab, sab, salp (the eight segment variables)
  std::vector<double> dseg;
  // dJ/d(wavelength) with the wavelength-scaled geometry held fixed
  double dlambda = 0.0;
  double host_ms = 0.0, device_ms = 0.0;
};

// The objective's value and the adjoint right-hand side g of one solve
// (current coefficients abc: 6n doubles), and what the back half reuses.
struct AdjointFront {
  double J = 0.0, gain_db = 0.0, dJdS = 0.0;
  zc zin = {0.0, 0.0}, usrc = {0.0, 0.0}, dJdI = {0.0, 0.0}, eth = {0.0, 0.0}, eph = {0.0, 0.0};
  std::vector<zc> dU;  // dJ/dU_j, 3 per segment
  std::vector<zc> g;   // adjoint right-hand side, n
};
void adjoint_front(const ScaledSegments& seg, const BasisTables& t, int src, zc volt, double wavelength,
                   const double* abc, const AdjointSpec& spec, AdjointFront* f);
// The gradient from the device's terms: lam (Z^T lam = g), x (the solution),
// gi (7 per segment) and gj (13 per segment) of the fill contraction.
void adjoint_back(const Geometry& geo, const ScaledSegments& seg, int src, zc volt, double wavelength,
                  const double* abc, const AdjointSpec& spec, const AdjointFront& f, const zc* lam, const zc* x,
                  const double* gi, const double* gj, AdjointResult* out);

// The last solve of `dev` must be the FP64 solve of this geometry at this
// wavelength with one voltage source `volt` on segment `src` (0-based) and
// current coefficients `abc` (6n: A, B, C real and imaginary parts).
void adjoint_gradient(const Geometry& geo, const ScaledSegments& seg, int src, zc volt, double wavelength,
                      double rkh, const std::vector<double>& abc, DeviceSolver& dev, const AdjointSpec& spec,
                      AdjointResult* out);

// Segment endpoint gradient (7 per segment: x1 y1 z1 x2 y2 z2 radius, per
// metre) and dJ/d(frequency in MHz) from an AdjointResult.
void adjoint_endpoints(const ScaledSegments& seg, double wavelength, double freq_mhz, const AdjointResult& r,
                       std::vector<double>* grad7, double* dfreq_mhz);

}  // namespace necgpu
