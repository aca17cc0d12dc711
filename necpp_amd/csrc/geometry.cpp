// geometry.cpp -- wire geometry of a nec_context (host side).
#include "geometry.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "nec_math.h"

namespace necgpu {

namespace {

// |a - b| in the L1 norm, the contact test of build_connections
// (src/c_geometry.cpp:1374-1378).
inline double manhattan(double ax, double ay, double az, double bx, double by, double bz) {
  return (std::fabs(ax - bx) + std::fabs(ay - by)) + std::fabs(az - bz);
}

std::string seg_msg(const char* head, int64_t seg, const char* tail) {
  return std::string(head) + std::to_string(seg) + tail;
}

}  // namespace

void Geometry::add_wire(int tag, int ns, double xw1, double yw1, double zw1, double xw2, double yw2, double zw2,
                        double rad, double rdel, double rrad) {
  // src/c_geometry.cpp:614-683
  if (ns < 0) throw NecError("GW: negative number of segments");
  if (ns == 0) return;
  complete_ = false;
  double dx = xw2 - xw1, dy = yw2 - yw1, dz = zw2 - zw1;
  double step, ratio;
  if (std::fabs(rdel - 1.0) >= 1.0e-6) {  // tapered wire
    step = std::sqrt((dx * dx + dy * dy) + dz * dz);
    dx /= step;
    dy /= step;
    dz /= step;
    step = step * (1. - rdel) / (1. - std::pow(rdel, ns));
    ratio = rdel;
  } else {
    double fns = ns;
    dx /= fns;
    dy /= fns;
    dz /= fns;
    step = 1.0;
    ratio = 1.0;
  }
  double r = rad, px = xw1, py = yw1, pz = zw1;
  for (int k = 0; k < ns; ++k) {
    double qx = px + dx * step, qy = py + dy * step, qz = pz + dz * step;
    tags.push_back(tag);
    x.push_back(px);
    y.push_back(py);
    z.push_back(pz);
    x2_.push_back(qx);
    y2_.push_back(qy);
    z2_.push_back(qz);
    this->rad.push_back(r);
    step = step * ratio;
    r = r * rrad;
    px = qx;
    py = qy;
    pz = qz;
  }
  n_ += ns;
  x2_.back() = xw2;
  y2_.back() = yw2;
  z2_.back() = zw2;
}

void Geometry::move(double rox, double roy, double roz, double xs, double ys, double zs, int its, int nrpt,
                    int itgi) {
  // c_geometry::move, src/c_geometry.cpp:830-900: the rotation matrix of
  // rotations about x, then y, then z, applied to both end points of every
  // segment from i1 on; copies are appended after the last segment and
  // each copy transforms the one before it
  const double sps = std::sin(rox), cps = std::cos(rox);
  const double sth = std::sin(roy), cth = std::cos(roy);
  const double sph = std::sin(roz), cph = std::cos(roz);
  const double xx = cph * cth, xy = cph * sth * sps - sph * cps, xz = cph * sth * cps + sph * sps;
  const double yx = sph * cth, yy = sph * sth * sps + cph * cps, yz = sph * sth * cps - cph * sps;
  const double zx = -sth, zy = cth * sps, zz = cth * cps;
  if (n_ <= 0) return;
  complete_ = false;
  int64_t i1 = std::max<int64_t>(1, segment_number(its, 1));  // throws for a tag no segment has
  const int nrp = nrpt == 0 ? 1 : nrpt;
  int64_t k = nrpt == 0 ? i1 - 1 : n_;
  if (nrpt != 0) {
    const size_t sz = static_cast<size_t>(n_ + (n_ + 1 - i1) * nrpt);
    x.resize(sz);
    y.resize(sz);
    z.resize(sz);
    x2_.resize(sz);
    y2_.resize(sz);
    z2_.resize(sz);
    rad.resize(sz);
    tags.resize(sz);
  }
  for (int ir = 0; ir < nrp; ++ir) {
    for (int64_t i = i1 - 1; i < n_; ++i, ++k) {
      double xi = x[i], yi = y[i], zi = z[i];
      x[k] = xi * xx + yi * xy + zi * xz + xs;
      y[k] = xi * yx + yi * yy + zi * yz + ys;
      z[k] = xi * zx + yi * zy + zi * zz + zs;
      xi = x2_[i];
      yi = y2_[i];
      zi = z2_[i];
      x2_[k] = xi * xx + yi * xy + zi * xz + xs;
      y2_[k] = xi * yx + yi * yy + zi * yz + ys;
      z2_[k] = xi * zx + yi * zy + zi * zz + zs;
      rad[k] = rad[i];
      tags[k] = tags[i] != 0 ? tags[i] + itgi : tags[i];
    }
    i1 = n_ + 1;
    n_ = k;
  }
}

void Geometry::scale(double f) {
  complete_ = false;
  for (int64_t i = 0; i < n_; ++i) {
    x[i] = x[i] * f;
    y[i] = y[i] * f;
    z[i] = z[i] * f;
    x2_[i] = x2_[i] * f;
    y2_[i] = y2_[i] * f;
    z2_[i] = z2_[i] * f;
    rad[i] = rad[i] * f;
  }
}

void Geometry::connect(int ignd) {
  // build_connections, src/c_geometry.cpp:1705-1915 (wires only).  For each
  // end, scan the other segments cyclically starting after this one and
  // record the first end within SMIN * |segment| (L1 distance).
  const int64_t n = n_;
  icon1.assign(n, 0);
  icon2.assign(n, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t iz = i + 1;
    const double za = z[i], zb = z2_[i];
    const double ex = x2_[i] - x[i], ey = y2_[i] - y[i], ez = z2_[i] - z[i];
    const double reach = std::sqrt((ex * ex + ey * ey) + ez * ez) * kSmin;
    bool grounded = false;
    if (ignd > 0) {
      if (za <= -reach) throw NecError(seg_msg("GEOMETRY DATA ERROR--SEGMENT ", iz, " EXTENDS BELOW GROUND"));
      if (za <= reach) {
        icon1[i] = static_cast<int32_t>(iz);
        z[i] = 0.;
        grounded = true;
      }
    }
    if (!grounded) {
      const double px = x[i], py = y[i], pz = z[i];
      int64_t c = i;
      for (int64_t step = 1; step < n; ++step) {
        c = (c + 1 == n) ? 0 : c + 1;
        if (manhattan(px, py, pz, x[c], y[c], z[c]) <= reach) {
          icon1[i] = static_cast<int32_t>(-(c + 1));
          break;
        }
        if (manhattan(px, py, pz, x2_[c], y2_[c], z2_[c]) <= reach) {
          icon1[i] = static_cast<int32_t>(c + 1);
          break;
        }
      }
    }
    if (ignd > 0 || grounded) {
      if (zb <= -reach) throw NecError(seg_msg("GEOMETRY DATA ERROR--SEGMENT ", iz, " EXTENDS BELOW GROUND"));
      if (zb <= reach) {
        if (icon1[i] == iz) throw NecError(seg_msg("GEOMETRY DATA ERROR--SEGMENT ", iz, " LIES IN GROUND PLANE"));
        icon2[i] = static_cast<int32_t>(iz);
        z2_[i] = 0.;
        continue;
      }
    }
    const double px = x2_[i], py = y2_[i], pz = z2_[i];
    int64_t c = i;
    for (int64_t step = 1; step < n; ++step) {
      c = (c + 1 == n) ? 0 : c + 1;
      if (manhattan(px, py, pz, x[c], y[c], z[c]) <= reach) {
        icon2[i] = static_cast<int32_t>(c + 1);
        break;
      }
      if (manhattan(px, py, pz, x2_[c], y2_[c], z2_[c]) <= reach) {
        icon2[i] = static_cast<int32_t>(-(c + 1));
        break;
      }
    }
  }
}

void Geometry::resolve_junctions() {
  // resolve_junctions, src/c_geometry.cpp:1918-2074.  Walk the ring of ends
  // that meet at each end of segment j; when j is the lowest-numbered member,
  // replace every member's coordinates by the mean, summed in ring order.
  const int64_t n = n_;
  std::vector<int64_t> ring;
  for (int64_t j = 0; j < n; ++j) {
    const int64_t self = j + 1;
    for (int end = 1; end <= 2; ++end) {
      ring.assign(1, end == 1 ? -self : self);
      double sx = end == 1 ? x[j] : x2_[j];
      double sy = end == 1 ? y[j] : y2_[j];
      double sz = end == 1 ? z[j] : z2_[j];
      int64_t link = end == 1 ? icon1[j] : icon2[j];
      int side = end == 1 ? -1 : 1;
      if (link == 0 || link == self || link > n) continue;
      bool handled_earlier = false;
      do {
        if (link < 0) link = -link; else side = -side;
        if (link == self) break;
        if (link < self) {
          handled_earlier = true;
          break;
        }
        ring.push_back(link * side);
        const int64_t k = link - 1;
        if (k < 0 || k >= n) throw NecError("CONNECT - SEGMENT CONNECTION ERROR (bad index): " + std::to_string(link));
        if (side != 1) {
          sx = sx + x[k];
          sy = sy + y[k];
          sz = sz + z[k];
          link = icon1[k];
        } else {
          sx = sx + x2_[k];
          sy = sy + y2_[k];
          sz = sz + z2_[k];
          link = icon2[k];
        }
      } while (link != 0);
      if (handled_earlier) {
        // The reference abandons the walk of this end and, when this was end
        // 1, moves on to end 2 (src/c_geometry.cpp:1998-2011).
        continue;
      }
      const double cnt = static_cast<double>(ring.size());
      sx = sx / cnt;
      sy = sy / cnt;
      sz = sz / cnt;
      for (int64_t e : ring) {
        if (e <= 0) {
          const int64_t k = -e - 1;
          x[k] = sx;
          y[k] = sy;
          z[k] = sz;
        } else {
          const int64_t k = e - 1;
          x2_[k] = sx;
          y2_[k] = sy;
          z2_[k] = sz;
        }
      }
    }
  }
}

void Geometry::complete(int gpflag) {
  // src/c_geometry.cpp:454-604
  if (n_ == 0) throw NecError("Geometry has no wires or patches.");
  if (n_ <= 1)
    throw NecError("GEOMETRY HAS ONE OR FEWER SEGMENTS and no patches.");
  connect(gpflag);
  resolve_junctions();
  const int64_t n = n_;
  len.assign(n, 0.);
  cab.assign(n, 0.);
  sab.assign(n, 0.);
  salp.assign(n, 0.);
  for (int64_t i = 0; i < n; ++i) {
    const double dx = x2_[i] - x[i], dy = y2_[i] - y[i], dz = z2_[i] - z[i];
    x[i] = (x[i] + x2_[i]) * .5;
    y[i] = (y[i] + y2_[i]) * .5;
    z[i] = (z[i] + z2_[i]) * .5;
    const double l2 = dx * dx + dy * dy + dz * dz;
    double l = std::sqrt(l2);
    l = (l2 / l + l) * .5;  // one Newton step, src/c_geometry.cpp:512-514
    len[i] = l;
    cab[i] = dx / l;
    sab[i] = dy / l;
    double sz = dz / l;
    if (sz > 1.) sz = 1.;
    if (sz < -1.) sz = -1.;
    salp[i] = sz;
    if (len[i] <= 1.e-20 || rad[i] <= 0.) throw NecError("SEGMENT DATA ERROR");
    // non-finite input would leave the device quadrature without an end
    if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(z[i]) || !std::isfinite(len[i]) ||
        !std::isfinite(rad[i]))
      throw NecError("SEGMENT DATA ERROR (non-finite coordinate or radius)");
  }
  complete_ = true;
}

int64_t Geometry::segment_number(int tag, int m) const {
  // src/c_geometry.cpp:76-99
  if (m <= 0)
    throw NecError("CHECK DATA, PARAMETER SPECIFYING SEGMENT POSITION IN A GROUP OF EQUAL TAGS MUST NOT BE ZERO");
  if (tag == 0) return m;
  int64_t seen = 0;
  for (int64_t i = 0; i < n_; ++i)
    if (tags[i] == tag && ++seen == m) return i + 1;
  throw NecError("NO SEGMENT HAS AN ITAG OF " + std::to_string(tag));
}

ScaledSegments Geometry::scaled(double freq_mhz) const {
  // src/c_geometry.cpp:3389-3418
  const double fr = (1.0e6 * freq_mhz) / em_speed_of_light();
  ScaledSegments s;
  const int64_t n = n_;
  s.x.resize(n);
  s.y.resize(n);
  s.z.resize(n);
  s.s.resize(n);
  s.b.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    s.x[i] = x[i] * fr;
    s.y[i] = y[i] * fr;
    s.z[i] = z[i] * fr;
    s.s[i] = len[i] * fr;
    s.b[i] = rad[i] * fr;
  }
  s.cab = cab;
  s.sab = sab;
  s.salp = salp;
  return s;
}

}  // namespace necgpu
