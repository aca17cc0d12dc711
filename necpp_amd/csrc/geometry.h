// geometry.h -- wire geometry of a nec_context (host side).
//
// Mirrors the parts of c_geometry that the wire MoM path needs:
//   wire()                 src/c_geometry.cpp:614-683
//   geometry_complete()    src/c_geometry.cpp:454-604
//   build_connections()    src/c_geometry.cpp:1705-1915
//   resolve_junctions()    src/c_geometry.cpp:1918-2074
//   get_segment_number()   src/c_geometry.cpp:76-99
//   frequency_scale()      src/c_geometry.cpp:3389-3418
// Junction averaging and the centre / length / direction-cosine formulas keep
// the reference's operation order, so the segment tables are bit-identical
// to nec2++'s.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace necgpu {

// nec_exception equivalent: card functions catch it and return 1.
class NecError : public std::runtime_error {
 public:
  explicit NecError(const std::string& m) : std::runtime_error(m) {}
};

// Structure-of-arrays segment table in wavelengths (one frequency).
struct ScaledSegments {
  std::vector<double> x, y, z, s, b, cab, sab, salp;
};

class Geometry {
 public:
  void add_wire(int tag, int ns, double xw1, double yw1, double zw1, double xw2, double yw2, double zw2,
                double rad, double rdel, double rrad);
  // GM: rotate about x, y, z by rox, roy, roz (radians), then translate,
  // the segments from the first one with tag its on (all with its = 0); with
  // nrpt > 0 append nrpt transformed copies instead, each of the previous
  // one, non-zero tags raised by itgi per copy (c_geometry::move,
  // src/c_geometry.cpp:830-900, wires only)
  void move(double rox, double roy, double roz, double xs, double ys, double zs, int its, int nrpt, int itgi);
  // GS: every segment's end points and radius times f (src/c_geometry.cpp:1308-1318)
  void scale(double f);
  void complete(int gpflag);
  bool is_complete() const { return complete_; }
  int64_t size() const { return n_; }
  int64_t segment_number(int tag, int m) const;
  ScaledSegments scaled(double freq_mhz) const;

  // Segment tables after complete() (metres, unscaled).
  std::vector<double> x, y, z, len, rad, cab, sab, salp;
  std::vector<int32_t> tags, icon1, icon2;

 private:
  void connect(int ignd);
  void resolve_junctions();

  int64_t n_ = 0;
  bool complete_ = false;
  // segment end points as generated (end 1 = x/y/z before complete, end 2 = x2/y2/z2)
  std::vector<double> x2_, y2_, z2_;
};

}  // namespace necgpu
