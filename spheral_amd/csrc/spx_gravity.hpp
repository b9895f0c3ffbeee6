// spx_gravity.hpp -- TreeGravity on the device (placeholder until the octree lands).
#pragma once
#include "spx_device.hpp"
namespace spx {
struct GravityTree {
  void init(int, int64_t, int, hipStream_t) {}
  void release() {}
  void build(double* const*, const double*, int) {}
  void evaluate(double* const*, const double*, double* const*, double*, int, double, double, double, Scalars*) {}
  void finishDt(Scalars*, double, double, double) {}
};
}  // namespace spx
