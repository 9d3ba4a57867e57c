// kf_setup.hpp -- host-side preparation of the kernel arguments: the
// device plane record and the physics constants.  Shared by the C ABI
// (kf_api.hip) and the CPU build of the per-track code used by the tests.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstring>

#include "kf_kernels.hpp"

namespace kfd {

// Physics constants evaluated the way the reference's unit literals do
// (long double literal times a double unit, rounded to double;
// Interactions.ipp:19-25, Units.hpp:145-153) and its float literals.
inline void physicsConstants(FitArgs &a) {
  a.kMe = (double)(0.001L * 0.5109989461L);
  a.kK = ((double)(0.001L * 0.307075L)) * 10.0 * 10.0;
  a.kPlasma = (double)(1e-9L * 28.816L);
  a.kHighland = (double)(0.001L * 13.6L);
  const float c = 0.038f * 2;
  a.kHighlandLog = (double)c;
  const float conv = 2 * std::sqrt(2 * std::log(2.0f));
  a.kLandauConv = (double)conv;
}

// Device plane record: frame, and the MaterialSlab (MaterialSlab.ipp:15-20)
// and Material (Material.ipp:80-85, Material.hpp:102-104) quantities, only
// when the slab is valid (MaterialSlab.hpp:46-48).
inline DevPlane devPlane(const kf_plane &p) {
  DevPlane d;
  std::memset(&d, 0, sizeof d);
  for (int k = 0; k < 3; ++k) d.c[k] = p.center[k];
  for (int k = 0; k < 9; ++k) d.r[k] = p.rot[k];
  const bool valid = p.thickness > 0.0 && 0.0f < p.ar;
  d.has_material = valid ? 1 : 0;
  if (valid) {
    const double eps = 2 * DBL_EPSILON;
    d.thickness = p.thickness;
    d.thicknessInX0 = (eps < p.x0) ? (p.thickness / p.x0) : 0.0;
    d.meanExcitation = (16.0 * 1e-9) * std::pow(p.z, (double)0.9f);
    d.molarElectronDensity = p.z * p.molar_rho;
  }
  d.geo_id = p.geo_id;
  return d;
}

}  // namespace kfd
