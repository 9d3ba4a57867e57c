// kfo_fitter.hpp -- FP64 CPU restatement of the gpuKalmanFitter track fit.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP fitter is checked
// against: the reference's own algorithm, restated in plain C++ from its
// sources (the reference needs Eigen, Boost.Range, ROOT and nvcc, none of which
// exist here, so it cannot be compiled).  It keeps the reference's call
// structure -- Propagator loop, DirectNavigator, KalmanFitter Actor/Aborter,
// GainMatrixUpdater, GainMatrixSmoother, EigenStepper, PointwiseMaterial-
// Interaction -- so each function can be read side by side with the file:line
// it cites.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg use it.
//
// Reference quirks kept on purpose (each is the reference's behaviour):
//  * the accuracy step never grows back after a rejected RK trial; it is reset
//    only by finalize() (EigenStepper.ipp:284, KalmanFitter.hpp:384-385);
//  * stepper.update() keeps the stepper charge q and the Jacobians; after the
//    smoother the backward propagation starts from jacToGlobal built at the
//    last measured surface (KalmanFitter.hpp:379-380, EigenStepper.ipp:554-564);
//  * phi is wrapped and theta clamped whenever bound parameters are built
//    (ParameterSet.hpp:276-281, ParameterTypes.hpp:72-115);
//  * the covariance update is (I - K H) P without symmetrisation
//    (GainMatrixUpdater.hpp:77-84);
//  * material noise is added to the bound covariance diagonal after the
//    update (PointwiseMaterialInteraction.hpp:137-148);
//  * float literals of the reference (0.9f, 0.038f, log(2.0f)) keep their
//    float values in the double build.
// Definitions this build adds (the reference has none):
//  * chi2 of a state = r^T S^-1 r with the predicted residual r;
//  * the fit succeeds when every input measurement was used (the reference
//    compares with the surface-sequence size; identical without holes);
//  * the smoother runs over the states that were filled.
#pragma once

#include <cfloat>
#include <cstdint>
#include <vector>

#include "kfo_linalg.hpp"

namespace kfo {

// ---------------------------------------------------------------- units ----
// include/Utilities/Units.hpp:87-142; unit literals are evaluated in long
// double and rounded to double, as the reference's literal operators do.
namespace units {
constexpr double mm = 1.0;
constexpr double cm = 10.0;
constexpr double um = 1e-3;
constexpr double MeV = 1e-3;
constexpr double eV = 1e-9;
constexpr double T = 0.000299792458;
constexpr double ns = 1e-9 * 299792458000.0;
constexpr double u = 0.93149410242;
constexpr double mol = 1.0;
}  // namespace units

constexpr double kOnSurfaceTolerance = 1e-4;         // Definitions.hpp:41
constexpr double kCurvilinearProjTolerance = 0.999995;  // Definitions.hpp:46
constexpr double kOverstepLimit = 0.01;              // EigenStepper.hpp:370
constexpr double kPi = 3.14159265358979323846;       // M_PI

// Interactions.ipp:19-25 (RPP2018 table 33.1)
constexpr double kMe = (double)(0.001L * 0.5109989461L);
constexpr double kK = ((double)(0.001L * 0.307075L)) * 10.0 * 10.0;
constexpr double kPlasmaEnergyScale = (double)(1e-9L * 28.816L);
constexpr double kHighland = (double)(0.001L * 13.6L);

// ------------------------------------------------------------- material ----
// include/Material/Material.hpp, detail/Material.ipp:80-85,
// detail/MaterialSlab.ipp:11-25
struct Material {
  Real x0 = 0.0, l0 = 0.0, ar = 0.0, z = 0.0, molarRho = 0.0;
  bool valid() const { return Real(0.0f) < ar; }
  Real meanExcitationEnergy() const {
    return Real(16.0 * units::eV) * kpow(z, Real((double)0.9f));
  }
  Real molarElectronDensity() const { return z * molarRho; }
};

struct MaterialSlab {
  Material material;
  Real thickness = 0.0, thicknessInX0 = 0.0, thicknessInL0 = 0.0;
  MaterialSlab() = default;
  MaterialSlab(const Material &m, Real t) : material(m), thickness(t) {
    const double eps = 2 * DBL_EPSILON;
    thicknessInX0 = (Real(eps) < m.x0) ? Real(t / m.x0) : Real(0.0);
    thicknessInL0 = (Real(eps) < m.l0) ? Real(t / m.l0) : Real(0.0);
  }
  explicit operator bool() const {
    return material.valid() && (Real(0.0f) < thickness);
  }
  void scaleThickness(Real s) {
    thickness = thickness * s;
    thicknessInX0 = thicknessInX0 * s;
    thicknessInL0 = thicknessInL0 * s;
  }
};

// ------------------------------------------------------------- surfaces ----
enum class SurfaceKind { Plane, Line };
enum class IStatus { Unreachable = 0, Reachable = 1, OnSurface = 2 };

struct Intersection {
  V3 position = V3::zero();
  Real pathLength = Real(HUGE_VAL);
  IStatus status = IStatus::Unreachable;
};

// PlaneSurface (include/Surfaces/detail/PlaneSurface.ipp) and LineSurface
// (include/Surfaces/detail/LineSurface.ipp) with a HomogeneousSurfaceMaterial
// (include/Material/HomogeneousSurfaceMaterial.hpp).
struct Surface {
  SurfaceKind kind = SurfaceKind::Plane;
  M33 rot = M33::identity();  // columns: local x, local y, normal
  V3 center = V3::zero();
  MaterialSlab slab;          // full material (no split factor on planes)
  int geoId = -1;
};

// ---------------------------------------------------------------- field ----
struct Field {
  bool grid = false;
  V3 b = V3::zero();
  const float *values = nullptr;
  int n[3] = {0, 0, 0};
  double lo[3] = {0, 0, 0};
  double hi[3] = {0, 0, 0};
  double scale = 1.0;

  // Test::ConstantBField::getField (include/Test/Helper.hpp:17-22) or the
  // trilinear grid lookup (InterpolatedBFieldMap.hpp:178-181, Grid.hpp:331-358,
  // interpolation_impl.hpp:98-131: the last axis is interpolated first).
  V3 get(const V3 &pos) const {
    if (!grid) {
      if (scale == 1.0) return b;
      return v3(Real(scale) * b[0], Real(scale) * b[1], Real(scale) * b[2]);
    }
    int idx[3];
    Real f[3];
    for (int d = 0; d < 3; ++d) {
      Real x = pos[d];
      if (x < Real(lo[d])) x = lo[d];
      if (Real(hi[d]) < x) x = hi[d];
      const double w = (hi[d] - lo[d]) / (n[d] - 1);
      int i = (int)kval(kfloor((x - Real(lo[d])) / Real(w)));
      if (i < 0) i = 0;
      if (i > n[d] - 2) i = n[d] - 2;
      idx[d] = i;
      Real ll = Real(lo[d]) + Real((double)i) * Real(w);
      Real ur = Real(lo[d]) + Real((double)(i + 1)) * Real(w);
      f[d] = (x - ll) / (ur - ll);
    }
    // corner c = (dx, dy, dz) with z fastest
    Real cz[4][3];
    for (int c = 0; c < 4; ++c) {
      const int dx = c >> 1, dy = c & 1;
      const int ix = idx[0] + dx, iy = idx[1] + dy;
      const size_t b0 = (((size_t)ix * n[1] + iy) * n[2] + idx[2]) * 3;
      const size_t b1 = b0 + 3;
      for (int k = 0; k < 3; ++k)
        cz[c][k] = (Real(1.0) - f[2]) * Real((double)values[b0 + k]) +
                   f[2] * Real((double)values[b1 + k]);
    }
    Real cy[2][3];
    for (int c = 0; c < 2; ++c)
      for (int k = 0; k < 3; ++k)
        cy[c][k] = (Real(1.0) - f[1]) * cz[2 * c][k] + f[1] * cz[2 * c + 1][k];
    V3 r;
    for (int k = 0; k < 3; ++k)
      r[k] = (Real(1.0) - f[0]) * cy[0][k] + f[0] * cy[1][k];
    if (scale != 1.0)
      for (int k = 0; k < 3; ++k) r[k] = Real(scale) * r[k];
    return r;
  }
};

// ----------------------------------------------------- surface geometry ----
inline V3 surfNormal(const Surface &s) { return col(s.rot, 2); }

// PlaneSurface.ipp:116-120 / LineSurface.ipp:46-61
inline M33 referenceFrame(const Surface &s, const V3 & /*pos*/, const V3 &mom) {
  if (s.kind == SurfaceKind::Plane) return s.rot;
  V3 measY = col(s.rot, 2);
  V3 measX = normalized(cross(measY, mom));
  V3 measDepth = cross(measX, measY);
  M33 m;
  for (int i = 0; i < 3; ++i) {
    m(i, 0) = measX[i];
    m(i, 1) = measY[i];
    m(i, 2) = measDepth[i];
  }
  return m;
}

// R^-1 (pos - c).  The surfaces here have orthonormal rotations, so R^T is
// the inverse of the reference's transform.inverse() (PlaneSurface.ipp:220).
inline V3 toLocal3(const Surface &s, const V3 &pos) {
  V3 d = pos - s.center;
  return v3(s.rot(0, 0) * d[0] + s.rot(1, 0) * d[1] + s.rot(2, 0) * d[2],
            s.rot(0, 1) * d[0] + s.rot(1, 1) * d[1] + s.rot(2, 1) * d[2],
            s.rot(0, 2) * d[0] + s.rot(1, 2) * d[1] + s.rot(2, 2) * d[2]);
}

// PlaneSurface.ipp:206-213, LineSurface.ipp:184-198
inline V3 localToGlobal(const Surface &s, const V2 &l, const V3 &mom) {
  if (s.kind == SurfaceKind::Plane) {
    V3 loc = v3(l[0], l[1], 0.0);
    return s.rot * loc + s.center;
  }
  V3 lineDir = col(s.rot, 2);
  V3 radiusAxis = cross(lineDir, mom);
  V3 locZ = s.rot * v3(0.0, 0.0, l[1]) + s.center;
  V3 ra = normalized(radiusAxis);
  return locZ + l[0] * ra;
}

// PlaneSurface.ipp:215-226, LineSurface.ipp:200-220
inline V2 globalToLocal(const Surface &s, const V3 &pos, const V3 &mom) {
  V3 l3 = toLocal3(s, pos);
  V2 l;
  if (s.kind == SurfaceKind::Plane) {
    l[0] = l3[0];
    l[1] = l3[1];
    return l;
  }
  l[0] = ksqrt(l3[0] * l3[0] + l3[1] * l3[1]);
  l[1] = l3[2];
  V3 lineDir = col(s.rot, 2);
  V3 dec = pos - s.center;
  Real sign = (dot(cross(lineDir, mom), dec) < Real(0.0)) ? Real(-1.0) : Real(1.0);
  l[0] = l[0] * sign;
  return l;
}

// PlanarHelper.hpp:28-49 and LineSurface.ipp:229-262
inline Intersection intersect(const Surface &s, const V3 &pos, const V3 &dir) {
  Intersection r;
  if (s.kind == SurfaceKind::Plane) {
    V3 n = col(s.rot, 2);
    Real denom = dot(dir, n);
    if (denom != Real(0.0)) {
      Real path = dot(n, s.center - pos) / denom;
      r.status = (path * path < Real(kOnSurfaceTolerance * kOnSurfaceTolerance))
                     ? IStatus::OnSurface
                     : IStatus::Reachable;
      r.position = pos + path * dir;
      r.pathLength = path;
    }
    return r;
  }
  V3 mb = s.center;
  V3 eb = col(s.rot, 2);
  V3 mab = mb - pos;
  Real eaTeb = dot(dir, eb);
  Real denom = Real(1.0) - eaTeb * eaTeb;
  if (denom * denom > Real(kOnSurfaceTolerance * kOnSurfaceTolerance)) {
    Real u = (dot(mab, dir) - dot(mab, eb) * eaTeb) / denom;
    r.status = (u * u < Real(kOnSurfaceTolerance * kOnSurfaceTolerance))
                   ? IStatus::OnSurface
                   : IStatus::Reachable;
    r.position = pos + u * dir;
    r.pathLength = u;
    return r;
  }
  r.position = pos;
  r.pathLength = DBL_MAX;
  r.status = IStatus::Unreachable;
  return r;
}

struct DirTrig {
  Real cosTheta, sinTheta, invSinTheta, cosPhi, sinPhi;
};
inline DirTrig dirTrig(const V3 &d) {
  DirTrig t;
  t.cosTheta = d[2];
  t.sinTheta = ksqrt(d[0] * d[0] + d[1] * d[1]);
  t.invSinTheta = Real(1.0) / t.sinTheta;
  t.cosPhi = d[0] * t.invSinTheta;
  t.sinPhi = d[1] * t.invSinTheta;
  return t;
}

// PlaneSurface.ipp:122-156, LineSurface.ipp:63-112
inline void initJacobianToGlobal(const Surface &s, M86 &jac, const V3 &pos,
                                 const V3 &dir, const V6 &pars) {
  DirTrig tr = dirTrig(dir);
  M33 rf = referenceFrame(s, pos, dir);
  for (int i = 0; i < 3; ++i) {
    jac(i, 0) = rf(i, 0);
    jac(i, 1) = rf(i, 1);
  }
  jac(3, 5) = 1.0;
  jac(4, 2) = (-tr.sinTheta) * tr.sinPhi;
  jac(4, 3) = tr.cosTheta * tr.cosPhi;
  jac(5, 2) = tr.sinTheta * tr.cosPhi;
  jac(5, 3) = tr.cosTheta * tr.sinPhi;
  jac(6, 3) = (-tr.sinTheta);
  jac(7, 4) = 1.0;
  if (s.kind == SurfaceKind::Line) {
    Real ipdn = Real(1.0) / dot(dir, col(rf, 2));
    V3 dPhi = v3(jac(4, 2), jac(5, 2), jac(6, 2));
    V3 dTheta = v3(jac(4, 3), jac(5, 3), jac(6, 3));
    V3 y = col(rf, 1), x = col(rf, 0);
    V3 dDPhiY = cross(y, dPhi);
    V3 dDThetaY = cross(y, dTheta);
    dDPhiY = dDPhiY - dot(x, dDPhiY) * x;
    dDThetaY = dDThetaY - dot(x, dDThetaY) * x;
    for (int i = 0; i < 3; ++i) {
      jac(i, 2) = dDPhiY[i] * pars[0] * ipdn;
      jac(i, 3) = dDThetaY[i] * pars[0] * ipdn;
    }
  }
}

// PlaneSurface.ipp:158-189, LineSurface.ipp:114-143; returns rframe^T
inline M33 initJacobianToLocal(const Surface &s, M68 &jac, const V3 &pos,
                               const V3 &dir) {
  DirTrig tr = dirTrig(dir);
  M33 rft = transpose(referenceFrame(s, pos, dir));
  for (int j = 0; j < 3; ++j) {
    jac(0, j) = rft(0, j);
    jac(1, j) = rft(1, j);
  }
  jac(5, 3) = 1.0;
  jac(2, 4) = -tr.sinPhi * tr.invSinTheta;
  jac(2, 5) = tr.cosPhi * tr.invSinTheta;
  jac(3, 4) = tr.cosPhi * tr.cosTheta;
  jac(3, 5) = tr.sinPhi * tr.cosTheta;
  jac(3, 6) = -tr.sinTheta;
  jac(4, 7) = 1.0;
  return rft;
}

// PlaneSurface.ipp:191-204, LineSurface.ipp:145-182
inline R6 derivativeFactors(const Surface &s, const V3 &pos, const V3 &dir,
                            const M33 &rft, const M86 &jac) {
  Mat<3, 6> top;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 6; ++j) top(i, j) = jac(i, j);
  if (s.kind == SurfaceKind::Plane) {
    Mat<1, 3> nv;
    for (int j = 0; j < 3; ++j) nv(0, j) = rft(2, j);
    Real dp = nv(0, 0) * dir[0] + nv(0, 1) * dir[1] + nv(0, 2) * dir[2];
    for (int j = 0; j < 3; ++j) nv(0, j) = nv(0, j) / dp;
    return nv * top;
  }
  Mat<1, 3> pc;
  for (int j = 0; j < 3; ++j) pc(0, j) = pos[j] - s.center[j];
  Mat<1, 3> locz;
  for (int j = 0; j < 3; ++j) locz(0, j) = rft(1, j);
  Real long_c = locz(0, 0) * dir[0] + locz(0, 1) * dir[1] + locz(0, 2) * dir[2];
  Mat<1, 3> norm_vec;
  for (int j = 0; j < 3; ++j) norm_vec(0, j) = dir[j] - long_c * locz(0, j);
  R6 s_vec = norm_vec * top;
  Mat<3, 6> dirRows;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 6; ++j) dirRows(i, j) = jac(4 + i, j);
  R6 d_vec = locz * dirRows;  // computed by the reference, unused there too
  (void)d_vec;
  Real nrm = Real(1.0) / (Real(1.0) - long_c * long_c);
  Mat<3, 6> long_mat = Mat<3, 6>::zero();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 6; ++j) long_mat(i, j) = long_mat(i, j) + locz(0, i);
  M66 diag = M66::zero();
  for (int k = 0; k < 6; ++k) diag(k, k) = s_vec(0, k);
  Mat<3, 6> diff = long_mat * diag;
  diff = diff - dirRows;
  R6 row_vec = pc * diff;
  R6 r;
  for (int j = 0; j < 6; ++j) r(0, j) = nrm * (s_vec(0, j) - row_vec(0, j));
  return r;
}

// PlaneSurface.ipp:228-236 (LineSurface.ipp:222-227 returns 1)
inline Real pathCorrection(const Surface &s, const V3 &dir) {
  if (s.kind == SurfaceKind::Line) return 1.0;
  return Real(1.0) / kabs(dot(surfNormal(s), dir));
}

// ------------------------------------------------------- bound parameters ----
// ParameterSet::setParameter applies the parameter-type correction twice
// (ParameterSet.hpp:276-281): phi cyclic in [-pi, pi), theta bound to [0, pi]
// (ParameterTypes.hpp:56-115, ParameterDefinitions.hpp:138-147).
inline Real cyclicPhi(Real x) {
  const Real mn = -kPi, mx = kPi;
  if (mn <= x && x < mx) return x;
  return x - (mx - mn) * kfloor((x - mn) / (mx - mn));
}
inline Real boundTheta(Real x) {
  const Real mn = 0.0, mx = kPi;
  return (x > mx) ? mx : ((x < mn) ? mn : x);
}
inline V6 correctParams(const V6 &p) {
  V6 r = p;
  r[2] = cyclicPhi(cyclicPhi(p[2]));
  r[3] = boundTheta(boundTheta(p[3]));
  return r;
}

// coordinate_transformations.hpp:66-76
inline V3 parameters2globalMomentum(const V6 &p) {
  Real pabs = kabs(Real(1.0) / p[4]);
  Real phi = p[2], theta = p[3];
  return v3(pabs * ksin(theta) * kcos(phi), pabs * ksin(theta) * ksin(phi),
            pabs * kcos(theta));
}

// Helpers / UnitVectors.hpp:49-58
inline V3 dirFromPhiTheta(Real phi, Real theta) {
  Real cosTheta = kcos(theta), sinTheta = ksin(theta);
  return v3(kcos(phi) * sinTheta, ksin(phi) * sinTheta, cosTheta);
}

inline Real vphi(const V3 &v) { return katan2(v[1], v[0]); }
inline Real vtheta(const V3 &v) {
  return katan2(ksqrt(v[0] * v[0] + v[1] * v[1]), v[2]);
}

// SingleBoundTrackParameters<ChargedPolicy>
// (include/EventData/SingleBoundTrackParameters.hpp:27-228)
struct BoundParams {
  V6 par = V6::zero();
  M66 cov = M66::zero();
  V3 position = V3::zero();
  V3 momentum = V3::zero();
  Real charge = 1.0;
  const Surface *surface = nullptr;
};

// ctor from (cov, parameter values, surface): SingleBoundTrackParameters.hpp:48-60
inline BoundParams makeBound(const M66 &cov, const V6 &pv, const Surface *s) {
  BoundParams b;
  b.momentum = parameters2globalMomentum(pv);
  V2 l;
  l[0] = pv[0];
  l[1] = pv[1];
  b.position = localToGlobal(*s, l, b.momentum);
  b.charge = (pv[4] > Real(0.0)) ? Real(1.0) : Real(-1.0);
  b.par = correctParams(pv);
  b.cov = cov;
  b.surface = s;
  return b;
}

// ctor from (cov, position, momentum, charge, time, surface):
// SingleBoundTrackParameters.hpp:62-76 + global2parameters,
// coordinate_transformations.hpp:118-130
inline BoundParams makeBoundGlobal(const M66 &cov, const V3 &pos, const V3 &mom,
                                   Real charge, Real time, const Surface *s) {
  V2 l = globalToLocal(*s, pos, mom);
  V6 pv;
  pv[0] = l[0];
  pv[1] = l[1];
  pv[2] = vphi(mom);
  pv[3] = vtheta(mom);
  pv[4] = ((kabs(charge) < Real(1e-4)) ? Real(1.0) : charge) / norm(mom);
  pv[5] = time;
  BoundParams b;
  b.position = pos;
  b.momentum = mom;
  b.charge = (pv[4] > Real(0.0)) ? Real(1.0) : Real(-1.0);
  b.par = correctParams(pv);
  b.cov = cov;
  b.surface = s;
  return b;
}

// boundParameters2freeParameters, coordinate_transformations.hpp:145-158
inline V8 boundToFree(const V6 &p, const Surface &s) {
  V3 mom = parameters2globalMomentum(p);
  V2 l;
  l[0] = p[0];
  l[1] = p[1];
  V3 pos = localToGlobal(s, l, mom);
  V3 d = dirFromPhiTheta(p[2], p[3]);
  V8 f;
  f[0] = pos[0];
  f[1] = pos[1];
  f[2] = pos[2];
  f[3] = p[5];
  f[4] = d[0];
  f[5] = d[1];
  f[6] = d[2];
  f[7] = p[4];
  return f;
}

// ------------------------------------------------------ ConstrainedStep ----
// include/Propagator/ConstrainedStep.hpp:22-139
enum StepType { kAccuracy = 0, kActor = 1, kAborter = 2, kUser = 3 };
struct ConstrainedStep {
  Real values[4];
  int direction = 1;
  ConstrainedStep() : ConstrainedStep(Real(DBL_MAX)) {}
  explicit ConstrainedStep(Real value) {
    direction = (value > Real(0.0)) ? 1 : -1;
    values[kAccuracy] = Real(DBL_MAX) * Real((double)direction);
    values[kActor] = Real(DBL_MAX) * Real((double)direction);
    values[kAborter] = Real(DBL_MAX) * Real((double)direction);
    values[kUser] = value;
  }
  Real maxv() const {
    Real m = -DBL_MAX;
    for (int i = 0; i < 4; ++i)
      if (values[i] > m) m = values[i];
    return m;
  }
  Real minv() const {
    Real m = DBL_MAX;
    for (int i = 0; i < 4; ++i)
      if (values[i] < m) m = values[i];
    return m;
  }
  void release(int type) { values[type] = (direction == 1) ? maxv() : minv(); }
  void update(Real value, int type, bool releaseStep = false) {
    if (releaseStep) release(type);
    Real c = values[type];
    values[type] = (kabs(c) < kabs(value)) ? c : value;
  }
  void assign(Real value) {
    values[kAccuracy] = value;
    direction = (value > Real(0.0)) ? 1 : -1;
  }
  Real get() const { return (direction == 1) ? minv() : maxv(); }
};

// ---------------------------------------------------------------- options ----
struct PropOptions {
  int absPdgCode = 13;
  Real mass = 105.6583755 * units::MeV;
  unsigned maxSteps = 1000;
  unsigned maxRungeKuttaStepTrials = 10000;
  Real maxStepSize = DBL_MAX;
  Real targetTolerance = kOnSurfaceTolerance;
  Real tolerance = 1e-4;
  Real stepSizeCutOff = 0.0;
  Real overstepLimit = kOverstepLimit;  // EigenStepper::m_overstepLimit
};

// --------------------------------------------------------- stepper state ----
// EigenStepper::State, include/Propagator/EigenStepper.hpp:27-120
struct StepperState {
  V3 pos = v3(0.0, 0.0, 0.0);
  V3 dir = v3(1.0, 0.0, 0.0);
  Real p = 0.0;
  int q = 1;
  Real t = 0.0;
  int navDir = 1;
  M66 jacobian = M66::identity();
  M86 jacToGlobal = M86::zero();
  M88 jacTransport = M88::identity();
  V8 derivative = V8::zero();
  bool covTransport = false;
  M66 cov = M66::zero();
  Real pathAccumulated = 0.0;
  ConstrainedStep stepSize;
  Real previousStepSize = 0.0;
  Real tolerance = kOnSurfaceTolerance;
  bool stepFailed = false;  // this build: records a failed step() return
};

struct NavState {
  const Surface *const *sequence = nullptr;  // ordered measurement surfaces
  unsigned size = 0;
  unsigned next = 0;
  const Surface *startSurface = nullptr;
  const Surface *currentSurface = nullptr;
  const Surface *targetSurface = nullptr;
  bool targetReached = false;
  bool navigationBreak = false;
};

struct PropState {
  PropOptions options;
  StepperState stepping;
  NavState navigation;
};

// ---------------------------------------------------------- track states ----
struct SourceLink {
  V2 meas;
  M22 cov;
  int geoId;
};

// TrackState, include/EventData/TrackState.hpp:141-168
struct TrackStateRec {
  SourceLink sl;
  BoundParams predicted, filtered, smoothed;
  M66 jacobian = M66::identity();
  Real pathLength = 0.0;
  Real chi2 = 0.0;
  bool outlier = false;  // left without an update by the outlier finder
};

// KalmanFitterResult, include/Fitter/KalmanFitter.hpp:98-120
struct FitResult {
  std::vector<TrackStateRec> states;  // fittedStates
  BoundParams fittedParameters;
  unsigned measurementStates = 0;
  bool smoothed = false;
  bool finished = false;
  bool result = true;
  bool smootherNaN = false;
  bool maxStepsReached = false;
  bool navBreak = false;
  bool stepFailed = false;
  unsigned steps = 0;
};

struct FitOptions {
  bool smoothing = true;
  bool multipleScattering = true;
  bool energyLoss = true;
  const Surface *referenceSurface = nullptr;
  // outlier finder: 0 = VoidOutlierFinder (VoidKalmanComponents.hpp:66-80);
  // > 0: a predicted-residual chi2 above it leaves the state unupdated
  double outlierChi2Cut = 0.0;
};

// ------------------------------------------------------------ the stepper ----
class Stepper {
 public:
  explicit Stepper(const Field *f) : field_(f) {}

  // EigenStepper::State ctor, EigenStepper.hpp:41-61
  StepperState makeState(const BoundParams &par, int ndir, Real ssize,
                         Real stol) const {
    StepperState s;
    s.pos = par.position;
    s.dir = normalized(par.momentum);
    s.p = norm(par.momentum);
    s.q = (int)kval(par.charge);
    s.t = par.par[5];
    s.navDir = ndir;
    s.stepSize = ConstrainedStep(Real((double)ndir) * kabs(ssize));
    s.tolerance = stol;
    s.covTransport = true;
    s.cov = par.cov;
    initJacobianToGlobal(*par.surface, s.jacToGlobal, s.pos, s.dir, par.par);
    return s;
  }

  V3 getField(const V3 &pos) const { return field_->get(pos); }

  // EigenStepper::step, EigenStepper.ipp:226-332
  bool step(PropState &st) const {
    StepperState &s = st.stepping;
    V3 B_first, B_middle, B_last, k1, k2, k3, k4;
    Real error_estimate = 0.0;
    const Real qop = Real((double)s.q) / s.p;

    B_first = getField(s.pos);
    k1 = qop * cross(s.dir, B_first);  // evaluatek i == 0, EigenStepper.ipp:18-33

    auto tryRungeKuttaStep = [&](Real h) -> bool {
      const Real h2 = h * h;
      const Real half_h = h * 0.5;
      V3 pos1 = s.pos + half_h * s.dir + (h2 * 0.125) * k1;
      B_middle = getField(pos1);
      k2 = qop * cross(s.dir + half_h * k1, B_middle);
      k3 = qop * cross(s.dir + half_h * k2, B_middle);
      V3 pos2 = s.pos + h * s.dir + (h2 * 0.5) * k3;
      B_last = getField(pos2);
      k4 = qop * cross(s.dir + h * k3, B_last);
      V3 e = k1 - k2 - k3 + k4;
      error_estimate = kmax(h2 * (kabs(e[0]) + kabs(e[1]) + kabs(e[2])), Real(1e-20));

      return error_estimate <= st.options.tolerance;
    };

    Real stepSizeScaling = 1.0;
    size_t nStepTrials = 0;
    while (!tryRungeKuttaStep(s.stepSize.get())) {
      stepSizeScaling = kmin(
          kmax(Real(0.25), kpow(st.options.tolerance / kabs(Real(2.0) * error_estimate),
                                Real(0.25))),
          Real(4.0));
      s.stepSize.assign(s.stepSize.get() * stepSizeScaling);
      if (s.stepSize.get() * s.stepSize.get() <
          st.options.stepSizeCutOff * st.options.stepSizeCutOff) {
        s.stepFailed = true;
        return false;
      }
      if (nStepTrials > st.options.maxRungeKuttaStepTrials) {
        s.stepFailed = true;
        return false;
      }
      nStepTrials++;
    }

    const Real h = s.stepSize.get();
    // propagationTime, EigenStepper.ipp:36-47
    const Real dtds = khypot(Real(1.0), st.options.mass / s.p);
    s.t = s.t + h * dtds;
    if (s.covTransport) s.derivative[3] = dtds;

    if (s.covTransport) {
      F32Region f32;
      M88 D = transportMatrix(st, B_first, B_middle, B_last, k1, k2, k3, h);
      s.jacTransport = D * s.jacTransport;
    }

    s.pos = s.pos + h * s.dir + (h * h / 6.0) * (k1 + k2 + k3);
    s.dir = s.dir + (h / 6.0) * (k1 + Real(2.0) * (k2 + k3) + k4);
    Real dn = norm(s.dir);
    s.dir = v3(s.dir[0] / dn, s.dir[1] / dn, s.dir[2] / dn);
    if (s.covTransport) {
      s.derivative[0] = s.dir[0];
      s.derivative[1] = s.dir[1];
      s.derivative[2] = s.dir[2];
      s.derivative[4] = k4[0];
      s.derivative[5] = k4[1];
      s.derivative[6] = k4[2];
    }
    s.pathAccumulated = s.pathAccumulated + h;
    return true;
  }

  // transportMatrix, EigenStepper.ipp:50-124
  M88 transportMatrix(const PropState &st, const V3 &Bf, const V3 &Bm,
                      const V3 &Bl, const V3 &k1, const V3 &k2, const V3 &k3,
                      Real h) const {
    const StepperState &s = st.stepping;
    M88 D = M88::identity();
    const V3 dir = s.dir;
    const Real qop = Real((double)s.q) / s.p;
    const Real half_h = h * 0.5;

    V3 dk1dL = cross(dir, Bf);
    V3 dk2dL = cross(dir + half_h * k1, Bm) + (qop * half_h) * cross(dk1dL, Bm);
    V3 dk3dL = cross(dir + half_h * k2, Bm) + (qop * half_h) * cross(dk2dL, Bm);
    V3 dk4dL = cross(dir + h * k3, Bl) + (qop * h) * cross(dk3dL, Bl);

    M33 dk1dT = M33::zero();
    dk1dT(0, 1) = Bf[2];
    dk1dT(0, 2) = -Bf[1];
    dk1dT(1, 0) = -Bf[2];
    dk1dT(1, 2) = Bf[0];
    dk1dT(2, 0) = Bf[1];
    dk1dT(2, 1) = -Bf[0];
    dk1dT = qop * dk1dT;
    M33 dk2dT = M33::identity() + half_h * dk1dT;
    dk2dT = qop * cross_cols(dk2dT, Bm);
    M33 dk3dT = M33::identity() + half_h * dk2dT;
    dk3dT = qop * cross_cols(dk3dT, Bm);
    M33 dk4dT = M33::identity() + h * dk3dT;
    dk4dT = qop * cross_cols(dk4dT, Bl);

    M33 dFdT = M33::identity() + (h / 6.0) * (dk1dT + dk2dT + dk3dT);
    dFdT = h * dFdT;
    V3 dFdL = ((h * h) / 6.0) * (dk1dL + dk2dL + dk3dL);
    M33 dGdT = M33::identity() + (h / 6.0) * (dk1dT + Real(2.0) * (dk2dT + dk3dT) + dk4dT);
    V3 dGdL = (h / 6.0) * (dk1dL + Real(2.0) * (dk2dL + dk3dL) + dk4dL);
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) {
        D(i, 4 + j) = dFdT(i, j);
        D(4 + i, 4 + j) = dGdT(i, j);
      }
      D(i, 7) = dFdL[i];
      D(4 + i, 7) = dGdL[i];
    }
    const Real m = st.options.mass;
    D(3, 7) = h * m * m * Real((double)s.q) /
              (s.p * khypot(Real(1.0), m / s.p));
    return D;
  }

  // detail::covarianceTransport + reinitializeJacobians,
  // CovarianceEngine.hpp:106-125, 169-201, 269-290
  void covarianceTransport(StepperState &s, const V8 &freePars,
                           const Surface &surf) const {
    {
      F32Region f32;
      s.jacToGlobal = s.jacTransport * s.jacToGlobal;
    }
    M68 jacToLocal = M68::zero();
    V3 pos = v3(freePars[0], freePars[1], freePars[2]);
    V3 dir = v3(freePars[4], freePars[5], freePars[6]);
    M33 rft = initJacobianToLocal(surf, jacToLocal, pos, dir);
    R6 sVec = derivativeFactors(surf, pos, dir, rft, s.jacToGlobal);
    Mat<8, 1> dv = s.derivative;
    {
      F32Region f32;
      s.jacToGlobal = s.jacToGlobal - dv * sVec;
      s.jacobian = jacToLocal * s.jacToGlobal;
      s.cov = s.jacobian * s.cov * transpose(s.jacobian);
    }
    // reinitializeJacobians
    s.jacTransport = M88::identity();
    s.derivative = V8::zero();
    s.jacToGlobal = M86::zero();
    V2 loc = globalToLocal(surf, pos, dir);
    V6 pars;
    pars[0] = loc[0];
    pars[1] = loc[1];
    pars[2] = vphi(dir);
    pars[3] = vtheta(dir);
    pars[4] = freePars[7];
    pars[5] = freePars[3];
    initJacobianToGlobal(surf, s.jacToGlobal, pos, dir, pars);
  }

  // EigenStepper::boundState, EigenStepper.ipp:516-540 -> detail::boundState,
  // CovarianceEngine.hpp:345-371
  void boundState(StepperState &s, const Surface &surf, BoundParams &out,
                  M66 &jacobian, Real &path) const {
    V8 fp;
    fp[0] = s.pos[0];
    fp[1] = s.pos[1];
    fp[2] = s.pos[2];
    fp[3] = s.t;
    fp[4] = s.dir[0];
    fp[5] = s.dir[1];
    fp[6] = s.dir[2];
    fp[7] = Real((double)s.q) / s.p;
    M66 cov = M66::zero();
    if (s.covTransport) {
      covarianceTransport(s, fp, surf);
      cov = s.cov;
    }
    V3 position = v3(fp[0], fp[1], fp[2]);
    Real pabs = kabs(Real(1.0) / fp[7]);
    V3 momentum = v3(pabs * fp[4], pabs * fp[5], pabs * fp[6]);
    Real charge = kcopysign(Real(1.0), fp[7]);
    out = makeBoundGlobal(cov, position, momentum, charge, fp[3], &surf);
    jacobian = s.jacobian;
    path = s.pathAccumulated;
  }

  // EigenStepper::update (free parameters + covariance), EigenStepper.ipp:554-564
  void update(StepperState &s, const V8 &f, const M66 &cov) const {
    s.pos = v3(f[0], f[1], f[2]);
    s.dir = normalized(v3(f[4], f[5], f[6]));
    s.p = kabs(Real(1.0) / f[7]);
    s.t = f[3];
    s.cov = cov;
  }

  // EigenStepper::update (pos, dir, p, t), EigenStepper.ipp:566-575
  void update(StepperState &s, const V3 &pos, const V3 &dir, Real p, Real t) const {
    s.pos = pos;
    s.dir = dir;
    s.p = p;
    s.t = t;
  }

  // detail::updateSingleSurfaceStatus, SteppingHelper.hpp:31-65
  IStatus updateSurfaceStatus(StepperState &s, const Surface &surf, Real overstepLimit) const {
    V3 d = Real((double)s.navDir) * s.dir;
    Intersection in = intersect(surf, s.pos, d);
    if (in.status == IStatus::OnSurface) {
      s.stepSize.release(kActor);
      return IStatus::OnSurface;
    } else if (in.status != IStatus::Unreachable) {
      Real pLimit = s.stepSize.values[kAborter];
      Real oLimit = -overstepLimit;
      Real cLimit = in.pathLength;
      bool accept = (cLimit > oLimit && cLimit * cLimit < pLimit * pLimit);
      if (accept) {
        // setStepSize, EigenStepper.hpp:243-248
        s.previousStepSize = s.stepSize.get();
        s.stepSize.update(Real((double)s.navDir) * cLimit, kActor, true);
        return IStatus::Reachable;
      }
    }
    return IStatus::Unreachable;
  }

 private:
  const Field *field_;
};

// -------------------------------------------------------------- navigator ----
// DirectNavigator::status / target, include/Propagator/DirectNavigator.hpp:158-225
struct Navigator {
  void status(PropState &st, const Stepper &stepper) const {
    NavState &n = st.navigation;
    n.currentSurface = nullptr;
    if (n.next != n.size) {
      const Surface *sp = n.sequence[n.next];
      if (stepper.updateSurfaceStatus(st.stepping, *sp, st.options.overstepLimit) ==
          IStatus::OnSurface) {
        n.currentSurface = sp;
        ++n.next;
        if (n.next != n.size) st.stepping.stepSize.release(kActor);
      }
    }
  }
  void target(PropState &st, const Stepper &stepper) const {
    NavState &n = st.navigation;
    n.currentSurface = nullptr;
    if (n.next != n.size) {
      const Surface *sp = n.sequence[n.next];
      if (stepper.updateSurfaceStatus(st.stepping, *sp, st.options.overstepLimit) ==
          IStatus::Unreachable)
        ++n.next;
    } else {
      n.navigationBreak = true;
      if (n.targetSurface == nullptr) n.targetReached = true;
    }
  }
};

// -------------------------------------------------------------- aborters ----
// PathLimitReached, StandardAborters.hpp:24-57
inline bool pathLimitReached(PropState &st) {
  if (st.navigation.targetReached) return true;
  Real distance = Real((double)st.stepping.navDir) * Real(DBL_MAX) -
                  st.stepping.pathAccumulated;
  Real tol = st.options.targetTolerance;
  st.stepping.stepSize.update(distance, kAborter);
  bool limitReached = (distance * distance < tol * tol);
  if (limitReached) st.navigation.targetReached = true;
  return limitReached;
}

// SurfaceReached, StandardAborters.hpp:61-133
inline bool surfaceReached(PropState &st, const Surface &target) {
  NavState &n = st.navigation;
  if (n.targetReached) return true;
  if (n.currentSurface && n.currentSurface == &target) {
    n.targetReached = true;
    return true;
  }
  StepperState &s = st.stepping;
  Intersection in = intersect(target, s.pos, Real((double)s.navDir) * s.dir);
  bool reached = (in.status == IStatus::OnSurface);
  Real distance = in.pathLength;
  if (reached) {
    n.currentSurface = &target;
    n.targetReached = true;
  } else {
    s.stepSize.update(Real((double)s.navDir) * distance, kAborter);
  }
  return reached;
}

// ------------------------------------------------------------- material ----
// Interactions.ipp:27-52 RelativisticQuantities
struct RelQ {
  Real q2OverBeta2, beta2, betaGamma, gamma;
  RelQ(Real mass, Real qOverP, Real q) {
    q2OverBeta2 = q * q + (mass * qOverP) * (mass * qOverP);
    Real mOverP = mass * kabs(qOverP / q);
    Real pOverM = Real(1.0f) / mOverP;
    beta2 = Real(1.0f) / (Real(1.0f) + mOverP * mOverP);
    betaGamma = pOverM;
    gamma = ksqrt(Real(1.0f) + pOverM * pOverM);
  }
};

// computeEnergyLossBethe, Interactions.ipp:150-176
inline Real energyLossBethe(const MaterialSlab &slab, Real m, Real qOverP, Real q) {
  if (!slab) return 0.0f;
  const Real I = slab.material.meanExcitationEnergy();
  const Real Ne = slab.material.molarElectronDensity();
  const Real thickness = slab.thickness;
  RelQ rq(m, qOverP, q);
  const Real eps = Real(0.5f) * Real(kK) * Ne * thickness * rq.q2OverBeta2;
  Real dhalf = 0.0f;
  if (!(rq.betaGamma < Real(10.0f))) {
    const Real plasmaEnergy = Real(kPlasmaEnergyScale) * ksqrt(Ne);
    dhalf = klog(rq.betaGamma) + klog(plasmaEnergy / I) - Real(0.5f);
  }
  const Real u = Real(2.0) * Real(kMe) * rq.betaGamma * rq.betaGamma;
  const Real mfrac = Real(kMe) / m;
  const Real nominator = Real(2.0) * Real(kMe) * rq.betaGamma * rq.betaGamma;
  const Real denominator = Real(1.0f) + Real(2.0) * rq.gamma * mfrac + mfrac * mfrac;
  const Real wmax = nominator / denominator;
  const Real running = Real(0.5f) * klog(u / I) + Real(0.5f) * klog(wmax / I) -
                       rq.beta2 - dhalf;
  return eps * running;
}

// computeEnergyLossLandauSigmaQOverP, Interactions.ipp:316-342; the FWHM to
// sigma factor uses the float log(2.0f) of the reference.
inline Real landauSigmaQOverP(const MaterialSlab &slab, Real m, Real qOverP, Real q) {
  if (!slab) return 0.0f;
  const Real Ne = slab.material.molarElectronDensity();
  const Real thickness = slab.thickness;
  RelQ rq(m, qOverP, q);
  const Real eps = Real(0.5f) * Real(kK) * Ne * thickness * rq.q2OverBeta2;
  const Real fwhm = Real(4.0) * eps;
  const float conv = 2 * std::sqrt(2 * std::log(2.0f));
  const Real sigmaE = fwhm / Real((double)conv);
  const Real pInv = qOverP / q;
  return ksqrt(rq.q2OverBeta2) * pInv * pInv * sigmaE;
}

// computeMultipleScatteringTheta0 (Highland), Interactions.ipp:501-546
inline Real multipleScatteringTheta0(const MaterialSlab &slab, Real m, Real qOverP,
                                     Real q) {
  if (!slab) return 0.0f;
  const Real xOverX0 = slab.thicknessInX0;
  const Real momentumInv = kabs(qOverP / q);
  const Real q2OverBeta2 = RelQ(m, qOverP, q).q2OverBeta2;
  const Real t = ksqrt(xOverX0 * q2OverBeta2);
  const float c = 0.038f * 2;
  return Real(kHighland) * momentumInv * t * (Real(1.0f) + Real((double)c) * klog(t));
}

// ---------------------------------------------------------- the fitter ----
struct KalmanFitter {
  Stepper stepper;
  Navigator navigator;

  explicit KalmanFitter(const Field *f) : stepper(f) {}

  // GainMatrixUpdater::operator(), GainMatrixUpdater.hpp:34-106
  bool gainMatrixUpdate(TrackStateRec &ts, const Surface *surface, double outlierCut) const {
    const BoundParams &pred = ts.predicted;
    const M66 &P = pred.cov;
    Mat<2, 6> H = Mat<2, 6>::zero();
    H(0, 0) = 1.0;
    H(1, 1) = 1.0;
    M22 cov = H * P * transpose(H) + ts.sl.cov;
    // get2DMatrixInverse, Helpers.hpp:387-399
    Real det = cov(0, 0) * cov(1, 1) - cov(0, 1) * cov(1, 0);
    M22 covInv = M22::zero();
    covInv(0, 0) = cov(1, 1) / det;
    covInv(0, 1) = -cov(0, 1) / det;
    covInv(1, 0) = -cov(1, 0) / det;
    covInv(1, 1) = cov(0, 0) / det;
    Mat<6, 2> K = P * transpose(H) * covInv;
    V2 res;
    res[0] = ts.sl.meas[0] - pred.par[0];
    res[1] = ts.sl.meas[1] - pred.par[1];
    V6 gain = K * res;
    V6 fpar = pred.par + gain;
    M66 KH = K * H;
    M66 C = M66::identity() - KH;
    M66 fcov = C * P;
    ts.filtered = makeBound(fcov, fpar, surface);
    // chi2 of the predicted residual (this build's definition)
    ts.chi2 = res[0] * (covInv(0, 0) * res[0] + covInv(0, 1) * res[1]) +
              res[1] * (covInv(1, 0) * res[0] + covInv(1, 1) * res[1]);
    // the outlier finder's decision: an outlier keeps its prediction
    if (outlierCut > 0.0 && kval(ts.chi2) > outlierCut) {
      ts.filtered = makeBound(P, pred.par, surface);
      ts.outlier = true;
    }
    return true;
  }

  // GainMatrixSmoother::operator(), GainMatrixSmoother.hpp:21-94; runs over
  // the first n filled states.  Returns the first state's smoothed params.
  const BoundParams *smooth(std::vector<TrackStateRec> &states, unsigned n,
                            bool &nanFound) const {
    TrackStateRec *prev = &states[n - 1];
    prev->smoothed = prev->filtered;
    for (int i = (int)n - 2; i >= 0; --i) {
      TrackStateRec &ts = states[i];
      M66 G = ts.filtered.cov * transpose(prev->jacobian) * inverse(prev->predicted.cov);
      if (has_nan(G)) {
        nanFound = true;
        return nullptr;
      }
      V6 prevDiff = prev->smoothed.par - prev->predicted.par;
      V6 gainPars = G * prevDiff;
      V6 sp = ts.filtered.par + gainPars;
      M66 sc = ts.filtered.cov -
               G * (prev->predicted.cov - prev->smoothed.cov) * transpose(G);
      ts.smoothed = makeBound(sc, sp, ts.filtered.surface);
      prev = &ts;
    }
    return &prev->smoothed;
  }

  // Actor::materialInteractor, KalmanFitter.hpp:405-449, with
  // PointwiseMaterialInteraction (.hpp:66-148, .ipp:14-53)
  void materialInteractor(const Surface *surface, PropState &st,
                          const FitOptions &fo) const {
    if (!(surface && surface->slab)) return;
    StepperState &s = st.stepping;
    const V3 pos = s.pos;
    const Real time = s.t;
    const V3 dir = s.dir;
    const Real momentum = s.p;
    const Real q = Real((double)s.q);
    const Real qOverP = q / momentum;
    const Real mass = st.options.mass;
    const int nav = s.navDir;
    // evaluateMaterialSlab, PointwiseMaterialInteraction.hpp:87-108: the start
    // surface gets postUpdate, the target surface preUpdate, others
    // fullUpdate; ISurfaceMaterial::factor (ISurfaceMaterial.hpp:126-134)
    // with the HomogeneousSurfaceMaterial default split factor 1.
    int stage = 0;
    if (surface == st.navigation.startSurface)
      stage = 1;
    else if (surface == st.navigation.targetSurface)
      stage = -1;
    MaterialSlab slab = st.navigation.currentSurface->slab;
    if (slab) {
      const double split = 1.0;
      const double factor = (stage == 0) ? 1.0 : ((nav * stage > 0) ? split : 1.0 - split);
      if (factor == 0.0)
        slab = MaterialSlab();
      else
        slab.scaleThickness(Real(factor));
    }
    Real Eloss = 0.0, variancePhi = 0.0, varianceTheta = 0.0, varianceQoverP = 0.0;
    {
      F32Region f32;  // the material formulas, not the state update below
      Real pc = pathCorrection(*surface, dir);
      slab.scaleThickness(pc);
      if (!slab) return;
      if (fo.energyLoss) Eloss = energyLossBethe(slab, mass, qOverP, q);
      if (s.covTransport) {
        if (fo.multipleScattering) {
          const Real theta0 = multipleScatteringTheta0(slab, mass, qOverP, q);
          const Real sigmaPhi = theta0 * (norm(dir) / ksqrt(dir[0] * dir[0] + dir[1] * dir[1]));
          variancePhi = sigmaPhi * sigmaPhi;
          varianceTheta = theta0 * theta0;
        }
        if (fo.energyLoss) {
          const Real sq = landauSigmaQOverP(slab, mass, qOverP, q);
          varianceQoverP = sq * sq;
        }
      }
    }
    // updateState
    const Real nextE = ksqrt(mass * mass + momentum * momentum) -
                       kcopysign(Eloss, Real((double)nav));
    const Real nextP = (mass < nextE) ? ksqrt(nextE * nextE - mass * mass) : Real(0.0);
    stepper.update(s, pos, dir, nextP, time);
    const Real mode = (nav == 1) ? 1.0 : -1.0;
    s.cov(2, 2) = kmax(Real(0.0), s.cov(2, 2) + kcopysign(variancePhi, mode));
    s.cov(3, 3) = kmax(Real(0.0), s.cov(3, 3) + kcopysign(varianceTheta, mode));
    s.cov(4, 4) = kmax(Real(0.0), s.cov(4, 4) + kcopysign(varianceQoverP, mode));
  }

  // Actor::filter, KalmanFitter.hpp:284-350
  bool filter(const Surface *surface, PropState &st, FitResult &r,
              const std::vector<SourceLink> &inputs, const FitOptions &fo) const {
    int found = -1;
    for (size_t i = 0; i < inputs.size(); ++i)
      if (inputs[i].geoId == surface->geoId) {
        found = (int)i;
        break;
      }
    if (found < 0) return true;
    if (r.measurementStates >= r.states.size()) return false;
    TrackStateRec &ts = r.states[r.measurementStates];
    ts = TrackStateRec();
    ts.sl = inputs[found];
    stepper.boundState(st.stepping, *surface, ts.predicted, ts.jacobian, ts.pathLength);
    if (!gainMatrixUpdate(ts, surface, fo.outlierChi2Cut)) return false;
    V8 fp = boundToFree(ts.filtered.par, *ts.filtered.surface);
    stepper.update(st.stepping, fp, ts.filtered.cov);
    materialInteractor(surface, st, fo);
    ++r.measurementStates;
    return true;
  }

  // Actor::finalize, KalmanFitter.hpp:360-393
  bool finalize(PropState &st, FitResult &r) const {
    kfo_mark_backward();
    r.smoothed = true;
    bool nanFound = false;
    const BoundParams *sp = smooth(r.states, r.measurementStates, nanFound);
    if (nanFound) r.smootherNaN = true;
    if (sp) {
      V8 fp = boundToFree(sp->par, *sp->surface);
      stepper.update(st.stepping, fp, sp->cov);
      st.stepping.navDir = -1;
      st.stepping.stepSize =
          ConstrainedStep(Real((double)st.stepping.navDir) * kabs(st.options.maxStepSize));
      st.stepping.pathAccumulated = 0.0;
      return true;
    }
    return false;
  }

  // Actor::operator(), KalmanFitter.hpp:216-273
  void actor(PropState &st, FitResult &r, const std::vector<SourceLink> &inputs,
             const FitOptions &fo) const {
    const Surface *surface = st.navigation.currentSurface;
    if (surface != nullptr && !r.smoothed && !r.finished) {
      if (!filter(surface, st, r, inputs, fo)) r.result = false;
    }
    if ((r.measurementStates == inputs.size() ||
         (r.measurementStates > 0 && st.navigation.navigationBreak)) &&
        !r.smoothed && !r.finished) {
      if (!fo.smoothing) {
        r.finished = true;
      } else {
        if (!finalize(st, r)) r.result = false;
      }
    }
    if (r.smoothed && !r.finished && surfaceReached(st, *fo.referenceSurface)) {
      M66 jac;
      Real path;
      stepper.boundState(st.stepping, *fo.referenceSurface, r.fittedParameters, jac, path);
      r.finished = true;
    }
  }

  // Aborter, KalmanFitter.hpp:682-694
  static bool kalmanAborter(const FitResult &r) { return !r.result || r.finished; }

  // Propagator::propagate, Propagator.ipp:3-96
  void propagate(const BoundParams &start, const PropOptions &po,
                 const Surface *const *seq, unsigned nSeq, const FitOptions &fo,
                 const std::vector<SourceLink> &inputs, FitResult &r) const {
    PropState st;
    st.options = po;
    st.stepping = stepper.makeState(start, 1, po.maxStepSize, po.tolerance);
    st.navigation.startSurface = start.surface;
    // DirectNavigatorInitializer, DirectNavigator.hpp:49-61
    st.navigation.sequence = seq;
    st.navigation.size = nSeq;
    st.navigation.targetSurface = fo.referenceSurface;

    navigator.status(st, stepper);
    actor(st, r, inputs, fo);
    bool terminatedNormally = false;
    unsigned steps = 0;
    if (!kalmanAborter(r) && !pathLimitReached(st)) {
      navigator.target(st, stepper);
      for (; steps < po.maxSteps; ++steps) {
        stepper.step(st);
        navigator.status(st, stepper);
        actor(st, r, inputs, fo);
        if (kalmanAborter(r) || pathLimitReached(st)) {
          terminatedNormally = true;
          break;
        }
        navigator.target(st, stepper);
      }
    }
    if (!terminatedNormally) {
      st.navigation.navigationBreak = true;
      r.maxStepsReached = (steps >= po.maxSteps);
    }
    actor(st, r, inputs, fo);
    r.navBreak = st.navigation.navigationBreak;
    r.stepFailed = st.stepping.stepFailed;
    r.steps = steps;
  }

  // KalmanFitter::fit, KalmanFitter.hpp:716-768 (success: every input
  // measurement was used; see the header comment)
  bool fit(const std::vector<SourceLink> &inputs, const BoundParams &start,
           const FitOptions &fo, const PropOptions &po, FitResult &r,
           const Surface *const *seq, unsigned nSeq) const {
    propagate(start, po, seq, nSeq, fo, inputs, r);
    if (!r.result || r.measurementStates != inputs.size()) return false;
    return true;
  }
};

}  // namespace kfo
