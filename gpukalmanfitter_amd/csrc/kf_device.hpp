// kf_device.hpp -- device-side building blocks of the MI355X Kalman fitter.
//
// One track per lane.  Every quantity of a track lives in VGPRs; the only
// memory a lane touches besides its inputs and outputs is its own state store
// (predicted / filtered / Jacobian per measured surface) that the RTS
// smoother reads back.
//
// Algorithm: the gpuKalmanFitter fit (KalmanFitter.hpp:216-393,716-768 with
// EigenStepper, DirectNavigator, GainMatrixUpdater, GainMatrixSmoother,
// PointwiseMaterialInteraction).  The per-lane operation sequence is the
// reference's; the data layout is ours:
//  * the free-from-bound Jacobian is accumulated directly, J <- D J, on its
//    25 non-constant entries (the reference keeps jacTransport = D_n..D_1 and
//    multiplies once at the surface; same matrix, fewer flops);
//  * covariances are kept as the 21-entry upper triangle;
//  * plane and line surfaces are special-cased, no virtual dispatch.
// The scalar type of the stepping/transport part is a template parameter
// (double for the plain FP64 fitter, float for the mixed-precision one);
// positions, directions, momentum, time and the Kalman update are double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/kf_abi.h"

namespace kfd {

#define KFD_INLINE __host__ __device__ __forceinline__

// Keeps the scheduler from hoisting the next phase's memory loads above this
// point (bounds the number of live registers in load-heavy phases).
#define KFD_PHASE_FENCE() __asm__ volatile("" ::: "memory")

// --------------------------------------------------------- lane groups ----
// A track may be fitted by a group of L = 1, 2 or 4 consecutive lanes of a
// quad (the counterpart of the reference's block-per-track kernel): every
// lane of a group runs the track's program on the same data, so a group never
// diverges, and the work of a few matrix products is split by rank (the lane's
// index in its group) and gathered back.  groupGet<L, SRC>(v) is the v of
// rank SRC of the caller's group: one DPP quad_perm move per 32 bits.
template <int L, int SRC>
KFD_INLINE int groupGet32(int v) {
  static_assert(L == 1 || L == 2 || L == 4, "lane groups of 1, 2 or 4 lanes");
  static_assert(SRC >= 0 && SRC < L, "source rank inside the group");
  if constexpr (L == 1) {
    return v;
  } else {
#if defined(__HIP_DEVICE_COMPILE__)
    // quad_perm: lane j of a quad reads lane perm[j]; groups of two lanes
    // are (0, 1) and (2, 3)
    constexpr int base2 = SRC | (SRC << 2) | ((2 | SRC) << 4) | ((2 | SRC) << 6);
    constexpr int ctrl = (L == 2) ? base2 : SRC * 0x55;
    return __builtin_amdgcn_update_dpp(v, v, ctrl, 0xF, 0xF, false);
#else
    return v;  // host build: groups of one lane only
#endif
  }
}
template <int L, int SRC>
KFD_INLINE float groupGet(float v) {
  return __builtin_bit_cast(float, groupGet32<L, SRC>(__builtin_bit_cast(int, v)));
}
template <int L, int SRC>
KFD_INLINE double groupGet(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const int lo = groupGet32<L, SRC>((int)(uint32_t)u);
  const int hi = groupGet32<L, SRC>((int)(uint32_t)(u >> 32));
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
// out[c] = the value of slot c / L of rank c % L, for c < N: the gather of N
// values computed in slots (rank r holds values r, r + L, r + 2L, ... in
// slot[0], slot[1], ...).
template <int L, int N, typename T>
KFD_INLINE void groupGather(const T *slot, T *out) {
#pragma unroll
  for (int c = 0; c < N; ++c) {
    if constexpr (L == 1) {
      out[c] = slot[c];
    } else if (c % L == 0) {
      out[c] = groupGet<L, 0>(slot[c / L]);
    } else if (c % L == 1) {
      out[c] = groupGet<L, (1 < L) ? 1 : 0>(slot[c / L]);
    } else if (c % L == 2) {
      out[c] = groupGet<L, (2 < L) ? 2 : 0>(slot[c / L]);
    } else {
      out[c] = groupGet<L, (3 < L) ? 3 : 0>(slot[c / L]);
    }
  }
}

constexpr double kOnSurfaceTolerance = 1e-4;
constexpr double kOverstepLimit = 0.01;
constexpr double kPi = 3.14159265358979323846;
constexpr double kDblMax = 1.7976931348623157e308;

// ---------------------------------------------------------------- plane ----
// Device plane table entry (built on the host from kf_plane).
struct DevPlane {
  double c[3];
  double r[9];           // row-major rotation, columns U, V, normal
  double thickness;      // <= 0: no material
  double thicknessInX0;  // precomputed MaterialSlab ratios
  double meanExcitation; // 16 eV * Z^0.9f
  double molarElectronDensity;
  int32_t geo_id;
  int32_t has_material;
};

// ---------------------------------------------------------------- field ----
struct DevField {
  int grid;
  double b[3];
  double scale;
  const float *values;
  int n[3];
  double lo[3], hi[3], w[3];
  double iw[3];  // 1 / w: cell index and fraction by multiplication
};

// Field kinds a kernel can be built for: either (chosen per lookup by
// DevField::grid), the constant field only, the grid only.
constexpr int kFieldAny = 0, kFieldConst = 1, kFieldGrid = 2;

// Field lookup in double.  Constant field: the vector times the scale.
KFD_INLINE void fieldConst(const DevField &f, double B[3]) {
  if (f.scale == 1.0) {
    B[0] = f.b[0];
    B[1] = f.b[1];
    B[2] = f.b[2];
  } else {
    B[0] = f.scale * f.b[0];
    B[1] = f.scale * f.b[1];
    B[2] = f.scale * f.b[2];
  }
}

// Grid: trilinear, last axis first (the reference's
// interpolation_impl.hpp:98-131), positions clamped to the box.
KFD_INLINE void fieldGrid(const DevField &f, double x, double y, double z, double B[3]) {
  const double p[3] = {x, y, z};
  int idx[3];
  double fr[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double v = p[d];
    if (v < f.lo[d]) v = f.lo[d];
    if (f.hi[d] < v) v = f.hi[d];
    int i = (int)floor((v - f.lo[d]) * f.iw[d]);
    i = i < 0 ? 0 : i;
    i = i > f.n[d] - 2 ? f.n[d] - 2 : i;
    idx[d] = i;
    const double ll = f.lo[d] + (double)i * f.w[d];
    fr[d] = (v - ll) * f.iw[d];
  }
  // the four (x, y) corner columns: six floats each (the z pair), all
  // loads issued before any is used, so one memory latency per lookup
  float v[4][6];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int ix = idx[0] + (c >> 1), iy = idx[1] + (c & 1);
    const size_t b0 = (((size_t)ix * f.n[1] + iy) * f.n[2] + idx[2]) * 3;
#pragma unroll
    for (int k = 0; k < 6; ++k) v[c][k] = f.values[b0 + k];
  }
  double cz[4][3];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      cz[c][k] = (1.0 - fr[2]) * (double)v[c][k] + fr[2] * (double)v[c][3 + k];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double y0 = (1.0 - fr[1]) * cz[0][k] + fr[1] * cz[1][k];
    const double y1 = (1.0 - fr[1]) * cz[2][k] + fr[1] * cz[3][k];
    B[k] = (1.0 - fr[0]) * y0 + fr[0] * y1;
    if (f.scale != 1.0) B[k] = f.scale * B[k];
  }
}

// Either kind, chosen by f.grid at run time.
KFD_INLINE void fieldAt(const DevField &f, double x, double y, double z, double B[3]) {
  if (!f.grid)
    fieldConst(f, B);
  else
    fieldGrid(f, x, y, z, B);
}

// Field and its spatial derivative dB[k][axis] = dB_k / dx_axis (the chi2
// adjoint).  Same interpolation as fieldAt; an axis clamped to the box edge
// has zero derivative.
KFD_INLINE void fieldAtGrad(const DevField &f, const double x[3], double B[3], double dB[3][3]) {
  if (!f.grid) {
    fieldAt(f, x[0], x[1], x[2], B);
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int d = 0; d < 3; ++d) dB[k][d] = 0.0;
    return;
  }
  int idx[3];
  double fr[3], dfr[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double v = x[d];
    bool clamped = false;
    if (v < f.lo[d]) {
      v = f.lo[d];
      clamped = true;
    }
    if (f.hi[d] < v) {
      v = f.hi[d];
      clamped = true;
    }
    int i = (int)floor((v - f.lo[d]) * f.iw[d]);
    i = i < 0 ? 0 : i;
    i = i > f.n[d] - 2 ? f.n[d] - 2 : i;
    idx[d] = i;
    const double ll = f.lo[d] + (double)i * f.w[d];
    fr[d] = (v - ll) * f.iw[d];
    dfr[d] = clamped ? 0.0 : f.iw[d];
  }
  const double sc = f.scale;
  // every corner value loaded before any is used (one memory latency):
  // column (a, b) holds the z pair, six floats
  float v[4][6];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const size_t o = ((((size_t)(idx[0] + a)) * f.n[1] + (idx[1] + b)) * f.n[2] + idx[2]) * 3;
#pragma unroll
      for (int m = 0; m < 6; ++m) v[2 * a + b][m] = f.values[o + m];
    }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double val[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) val[a][b][c] = (double)v[2 * a + b][3 * c + k];
    const double wx[2] = {1.0 - fr[0], fr[0]}, wy[2] = {1.0 - fr[1], fr[1]},
                 wz[2] = {1.0 - fr[2], fr[2]};
    double s = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const double v = val[a][b][c];
          s += wx[a] * wy[b] * wz[c] * v;
          gx += (a ? 1.0 : -1.0) * wy[b] * wz[c] * v;
          gy += wx[a] * (b ? 1.0 : -1.0) * wz[c] * v;
          gz += wx[a] * wy[b] * (c ? 1.0 : -1.0) * v;
        }
    B[k] = sc * s;
    dB[k][0] = sc * gx * dfr[0];
    dB[k][1] = sc * gy * dfr[1];
    dB[k][2] = sc * gz * dfr[2];
  }
}

// ---------------------------------------------- out-of-line library math ----
// The library double sin / cos / atan2 / log as calls rather than inline
// polynomial evaluations: in the chi2 adjoint, whose register use is at the
// 512-register limit, this lowers the spills (measured 7 % faster); the fit
// kernels keep the inline forms (measured faster there).  The same library
// functions, so the same results.
struct SinCos {
  double s, c;
};
__host__ __device__ __attribute__((noinline)) inline SinCos oolSinCos(double x) {
  return SinCos{sin(x), cos(x)};
}
__host__ __device__ __attribute__((noinline)) inline double oolAtan2(double y, double x) {
  return atan2(y, x);
}
__host__ __device__ __attribute__((noinline)) inline double oolLog(double x) { return log(x); }

// ------------------------------------------------------- small helpers ----
template <typename T>
KFD_INLINE void cross(const T a[3], const T b[3], T r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
template <typename T>
KFD_INLINE T dot3(const T a[3], const T b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// packed upper-triangle index of (i, j)
KFD_INLINE constexpr int sidx(int i, int j) {
  return (i <= j) ? (i * 6 - (i * (i - 1)) / 2 + (j - i)) : (j * 6 - (j * (j - 1)) / 2 + (i - j));
}

// phi cyclic in [-pi, pi), theta clamped to [0, pi], each applied twice as
// ParameterSet::setParameter does (ParameterSet.hpp:276-281).
KFD_INLINE double cyclicPhi(double x) {
  const double mn = -kPi, mx = kPi;
  if (mn <= x && x < mx) return x;
  return x - (mx - mn) * floor((x - mn) / (mx - mn));
}
KFD_INLINE double boundTheta(double x) {
  return (x > kPi) ? kPi : ((x < 0.0) ? 0.0 : x);
}
KFD_INLINE void correctParams(double p[6]) {
  p[2] = cyclicPhi(cyclicPhi(p[2]));
  p[3] = boundTheta(boundTheta(p[3]));
}

// ------------------------------------------------------ ConstrainedStep ----
// include/Propagator/ConstrainedStep.hpp:22-139
enum { kAccuracy = 0, kActor = 1, kAborter = 2, kUser = 3 };
struct CStep {
  double v[4];
  int dir;
  KFD_INLINE void init(double value) {
    dir = (value > 0.0) ? 1 : -1;
    v[kAccuracy] = kDblMax * (double)dir;
    v[kActor] = kDblMax * (double)dir;
    v[kAborter] = kDblMax * (double)dir;
    v[kUser] = value;
  }
  // s = +1: the minimum of the values, from DBL_MAX with a strict
  // comparison; s = -1: their maximum, from -DBL_MAX.  One comparison loop
  // over s * v (negation is exact), so NaNs are skipped and signed zeros kept
  // as by the two loops of the reference's std::min / std::max.
  KFD_INLINE double signedMin(double s) const {
    double m = kDblMax;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double x = s * v[i];
      m = (x < m) ? x : m;
    }
    return s * m;
  }
  KFD_INLINE void release(int t) {
    const double r = signedMin(-(double)dir);  // the most permissive value
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i == t) v[i] = r;
  }
  KFD_INLINE void update(double value, int t) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i == t) v[i] = (fabs(v[i]) < fabs(value)) ? v[i] : value;
  }
  KFD_INLINE void assign(double value) {
    v[kAccuracy] = value;
    dir = (value > 0.0) ? 1 : -1;
  }
  KFD_INLINE double get() const { return signedMin((double)dir); }
};

}  // namespace kfd
