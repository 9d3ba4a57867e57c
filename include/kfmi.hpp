// kfmi.hpp -- C++ facade of the MI355X Kalman fitter over the C ABI
// (kf_abi.h), mirroring the reference's host-side fit interface:
//
//   KalmanFitter<...>::fit(sourcelinks, sParameters, kfOptions, kfResult,
//                          surfaceSequence, surfaceSequenceSize) -> bool
//       include/Fitter/KalmanFitter.hpp:716-768       -> kfmi::KalmanFitter::fit
//   KalmanFitterOptions{referenceSurface, smoothing, multipleScattering,
//                       energyLoss}  KalmanFitter.hpp:48-96 -> kfmi::KalmanFitterOptions
//   KalmanFitterResult{fittedStates, fittedParameters, measurementStates,
//                      smoothed, finished, result}  KalmanFitter.hpp:98-120
//                                                    -> kfmi::KalmanFitterResult
//   PixelSourceLink{values, cov, geometryId}  PixelSourceLink.hpp:22-75
//                                                    -> kfmi::PixelSourceLink
//   TrackState{sourceLink, parameter.{predicted, filtered, smoothed,
//              jacobian, pathLength}}  TrackState.hpp:38-169 -> kfmi::TrackState
//   fitKernelThreadPerTrack launch  Run/GPU/Kernels.cu:8-15
//                                                    -> kfmi::KalmanFitter::fitBatch
//
// The per-track fit() stages one track through device memory owned by the
// fitter and runs the same HIP kernels as a batch; fitBatch() takes caller
// device buffers and is stream-ordered.  Setup failures throw
// std::runtime_error; fit() returns the reference's bool and never throws for
// a track that does not fit (the reason is in KalmanFitterResult::status).
//
// Host code only: needs the HIP runtime headers (-D__HIP_PLATFORM_AMD__
// -I/opt/rocm/include) and links libkfmi.so and libamdhip64.so.
#pragma once

#include <hip/hip_runtime_api.h>

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "kf_abi.h"

namespace kfmi {

using Vector6 = std::array<double, 6>;
using SymMatrix6 = std::array<std::array<double, 6>, 6>;

// 2-D measurement on a plane (PixelSourceLink.hpp:22-75).
struct PixelSourceLink {
  std::array<double, 2> values{};
  std::array<std::array<double, 2>, 2> cov{};
  int32_t geometryId = 0;
};

// The z-parallel perigee line through `center` (LineSurface.ipp:19-21, built
// at the truth vertex by Run/Common/Processor.hpp:121-129).
struct LineSurface {
  std::array<double, 3> center{};
};

// (loc0, loc1, phi, theta, q/p, t) with covariance, on the surface the
// caller associates with them.
struct BoundParameters {
  Vector6 parameters{};
  SymMatrix6 covariance{};
};

struct KalmanFitterOptions {
  const LineSurface *referenceSurface = nullptr;  // where the fitted parameters are expressed
  bool smoothing = true;
  bool multipleScattering = true;
  bool energyLoss = true;
  // PropagatorOptions of the reference (Propagator.hpp:52-111) and this
  // build's precision / overstep settings; kf_options_default() values.
  kf_options propagator = defaults();

  static kf_options defaults() {
    kf_options o;
    kf_options_default(&o);
    return o;
  }
};

struct TrackState {
  PixelSourceLink sourceLink;
  BoundParameters predicted, filtered, smoothed;
  std::array<std::array<double, 6>, 6> jacobian{};  // bound-to-bound from the previous state
  double pathLength = 0.0;
  double chi2 = 0.0;  // predicted-residual chi2 of this state
};

struct KalmanFitterResult {
  std::vector<TrackState> fittedStates;
  BoundParameters fittedParameters;  // on kfOptions.referenceSurface
  size_t measurementStates = 0;
  bool smoothed = false;
  bool finished = false;
  bool result = true;
  uint32_t status = 0;  // KF_STATUS_* bits
  double chi2 = 0.0;
  int32_t ndf = 0;
};

// Magnetic field description handed to the device context.
struct MagneticField {
  bool grid = false;
  std::array<double, 3> b{};  // constant field [GeV/(e mm)]
  const float *values = nullptr;  // grid, [nx][ny][nz][3]
  std::array<int32_t, 3> n{};
  std::array<double, 3> lo{}, hi{};

  static MagneticField constant(double bx, double by, double bz) {
    MagneticField f;
    f.b = {bx, by, bz};
    return f;
  }
  static MagneticField interpolated(const float *values, int32_t nx, int32_t ny, int32_t nz,
                                    std::array<double, 3> lo, std::array<double, 3> hi) {
    MagneticField f;
    f.grid = true;
    f.values = values;
    f.n = {nx, ny, nz};
    f.lo = lo;
    f.hi = hi;
    return f;
  }
};

inline void packSym(const SymMatrix6 &m, double *out21, size_t stride = 1) {
  size_t k = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) out21[(k++) * stride] = m[i][j];
}
inline SymMatrix6 unpackSym(const double *in21, size_t stride = 1) {
  SymMatrix6 m{};
  size_t k = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) {
      m[i][j] = m[j][i] = in21[k * stride];
      ++k;
    }
  return m;
}

class KalmanFitter {
 public:
  // `planes` is the plane table every surface sequence points into.
  KalmanFitter(std::vector<kf_plane> planes, const MagneticField &field, int device = 0)
      : planes_(std::move(planes)), device_(device) {
    check(kf_ctx_create(device, &ctx_), "kf_ctx_create");
    check(kf_set_geometry(ctx_, planes_.data(), (int32_t)planes_.size()), "kf_set_geometry");
    if (field.grid)
      check(kf_set_field_grid(ctx_, field.values, field.n[0], field.n[1], field.n[2],
                              field.lo.data(), field.hi.data()),
            "kf_set_field_grid");
    else
      check(kf_set_field_constant(ctx_, field.b[0], field.b[1], field.b[2]),
            "kf_set_field_constant");
  }
  KalmanFitter(const KalmanFitter &) = delete;
  KalmanFitter &operator=(const KalmanFitter &) = delete;
  ~KalmanFitter() {
    (void)hipSetDevice(device_);
    for (void *p : stage_) (void)hipFree(p);
    kf_ctx_destroy(ctx_);
  }

  const std::vector<kf_plane> &planes() const { return planes_; }
  kf_ctx *context() const { return ctx_; }
  void setFieldScale(double s) { check(kf_set_field_scale(ctx_, s), "kf_set_field_scale"); }

  // Per-track fit.  `surfaceSequence` points into planes() (nullptr: the whole
  // table); the start parameters live on kfOptions.referenceSurface.
  bool fit(const std::vector<PixelSourceLink> &sourcelinks, const BoundParameters &sParameters,
           const KalmanFitterOptions &kfOptions, KalmanFitterResult &kfResult,
           const kf_plane *surfaceSequence = nullptr, size_t surfaceSequenceSize = 0) {
    if (!kfOptions.referenceSurface)
      throw std::runtime_error("kfmi::KalmanFitter::fit: kfOptions.referenceSurface is null");
    int32_t first = 0, count = (int32_t)planes_.size();
    if (surfaceSequence) {
      const ptrdiff_t off = surfaceSequence - planes_.data();
      if (off < 0 || (size_t)off + surfaceSequenceSize > planes_.size())
        throw std::runtime_error("kfmi::KalmanFitter::fit: sequence is not in the plane table");
      first = (int32_t)off;
      count = (int32_t)surfaceSequenceSize;
    }
    const int64_t m = (int64_t)sourcelinks.size();
    const int32_t states = count > 0 ? count : 1;
    // host staging: one track, SoA
    std::vector<double> sp(sParameters.parameters.begin(), sParameters.parameters.end());
    std::vector<double> sc(21);
    packSym(sParameters.covariance, sc.data());
    std::vector<double> pg(kfOptions.referenceSurface->center.begin(),
                           kfOptions.referenceSurface->center.end());
    std::vector<int64_t> off = {0, m};
    std::vector<int32_t> geo(m);
    std::vector<double> meas(2 * m), hcov(3 * m);
    for (int64_t h = 0; h < m; ++h) {
      const PixelSourceLink &sl = sourcelinks[h];
      geo[h] = sl.geometryId;
      meas[h] = sl.values[0];
      meas[m + h] = sl.values[1];
      hcov[h] = sl.cov[0][0];
      hcov[m + h] = sl.cov[0][1];
      hcov[2 * m + h] = sl.cov[1][1];
    }
    check(hipSetDevice(device_) == hipSuccess ? KF_OK : KF_E_HIP, "hipSetDevice");
    const size_t ts = (size_t)states * KF_TS_STRIDE;
    const size_t bytes[] = {6 * 8, 21 * 8, 3 * 8, 4, 4, 2 * 8, (size_t)m * 4 + 4,
                            (size_t)m * 16 + 8, (size_t)m * 24 + 8, 6 * 8, 21 * 8, 8, 4, 4, 4,
                            ts * 8};
    void *d[16];
    for (int k = 0; k < 16; ++k) d[k] = stage(k, bytes[k]);
    copyIn(d[0], sp.data(), bytes[0]);
    copyIn(d[1], sc.data(), bytes[1]);
    copyIn(d[2], pg.data(), bytes[2]);
    copyIn(d[3], &first, 4);
    copyIn(d[4], &count, 4);
    copyIn(d[5], off.data(), bytes[5]);
    if (m > 0) {
      copyIn(d[6], geo.data(), (size_t)m * 4);
      copyIn(d[7], meas.data(), (size_t)m * 16);
      copyIn(d[8], hcov.data(), (size_t)m * 24);
    }
    kf_tracks t{};
    t.n_tracks = 1;
    t.n_hits = m;
    t.start_params = (const double *)d[0];
    t.start_cov = (const double *)d[1];
    t.perigee = (const double *)d[2];
    t.seq_first = (const int32_t *)d[3];
    t.seq_count = (const int32_t *)d[4];
    t.hit_offset = (const int64_t *)d[5];
    t.hit_geo_id = (const int32_t *)d[6];
    t.hit_meas = (const double *)d[7];
    t.hit_cov = (const double *)d[8];
    kf_results r{};
    r.params = (double *)d[9];
    r.cov = (double *)d[10];
    r.chi2 = (double *)d[11];
    r.ndf = (int32_t *)d[12];
    r.n_measured = (int32_t *)d[13];
    r.status = (uint32_t *)d[14];
    r.track_states = (double *)d[15];
    kf_options o = options(kfOptions);
    o.max_states = states;
    check(kf_fit(ctx_, &t, &o, &r, nullptr), "kf_fit");
    check(hipDeviceSynchronize() == hipSuccess ? KF_OK : KF_E_HIP, "hipDeviceSynchronize");
    double par[6], cov[21], chi2;
    int32_t ndf, nm;
    uint32_t status;
    std::vector<double> tsv(ts);
    copyOut(par, d[9], sizeof par);
    copyOut(cov, d[10], sizeof cov);
    copyOut(&chi2, d[11], 8);
    copyOut(&ndf, d[12], 4);
    copyOut(&nm, d[13], 4);
    copyOut(&status, d[14], 4);
    copyOut(tsv.data(), d[15], ts * 8);
    kfResult = KalmanFitterResult{};
    for (int k = 0; k < 6; ++k) kfResult.fittedParameters.parameters[k] = par[k];
    kfResult.fittedParameters.covariance = unpackSym(cov);
    kfResult.measurementStates = (size_t)nm;
    kfResult.smoothed = (status & KF_STATUS_SMOOTHED) != 0;
    kfResult.finished = (status & KF_STATUS_FINISHED) != 0;
    kfResult.result = (status & KF_STATUS_RESULT_FAIL) == 0 && (status & KF_STATUS_BAD_INPUT) == 0;
    kfResult.status = status;
    kfResult.chi2 = chi2;
    kfResult.ndf = ndf;
    for (int32_t k = 0; k < nm && k < states; ++k) {
      const double *rec = tsv.data() + (size_t)k * KF_TS_STRIDE;
      TrackState s;
      const int32_t g = (int32_t)rec[KF_TS_GEO_ID];
      for (const PixelSourceLink &sl : sourcelinks)
        if (sl.geometryId == g) {
          s.sourceLink = sl;
          break;
        }
      BoundParameters *bp[3] = {&s.predicted, &s.filtered, &s.smoothed};
      const int offs[3] = {KF_TS_PREDICTED, KF_TS_FILTERED, KF_TS_SMOOTHED};
      for (int q = 0; q < 3; ++q) {
        for (int c = 0; c < 6; ++c) bp[q]->parameters[c] = rec[offs[q] + c];
        bp[q]->covariance = unpackSym(rec + offs[q] + 6);
      }
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) s.jacobian[a][b] = rec[KF_TS_JACOBIAN + 6 * a + b];
      s.pathLength = rec[KF_TS_PATH];
      s.chi2 = rec[KF_TS_CHI2];
      kfResult.fittedStates.push_back(s);
    }
    // KalmanFitter.hpp:760-764
    return kfResult.result && kfResult.measurementStates == sourcelinks.size();
  }

  // Batch fit on caller-owned device buffers, asynchronous on `stream`.
  void fitBatch(const kf_tracks &tracks, const KalmanFitterOptions &kfOptions,
                const kf_results &results, hipStream_t stream = nullptr) {
    kf_options o = options(kfOptions);
    check(kf_fit(ctx_, &tracks, &o, &results, (void *)stream), "kf_fit");
  }

  // (forward filter, smoother + back-propagation) durations of the last fit [ms].
  std::pair<double, double> lastPhaseMs() const {
    double f = -1.0, b = -1.0;
    kf_last_phase_ms(ctx_, &f, &b);
    return {f, b};
  }

 private:
  static kf_options options(const KalmanFitterOptions &k) {
    kf_options o = k.propagator;
    o.smoothing = k.smoothing ? 1 : 0;
    o.multiple_scattering = k.multipleScattering ? 1 : 0;
    o.energy_loss = k.energyLoss ? 1 : 0;
    return o;
  }
  void check(int rc, const char *what) const {
    if (rc != KF_OK)
      throw std::runtime_error(std::string("kfmi: ") + what + " failed: " + kf_last_error(ctx_));
  }
  void *stage(int k, size_t bytes) {
    if (stage_.size() <= (size_t)k) {
      stage_.resize(k + 1, nullptr);
      stageBytes_.resize(k + 1, 0);
    }
    if (stageBytes_[k] < bytes) {
      if (stage_[k]) (void)hipFree(stage_[k]);
      stage_[k] = nullptr;
      if (hipMalloc(&stage_[k], bytes) != hipSuccess)
        throw std::runtime_error("kfmi: hipMalloc of the per-track staging buffers failed");
      stageBytes_[k] = bytes;
    }
    return stage_[k];
  }
  static void copyIn(void *dst, const void *src, size_t n) {
    if (hipMemcpy(dst, src, n, hipMemcpyHostToDevice) != hipSuccess)
      throw std::runtime_error("kfmi: hipMemcpy to the device failed");
  }
  static void copyOut(void *dst, const void *src, size_t n) {
    if (hipMemcpy(dst, src, n, hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("kfmi: hipMemcpy from the device failed");
  }

  std::vector<kf_plane> planes_;
  int device_ = 0;
  kf_ctx *ctx_ = nullptr;
  std::vector<void *> stage_;
  std::vector<size_t> stageBytes_;
};

}  // namespace kfmi
