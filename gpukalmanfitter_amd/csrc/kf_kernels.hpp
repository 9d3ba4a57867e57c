// kf_kernels.hpp -- launch interface between the C ABI (kf_api.hip) and the
// fit kernel (kf_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kf_device.hpp"

#define KF_BLOCK 128

namespace kfd {

// Kernel argument block (passed by value).  All pointers are device memory.
struct FitArgs {
  int64_t n;
  int64_t n_hits;
  const double *start_params;
  const double *start_cov;
  const double *perigee;
  const int32_t *seq_first;
  const int32_t *seq_count;
  const int64_t *hit_offset;
  const int32_t *hit_geo_id;
  const double *hit_meas;
  const double *hit_cov;
  const DevPlane *planes;
  int32_t n_planes;
  DevField field;
  // options
  int32_t smoothing, ms, eloss, max_steps;
  uint32_t max_rk_trials;
  int32_t max_states;
  double mass, tolerance, step_size_cutoff, max_step_size, target_tolerance, overstep_limit;
  // physics constants (host-evaluated exactly as the reference's literals)
  double kK, kMe, kPlasma, kHighland, kHighlandLog, kLandauConv;
  // outputs
  double *out_params;
  double *out_cov;
  double *out_chi2;
  int32_t *out_ndf;
  int32_t *out_nmeas;
  uint32_t *out_status;
  // per-state store: element f of state k of track i at
  // ts[(k * ts_fields + f) * n + i]
  double *ts;
  int32_t ts_fields;
  int32_t off_pred, off_filt, off_jac;
  int32_t full_record;  // 1: ts is kf_results.track_states (KF_TS_* layout)
};

}  // namespace kfd

extern "C" hipError_t kfd_launch_fit(const kfd::FitArgs *args, int precision, hipStream_t stream);
