// kf_kernels.hpp -- launch interface between the C ABI (kf_api.hip) and the
// fit kernel (kf_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kf_device.hpp"

#define KF_BLOCK 128  // multiple of 64: a wave covers one 64-track store chunk

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
  // per-state store, chunked by 64 tracks: element f of state k of track i at
  // ts[((i / 64) * max_states * ts_fields + k * ts_fields + f) * 64 + i % 64]
  double *ts;
  int32_t ts_fields;
  int32_t off_pred, off_filt, off_jac;
  int32_t full_record;  // 1: fields use the KF_TS_* layout (120), exported after the fit
  // forward -> backward kernel hand-over, [field][n]
  int32_t *carry_i;  // 6 int fields
  double *carry_d;   // 30 double fields
};

}  // namespace kfd

extern "C" hipError_t kfd_launch_fit(const kfd::FitArgs *args, int precision, hipStream_t stream,
                                     double *track_states_out);
