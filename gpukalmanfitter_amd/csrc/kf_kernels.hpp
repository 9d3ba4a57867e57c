// kf_kernels.hpp -- launch interface between the C ABI (kf_api.hip) and the
// fit kernel (kf_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kf_device.hpp"

#define KF_BLOCK 128  // multiple of 64: a wave covers one 64-track store chunk
// Plane tables up to this size are copied to LDS by every fit block (plane
// reads then wait on LDS only, not behind the store writes' vector-memory
// count); larger tables are read from global memory.
#define KF_LDS_PLANES 256

// per-state store elements (see FitArgs).  A measured state sits in the
// store slot of its plane's position in the track's surface sequence (when
// the sequence fits max_states slots; else in slot = its measurement index),
// so the lanes of a wave that cross the same plane write and read the same
// 64-lane chunk of the store even when their tracks have holes.
#define KF_SD_XP 0   // predicted parameters (6)
#define KF_SD_XF 6   // filtered parameters (6)
#define KF_SD_N 12
#define KF_SS_J 0    // Jacobian rows 0, 1, 2, 3, 5 x columns 0..4 (25, padded to 28)
#define KF_SS_LINK 25  // in the padding: slot of the measured state two before (bits)
#define KF_SS_CF 28  // filtered covariance (21, packed upper triangle)
#define KF_SS_MD 49  // covariance (2,2), (3,3), (4,4) after the material update
#define KF_SS_N 52
#define KF_SS_USED 50  // elements that carry data (the rest is padding)
// forward -> backward hand-over record per track: int32 and double fields
#define KF_CARRY_I 8
#define KF_CARRY_D 30
// per-surface checkpoints of the chi2 adjoint (state after the update and
// material at a measured surface; see FitArgs::tsc)
#define KF_CK_POS 0    // 3
#define KF_CK_DIR 3    // 3
#define KF_CK_P 6
#define KF_CK_T 7
#define KF_CK_PATH 8
#define KF_CK_SS 9     // ConstrainedStep values (4)
#define KF_CK_COV 13   // 21
#define KF_CK_J 34     // A (15), b, G (9)
#define KF_CK_NEXT 59  // navigation index (as a double)
#define KF_CK_N 60

namespace kfd {

// Kernel argument block (passed by value).  All pointers are device memory.
struct FitArgs {
  int64_t n;  // tracks in the batch (the stride of the [component][n] arrays)
  // this launch covers track positions first .. first + count - 1; the
  // per-state stores are indexed by the position within the launch
  int64_t first, count;
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
  // the geo ids of the plane table are distinct (lets the hit search of a
  // track resume after its last match)
  int32_t unique_geo;
  DevField field;
  // options
  int32_t smoothing, ms, eloss, max_steps;
  uint32_t max_rk_trials;
  int32_t max_states;
  double mass, tolerance, step_size_cutoff, max_step_size, target_tolerance, overstep_limit;
  double outlier_chi2_cut;  // 0: no outlier finder
  // physics constants (host-evaluated exactly as the reference's literals)
  double kK, kMe, kPlasma, kHighland, kHighlandLog, kLandauConv;
  // outputs
  double *out_params;
  double *out_cov;
  double *out_chi2;
  int32_t *out_ndf;
  int32_t *out_nmeas;
  uint32_t *out_status;
  // Per-state store of the smoother, chunked by 64 track positions and kept
  // in 16-byte groups per lane (W = 16 / sizeof(element) elements): element
  // f of the state in slot s (KF_SD_* above) of position i sits at
  //   base[(((i / 64) * max_states * F + s * F + f - f % W) * 64 + (i % 64) * W) + f % W],
  // so one 16-byte vector access per lane moves W elements and a wave's
  // access is 1 KB contiguous.
  //  tsd  double, F = KF_SD_N: predicted and filtered parameters;
  //  tss  the transport scalar S (double in FP64 mode, float in mixed mode),
  //       F = KF_SS_N: the 25 free entries of the bound-to-bound transport
  //       Jacobian (row 4 is e4, column 5 is e5 on every surface), the
  //       filtered covariance and the three variances after the material
  //       update;
  //  tsf  double, F = KF_TS_STRIDE: the public track-state record, written
  //       only when the caller asks for track states (else NULL), one
  //       element per lane, by measurement index k:
  //       base[((i / 64) * max_states * F + k * F + f) * 64 + i % 64].
  // The predicted covariance is not stored: the smoother recomputes it from
  // the stored Jacobian and post-material covariance with the same operation
  // sequence, so it is bit-identical to the forward pass's.
  double *tsd;
  void *tss;
  double *tsf;
  //  tsc  double, F = KF_CK_N: per-surface checkpoints of the chi2 adjoint
  //       (NULL outside the adjoint).
  double *tsc;
  // forward -> backward kernel hand-over, [field][n]
  int32_t *carry_i;  // KF_CARRY_I int fields
  double *carry_d;   // KF_CARRY_D double fields
  // Scheduling.  Tracks are processed at positions 0..n-1; the track at
  // position p is order[p] (identity when NULL) and its store slot is that
  // of position p.  persistent = 1: a grid of resident waves pulls
  // chunks_per_grab 64-position chunks at a time from queue[0] (forward) and
  // queue[1] (backward), both zeroed before the launches.
  const int32_t *order;
  int32_t persistent;
  int32_t chunks_per_grab;
  uint32_t *queue;
};

}  // namespace kfd

// Step-stack depth per track of the chi2 adjoint: a segment between two
// measured surfaces may take this many RK steps.
#define KF_ADJ_MAX_SEGMENT_STEPS 48
// Track slices of the adjoint's first reduction stage.
#define KF_ADJ_RED_SLICES 64

namespace kfd {
struct AdjArgs;
}
// Launches the reverse sweep of the chi2 adjoint over the positions
// args->first .. first + count - 1 and its reduction into sums (added onto
// the previous launches' sums when accumulate != 0); the forward filter with
// checkpoints must have run over the same positions.
extern "C" hipError_t kfd_launch_adjoint(const kfd::FitArgs *args, const kfd::AdjArgs *adj,
                                         double *sums, int accumulate, hipStream_t stream);
// Converts the summed raw adjoints to per-plane alignment gradients.
extern "C" hipError_t kfd_launch_adjoint_finish(const kfd::FitArgs *args, const double *sums,
                                                double *grad_planes, double *grad_scale,
                                                double *chi2, hipStream_t stream);

// Batched ray x polygon-plane intersection (kf_intersect.hip).
extern "C" int64_t kfd_intersect_scratch_bytes(int64_t n_rays, int64_t n_surfaces);
extern "C" hipError_t kfd_launch_intersect(const kf_polygon_plane *surfaces, int64_t n_surfaces,
                                           const double *ray_pos, const double *ray_dir,
                                           int64_t n_rays, int bcheck, kf_intersection *hits,
                                           int64_t capacity, int64_t *n_hits, void *scratch,
                                           hipStream_t s);

// Launch shape of kfd_launch_fit.
struct KfdLaunch {
  int32_t sort_by_length;  // order positions by decreasing sequence length first
  int32_t persistent_blocks;  // > 0: persistent grid of this many blocks
  int32_t lanes_per_track;  // 1, 2 or 4 lanes per position (one-shot grid only)
  int32_t store_by_plane;   // store slots by plane position (one lane per position)
  uint32_t *sort_scratch;  // 256 + 256 counters (length histogram, bin offsets)
  int32_t *order;          // [n] positions -> tracks (when sorting)
};

// Launches the fit of the positions args->first .. first + count - 1:
// [work-queue reset] -> forward kernel -> record `between` (if not NULL) ->
// smoother/back-propagation kernel -> track-state export (if
// track_states_out is not NULL).  `launch` may be NULL (one thread per
// track, identity order).  With a persistent grid, args->queue and
// args->chunks_per_grab must be set; with sorting, args->order.
extern "C" hipError_t kfd_launch_fit(const kfd::FitArgs *args, int precision, hipStream_t stream,
                                     double *track_states_out, hipEvent_t between,
                                     const KfdLaunch *launch);

// The length sort of the whole batch (positions -> tracks into launch->order),
// before the per-slice kfd_launch_fit calls.
extern "C" hipError_t kfd_launch_sort(const kfd::FitArgs *args, const KfdLaunch *launch,
                                      hipStream_t stream);
