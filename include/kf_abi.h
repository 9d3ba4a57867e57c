/*
 * kf_abi.h -- C ABI of the MI355X Kalman track fitter (libkfmi.so).
 *
 * This is the drop-in boundary for the per-track Kalman fit of
 * gpuKalmanFitter.  One call of kf_fit() fits a batch of tracks; every entry
 * point takes plain pointers and sizes, no C++ or torch types.
 *
 * Reference interfaces replaced (paths are inside the gpuKalmanFitter tree):
 *   KalmanFitter::fit(sourcelinks, startParameters, kfOptions, kfResult,
 *                     surfaceSequence, surfaceSequenceSize) -> bool
 *       include/Fitter/KalmanFitter.hpp:716-768          -> kf_fit (n_tracks = 1
 *                                                            is the per-track call)
 *   fitKernelThreadPerTrack / fitKernelBlockPerTrack
 *       Run/GPU/Kernels.cu:6-94                           -> kf_fit (batch)
 *   KalmanFitterOptions{smoothing, multipleScattering, energyLoss, referenceSurface}
 *       include/Fitter/KalmanFitter.hpp:48-96             -> kf_options
 *   PropagatorOptions{mass, absPdgCode, maxSteps, maxRungeKuttaStepTrials,
 *                     maxStepSize, targetTolerance, tolerance, stepSizeCutOff}
 *       include/Propagator/Propagator.hpp:52-111          -> kf_options
 *   PlaneSurface(center, normal, HomogeneousSurfaceMaterial) + GeometryID
 *       include/Surfaces/detail/PlaneSurface.ipp:48-69,
 *       include/Geometry/GeometryID.hpp:30-123            -> kf_plane, kf_set_geometry
 *   Test::ConstantBField / InterpolatedBFieldMap::getField
 *       include/Test/Helper.hpp:17-22,
 *       include/MagneticField/InterpolatedBFieldMap.hpp:318-320
 *                                                         -> kf_set_field_constant,
 *                                                            kf_set_field_grid
 *   PixelSourceLink{values, cov, geometryId}
 *       include/EventData/PixelSourceLink.hpp:22-75        -> kf_tracks.hit_*
 *   BoundState{boundParams, boundCov} + LineSurface(perigee)
 *       Run/Common/FitData.hpp:45-48, Run/Common/Processor.hpp:121-129
 *                                                         -> kf_tracks.start_*, perigee
 *   KalmanFitterResult{fittedStates, fittedParameters, measurementStates,
 *                      smoothed, finished, result}
 *       include/Fitter/KalmanFitter.hpp:98-120             -> kf_results
 *
 * Units are the reference's native units: mm, GeV, elementary charge, time in
 * mm/c, magnetic field in GeV/(e*mm) (1 T = 0.000299792458).
 *
 * Layout: every per-track array is structure-of-arrays, component k of track i
 * at [k * n_tracks + i].  Per-hit arrays are CSR over tracks: the hits of
 * track i are hit_offset[i] .. hit_offset[i+1]-1, component k of hit h at
 * [k * n_hits + h].  Symmetric 6x6 matrices are packed as the upper triangle,
 * row-major: (0,0) (0,1) .. (0,5) (1,1) .. (1,5) (2,2) .. (5,5)  -> 21 values.
 * Bound parameters are (loc0, loc1, phi, theta, q/p, t).
 *
 * Errors: every function returns 0 on success and a negative KF_E_* code on
 * failure; kf_last_error() then holds a message.  Nothing aborts the process.
 * Per-track outcome is a status bit mask (KF_STATUS_*).
 */
#ifndef KF_ABI_H
#define KF_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KF_ABI_VERSION 4

/* return codes */
#define KF_OK 0
#define KF_E_INVALID (-1)   /* bad argument (null pointer, size, layout) */
#define KF_E_HIP (-2)       /* a HIP runtime call failed */
#define KF_E_STATE (-3)     /* context not ready (no geometry / field) */
#define KF_E_NOMEM (-4)     /* device allocation failed */

/* per-track status bits (kf_results.status) */
#define KF_STATUS_OK 0x001u          /* fit() returned true: result && all hits used */
#define KF_STATUS_SMOOTHED 0x002u    /* KalmanFitterResult::smoothed */
#define KF_STATUS_FINISHED 0x004u    /* KalmanFitterResult::finished (perigee params set) */
#define KF_STATUS_RESULT_FAIL 0x008u /* KalmanFitterResult::result == false */
#define KF_STATUS_STEP_FAIL 0x010u   /* an RK step hit the trial limit or step cut-off */
#define KF_STATUS_MAX_STEPS 0x020u   /* propagation stopped at maxSteps */
#define KF_STATUS_NAV_BREAK 0x040u   /* surface sequence exhausted before all hits used */
#define KF_STATUS_SMOOTHER_NAN 0x080u /* smoother gain had a NaN */
#define KF_STATUS_BAD_INPUT 0x100u   /* empty sequence, too many states, q/p == 0 */

/* kf_options.precision */
#define KF_PRECISION_FP64 0  /* everything in double */
#define KF_PRECISION_MIXED 1 /* fp32 stepping / transport, fp64 update + smoother gain */

/* per-state record of kf_results.track_states: component offsets */
#define KF_TS_PREDICTED 0   /* 6 params + 21 packed cov */
#define KF_TS_FILTERED 27   /* 6 params + 21 packed cov */
#define KF_TS_SMOOTHED 54   /* 6 params + 21 packed cov */
#define KF_TS_JACOBIAN 81   /* 36, row-major bound-to-bound transport from the previous state */
#define KF_TS_PATH 117      /* path length accumulated since the start surface */
#define KF_TS_CHI2 118      /* predicted-residual chi2 r^T S^-1 r */
#define KF_TS_GEO_ID 119    /* geometry id of the surface (as a double) */
#define KF_TS_STRIDE 120

/* Planar measurement surface with an optional homogeneous material slab.
 * rot is row-major; its columns are the local U axis, the local V axis and
 * the plane normal T (PlaneSurface.ipp:48-69 builds U = z x T (or x x T), V = T x U). */
typedef struct kf_plane {
  double center[3];
  double rot[9];
  double thickness; /* slab thickness [mm]; <= 0: no material */
  double x0;        /* radiation length [mm] */
  double l0;        /* nuclear interaction length [mm] */
  double ar;        /* relative atomic mass */
  double z;         /* nuclear charge */
  double molar_rho; /* molar density [mol/mm^3] */
  double half_u;    /* rectangle half lengths (informational: the direct */
  double half_v;    /* navigator intersects without a bounds check) */
  int32_t geo_id;   /* matched against kf_tracks.hit_geo_id */
  int32_t reserved;
} kf_plane;

typedef struct kf_options {
  int32_t smoothing;           /* run the RTS smoother + back-propagation (default 1) */
  int32_t multiple_scattering; /* default 1 */
  int32_t energy_loss;         /* default 1 */
  int32_t precision;           /* KF_PRECISION_* (default FP64) */
  int32_t abs_pdg;             /* default 13 */
  int32_t max_steps;           /* default 1000 */
  int32_t max_rk_trials;       /* default 10000 */
  int32_t max_states;          /* capacity of the per-track state store, at least the
                                  longest surface sequence of the batch; 0 = the size
                                  of the plane table (which holds every sequence);
                                  the store and the memory slices are sized by it */
  double mass;                 /* default 105.6583755 MeV = 0.1056583755 */
  double tolerance;            /* RK error tolerance, default 1e-4 */
  double step_size_cutoff;     /* default 0 */
  double max_step_size;        /* default DBL_MAX */
  double target_tolerance;     /* default 1e-4 (path-limit aborter) */
  double overstep_limit;       /* default 0.01 mm: a surface more than this far behind
                                  is unreachable (EigenStepper.hpp:370, hard-coded
                                  m_overstepLimit in the reference) */
  double outlier_chi2_cut;     /* outlier finder: 0 (default) = VoidOutlierFinder
                                  (VoidKalmanComponents.hpp:66-80, nothing is an outlier);
                                  > 0: a measurement whose predicted-residual chi2
                                  exceeds it is an outlier: no update (filtered =
                                  predicted), left out of chi2 and ndf, still a
                                  measurement state (smoothed, counted in n_measured).
                                  Outliers of a track = n_measured - ndf / 2; a state is
                                  one when its KF_TS_CHI2 exceeds the cut.  Not
                                  supported by kf_chi2_gradient. */
} kf_options;

/* One batch of tracks.  Every pointer is a device pointer for kf_fit(). */
typedef struct kf_tracks {
  int64_t n_tracks;
  int64_t n_hits;
  const double *start_params; /* [6][n_tracks] on the perigee line surface */
  const double *start_cov;    /* [21][n_tracks] packed */
  const double *perigee;      /* [3][n_tracks] points of the z-parallel perigee lines */
  const int32_t *seq_first;   /* [n_tracks] first plane-table index of the sequence */
  const int32_t *seq_count;   /* [n_tracks] number of planes in the sequence */
  const int64_t *hit_offset;  /* [n_tracks + 1] CSR offsets into the hit arrays */
  const int32_t *hit_geo_id;  /* [n_hits] */
  const double *hit_meas;     /* [2][n_hits] local (u, v) measurement */
  const double *hit_cov;      /* [3][n_hits] V00, V01, V11 */
} kf_tracks;

/* Outputs, device pointers.  track_states may be NULL. */
typedef struct kf_results {
  double *params;     /* [6][n_tracks] fitted parameters on the perigee line */
  double *cov;        /* [21][n_tracks] */
  double *chi2;       /* [n_tracks] sum of predicted-residual chi2 */
  int32_t *ndf;       /* [n_tracks] 2 * (measured - outliers) */
  int32_t *n_measured;/* [n_tracks] KalmanFitterResult::measurementStates */
  uint32_t *status;   /* [n_tracks] KF_STATUS_* */
  double *track_states; /* optional [max_states][KF_TS_STRIDE][n_tracks]; a fit
                           that asks for them runs the kernel build that writes them,
                           whose results agree with a fit without them to rounding
                           (FP64 ~1e-12), not bit for bit */
} kf_results;

/* How kf_fit lays the batch onto the GPU (this build's; the reference
 * launches one thread per track on a fixed grid, Run/GPU/Kernels.cu:6-55). */
typedef struct kf_schedule {
  int32_t sort_by_length;  /* fit tracks in order of decreasing sequence length, so a
                              wave's tracks take similar numbers of steps (default 0) */
  int32_t persistent;      /* 1: resident waves pull 64-track chunks from a work queue */
  int32_t waves_per_cu;    /* persistent grid size, waves per compute unit (default 4) */
  int32_t chunks_per_grab; /* work-queue granularity, 64-track chunks per grab (default 1) */
  int32_t lanes_per_track; /* lanes cooperating on one track (the counterpart of the
                              reference's block-per-track kernel, Kernels.cu:47-94):
                              1 (default), 2 or 4 consecutive lanes, one-shot grid only */
  int32_t graph;           /* 1: capture the launch sequence as a hipGraph on the first
                              call and replay it while the arguments are unchanged */
  int32_t store_by_plane;  /* 1: keep a measured state in the smoother store at its
                              plane's position in the sequence rather than at its
                              measurement index, so the tracks of a wave write and read
                              one store chunk per plane when they have holes (for
                              batches with holes; one lane per track; default 0) */
} kf_schedule;

typedef struct kf_ctx kf_ctx;

int kf_abi_version(void);

/* Fills *opt with the reference defaults (Propagator.hpp:66-95,
 * KalmanFitter.hpp:72). */
void kf_options_default(kf_options *opt);

/* Creates a fitting context on HIP device `device`. */
int kf_ctx_create(int device, kf_ctx **out);
int kf_ctx_destroy(kf_ctx *ctx);
/* Message of the last failing call on ctx (or of kf_ctx_create when ctx is NULL). */
const char *kf_last_error(const kf_ctx *ctx);

/* Copies the plane table (host memory) to the device. */
int kf_set_geometry(kf_ctx *ctx, const kf_plane *planes, int32_t n_planes);
/* Uniform field, GeV/(e*mm). */
int kf_set_field_constant(kf_ctx *ctx, double bx, double by, double bz);
/* Trilinear grid field (host memory): b[((ix*ny + iy)*nz + iz)*3 + c] in
 * GeV/(e*mm) at lo + i*(hi-lo)/(n-1); positions outside are clamped. */
int kf_set_field_grid(kf_ctx *ctx, const float *b, int32_t nx, int32_t ny,
                      int32_t nz, const double lo[3], const double hi[3]);
/* Multiplies the field (constant or grid) by s. */
int kf_set_field_scale(kf_ctx *ctx, double s);

/* Fits the batch asynchronously on `stream` (a hipStream_t; NULL = default
 * stream).  Inputs and outputs are device memory owned by the caller. */
int kf_fit(kf_ctx *ctx, const kf_tracks *tracks, const kf_options *opt,
           const kf_results *res, void *stream);

/* kf_fit without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  The working storage (smoother
 * store, hand-over records, track-state staging) comes from a fixed
 * device-resident workspace of KF_NOALLOC_WORKSPACE_BYTES that the library
 * holds on every device it runs on until the process ends; batches larger
 * than it holds run in consecutive slices of 64-track multiples, with the
 * same results as kf_fit bit for bit.  Always the one-thread-per-track
 * schedule (no length sort, work queue or graph replay).  Returns
 * KF_E_INVALID when the working storage of one 64-track slice,
 * 64 x (max_states x (512 B in FP64, 304 B mixed, + 960 B with track
 * states) + 272 B), does not fit the workspace (max_states: opt->max_states,
 * else the plane-table size; in FP64 without track states up to 8,191).  Calls on
 * one device share its workspace: calls that overlap in time on one GPU
 * (different streams or host threads) race and must be serialised by the
 * caller.  The context must have been created (kf_ctx_create resolves the
 * workspace of its device). */
#define KF_NOALLOC_WORKSPACE_BYTES (256ull << 20)
int kf_fit_noalloc(kf_ctx *ctx, const kf_tracks *tracks, const kf_options *opt,
                   const kf_results *res, void *stream);

/* Gradient of the summed chi2 of the forward filter (the fit's results.chi2
 * summed over the batch) with respect to the alignment of every plane of the
 * table and to the field scale, by a hand-written adjoint (reverse sweep over
 * per-surface checkpoints), asynchronously on `stream`, always in FP64.
 *   grad_planes       device [n_planes][6]: d/d(du, dv, dw, a, b, c) at zero,
 *                     with the plane moved as centre' = centre + R (du, dv, dw)
 *                     and R' = R Rx(a) Ry(b) Rz(c) (local shifts [mm] and
 *                     rotations about the local U, V, normal axes [rad]);
 *   grad_field_scale  device [1]: d/ds of the field B -> s B at the current s;
 *   chi2_sum          device [1] or NULL: the objective;
 *   grad_raw          device [n_planes][12] or NULL: dX/d(centre) (3) and
 *                     dX/d(rot) (9, row-major) of every plane, for callers
 *                     with their own alignment parameterisation;
 *   track_flags       device [n_tracks] or NULL: 0, or KF_ADJ_* bits of a
 *                     track left out of the gradient.
 * Material effects (energy loss, scattering and straggling variances) are
 * differentiated with the rest.  No reference
 * counterpart (this build's addition; the objective is defined in
 * results.chi2). */
#define KF_ADJ_SKIPPED 0x1u   /* no measured surface */
#define KF_ADJ_OVERFLOW 0x2u  /* a segment took more RK steps than the adjoint stacks */
#define KF_ADJ_LOST 0x4u      /* the recomputed segment did not reach its surface */
int kf_chi2_gradient(kf_ctx *ctx, const kf_tracks *tracks, const kf_options *opt,
                     double *grad_planes, double *grad_field_scale, double *chi2_sum,
                     double *grad_raw, uint32_t *track_flags, void *stream);

/* Schedule of later kf_fit calls on ctx (kf_schedule_default: one thread per
 * track in input order, no graph). */
void kf_schedule_default(kf_schedule *s);
int kf_set_schedule(kf_ctx *ctx, const kf_schedule *s);
int kf_get_schedule(const kf_ctx *ctx, kf_schedule *s);

/* Autotuner: fits the batch with each candidate schedule (sort on/off x
 * store slots by measurement or by plane x one-shot grid or persistent work
 * queue with 2/4/8 waves per CU and 1/4 chunks per grab, graph replay), times
 * every candidate on `stream` with HIP events (best of `repeats` fits), keeps
 * the fastest as ctx's schedule and returns it with its time.  The results
 * buffers hold the last fit.  The one-shot candidates with slots by
 * measurement also run with 2 and 4 lanes per track. */
int kf_autotune(kf_ctx *ctx, const kf_tracks *tracks, const kf_options *opt,
                const kf_results *res, void *stream, int32_t repeats, kf_schedule *best,
                double *best_ms);

/* Device-memory budget of ctx's per-batch working buffers (per-state stores,
 * kernel hand-over, adjoint stacks and per-track adjoints); 0 restores the
 * default, KF_DEFAULT_MEMORY_BUDGET.  A batch whose working set exceeds the
 * budget is fitted in consecutive slices of positions (multiples of 64
 * tracks) that reuse the same buffers; results do not depend on the slicing.
 * Buffers only grow: lowering the budget does not release memory (a new
 * context does).  The reference divides a batch into nStreams equal parts
 * for copy/compute overlap (Run/GPU/KalmanFitterGPUTest.cu:212-219); here the
 * parts are sized by memory. */
#define KF_DEFAULT_MEMORY_BUDGET (16ull << 30)
int kf_set_memory_budget(kf_ctx *ctx, uint64_t bytes);

/* Duration of the last fit kernel on ctx in milliseconds (HIP events recorded
 * on the launch stream); < 0 if unknown. */
double kf_last_kernel_ms(const kf_ctx *ctx);

/* The same split at the hand-over between the two kernels of a fit: the
 * forward filter (up to the reference's finalize()) and the smoother plus
 * back-propagation to the perigee.  The backward time includes the
 * track-state export when track states were requested.  No reference
 * counterpart (the reference times the whole batch from the host,
 * Run/GPU/KalmanFitterGPUTest.cu:382-513). */
int kf_last_phase_ms(const kf_ctx *ctx, double *forward_ms, double *backward_ms);

/* Number of launch sequences ctx has captured as a hipGraph (kf_schedule.graph);
 * a kf_fit whose arguments equal the previous call's replays the captured
 * graph and captures nothing.  No reference counterpart (the reference has no
 * graph replay); for callers that check their calls replay. */
int64_t kf_graph_captures(const kf_ctx *ctx);

/* ---- Batched ray x bounded-plane intersection ------------------------------
 * The reference's intersection test (Run/GPU/IntersectGPUTest.cu:22-48,
 * 116-172: one ray against every surface of a triangular mesh, then a
 * compaction of the valid intersections) as one batched call over many rays.
 *
 * A plane surface with convex polygon bounds (ConvexPolygonBounds<N>,
 * include/Surfaces/ConvexPolygonBounds.hpp): 3..KF_POLY_MAX vertices in the
 * local (u, v) frame, in order around the polygon (the kernel reads a
 * n_vertices outside 3..KF_POLY_MAX as the nearest bound: the surfaces are
 * device memory, the host cannot check them). */
#define KF_POLY_MAX 8
typedef struct kf_polygon_plane {
  double center[3];
  double rot[9];                    /* row-major, columns U, V, normal (as kf_plane) */
  double vertices[2 * KF_POLY_MAX]; /* (u, v) of vertex i at [2 i], [2 i + 1] */
  int32_t n_vertices;
  int32_t pad;
} kf_polygon_plane;

typedef struct kf_intersection {
  int32_t ray, surface;
  double path;        /* path length along the ray direction */
  double position[3]; /* global position of the intersection */
} kf_intersection;

/* Intersects every ray with every surface as PlaneSurface::intersect
 * (PlaneSurface.ipp:239-259, PlanarHelper::intersectionEstimate,
 * PlanarHelper.hpp:29-48) and keeps what the reference driver keeps: status
 * reachable (not on the surface within s_onSurfaceTolerance), path >= 0 and,
 * with bcheck != 0, the local position inside the polygon
 * (BoundaryCheck::isInsidePolygon with zero tolerance, BoundaryCheck.hpp).
 * Arrays are device memory: surfaces [n_surfaces], ray_pos / ray_dir
 * [3][n_rays], hits [capacity], n_hits [1], scratch of
 * kf_intersect_scratch_bytes(n_rays, n_surfaces) bytes.  The kept intersections are
 * written in (ray, surface) order; *n_hits receives their number, of which
 * the first min(*n_hits, capacity) are written.  Asynchronous on `stream`.
 * No kf_ctx: errors are reported by kf_last_error(NULL). */
int64_t kf_intersect_scratch_bytes(int64_t n_rays, int64_t n_surfaces);
int kf_intersect(const kf_polygon_plane *surfaces, int64_t n_surfaces, const double *ray_pos,
                 const double *ray_dir, int64_t n_rays, int32_t bcheck, kf_intersection *hits,
                 int64_t capacity, int64_t *n_hits, void *scratch, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* KF_ABI_H */
