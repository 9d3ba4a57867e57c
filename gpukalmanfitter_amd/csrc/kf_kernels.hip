// kf_kernels.hip -- the batched Kalman-fit kernels for gfx950 and their
// launcher: the forward filter, the smoother plus back-propagation, and the
// track-state export.  The per-track program is TrackFit (kf_trackfit.hpp).
#include <hip/hip_runtime.h>

#include "kf_trackfit.hpp"

namespace kfd {

// Bases of the state stores of the 64-track chunk this wave works on.
// Blocks are multiples of 64 work-items, so a wave covers exactly one chunk.
template <typename S>
struct ChunkBases {
  double *d;
  S *s;
  double *f;
  double *c;
  KFD_INLINE ChunkBases(const FitArgs &a, int64_t i) {
    const size_t chunk = __builtin_amdgcn_readfirstlane((uint32_t)(i >> 6));
    const size_t per = (size_t)a.max_states * 64u;
    d = a.tsd + chunk * per * KF_SD_N;
    s = (S *)a.tss + chunk * per * KF_SS_N;
    f = a.tsf ? a.tsf + chunk * per * KF_TS_STRIDE : nullptr;
    c = a.tsc ? a.tsc + chunk * per * KF_CK_N : nullptr;
  }
};

// Fits the track at position pos (pass FWD of the two kernels); `rank` is
// the lane's index in the track's group of L lanes.
template <typename S, bool FWD, int L, bool PS, int FK, bool REC>
KFD_INLINE void fitPosition(const FitArgs &a, int64_t local, uint32_t rank,
                            const DevPlane *planes) {
  const ChunkBases<S> cb(a, local);
  if (local >= a.count) return;
  const int64_t pos = a.first + local;
  const int64_t i = a.order ? (int64_t)a.order[pos] : pos;
  TrackFit<S, FWD, L, PS, false, FK, REC> tf(a, i, cb.d, cb.s, cb.f, cb.c,
                                             (uint32_t)(local & 63), local);
  tf.setRank(rank);
  tf.pls = planes;
  if constexpr (FWD)
    tf.runForward();
  else
    tf.runBackward();
}

// One thread per position, or (persistent) resident waves that take
// chunks_per_grab 64-position chunks at a time from a work-queue counter
// until every chunk is taken; the grab is one vector atomic by lane 0.
//
// LDS = true: the block first copies the plane table (at most KF_LDS_PLANES
// planes) to LDS and reads planes there.
// PERS = false: one thread per position; true: the persistent work queue
// (a separate instantiation, so the one-shot kernel does not carry the
// queue loop's registers).
// L: lanes per track (one-shot grid only): L consecutive lanes fit one
// position, a wave covers 64 / L positions of one 64-position store chunk.
// PS: store slots by plane position (kf_schedule.store_by_plane; one lane
// per track).
// FK, REC: the field kind and whether track states or checkpoints are
// written.  The launcher builds on (kFieldConst, false) or (kFieldGrid,
// false) for the batch's field when neither is asked for, so the step loop
// carries no code of the other field kind and the filter none of the
// recording; otherwise, and for plane tables read from global memory, on
// (kFieldAny, true).
template <typename S, bool FWD, bool LDS, bool PERS, int L, bool PS, int FK, bool REC>
__global__ void __launch_bounds__(KF_BLOCK) kf_fit_kernel(FitArgs args) {
  static_assert(L == 1 || !PERS, "lane groups run on the one-shot grid");
  static_assert(L == 1 || !PS, "store slots by plane with one lane per track");
  const DevPlane *planes = args.planes;
  if constexpr (LDS) {
    __shared__ DevPlane lds_planes[KF_LDS_PLANES];
    const double *src = (const double *)args.planes;
    double *dst = (double *)lds_planes;
    const int words = args.n_planes * (int)(sizeof(DevPlane) / sizeof(double));
    for (int w = threadIdx.x; w < words; w += KF_BLOCK) dst[w] = src[w];
    __syncthreads();
    planes = lds_planes;
  }
  const uint32_t lane = threadIdx.x & 63u;
  if constexpr (!PERS) {
    const int64_t t = (int64_t)blockIdx.x * KF_BLOCK + threadIdx.x;
    fitPosition<S, FWD, L, PS, FK, REC>(args, t / L, (uint32_t)(t % L), planes);
  } else {
    const uint32_t nchunks = (uint32_t)((args.count + 63) >> 6);
    const uint32_t grab = (uint32_t)args.chunks_per_grab;
    uint32_t *q = args.queue + (FWD ? 0 : 1);
    for (;;) {
      uint32_t c0 = 0;
      if (lane == 0) c0 = atomicAdd(q, grab);
      c0 = (uint32_t)__shfl((int)c0, 0, 64);
      if (c0 >= nchunks) break;
      const uint32_t c1 = (c0 + grab < nchunks) ? c0 + grab : nchunks;
      for (uint32_t c = c0; c < c1; ++c)
        fitPosition<S, FWD, 1, PS, FK, REC>(args, (int64_t)c * 64 + lane, 0u, planes);
    }
  }
}

// Copies the public track-state records into kf_results.track_states
// ([state][KF_TS_STRIDE][n]).
__global__ void __launch_bounds__(256) kf_export_states(FitArgs args, double *out) {
  const int64_t local = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (local >= args.count) return;
  const int64_t pos = args.first + local;
  const int64_t i = args.order ? (int64_t)args.order[pos] : pos;
  const int nm = args.out_nmeas[i];
  const double *base =
      args.tsf + (size_t)(local >> 6) * (size_t)args.max_states * KF_TS_STRIDE * 64u;
  const uint32_t lane = (uint32_t)(local & 63);
  for (int k = 0; k < nm && k < args.max_states; ++k)
    for (int f = 0; f < KF_TS_STRIDE; ++f)
      out[((size_t)k * KF_TS_STRIDE + f) * (size_t)args.n + i] =
          base[((uint32_t)k * KF_TS_STRIDE + (uint32_t)f) * 64u + lane];
}

// ------------------------------------------------------- length sort ----
// Counting sort of the tracks by decreasing sequence length (lengths >= 255
// share the top bin): tracks of one wave then take similar numbers of steps,
// and a persistent grid starts on the longest.  Per-block LDS histograms keep
// the global atomics to one per block and bin.
constexpr int kLenBins = 256;

KFD_INLINE int lengthBin(const FitArgs &a, int64_t i) {
  int l = a.seq_count[i];
  return l < 0 ? 0 : (l >= kLenBins ? kLenBins - 1 : l);
}

// zeroes the histogram and bin offsets
__global__ void __launch_bounds__(256) kf_sort_reset(uint32_t *scratch) {
  scratch[threadIdx.x] = 0u;
  scratch[kLenBins + threadIdx.x] = 0u;
}

// zeroes the work-queue counters (forward, backward)
__global__ void __launch_bounds__(64) kf_queue_reset(uint32_t *queue) {
  if (threadIdx.x < 2) queue[threadIdx.x] = 0u;
}

__global__ void __launch_bounds__(256) kf_len_hist(FitArgs a, uint32_t *hist) {
  __shared__ uint32_t h[kLenBins];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n) atomicAdd(&h[lengthBin(a, i)], 1u);
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// off[len] = number of tracks longer than len (one block of 256)
__global__ void __launch_bounds__(256) kf_len_scan(uint32_t *hist, uint32_t *off) {
  __shared__ uint32_t s[kLenBins];
  const int t = threadIdx.x;
  s[t] = hist[kLenBins - 1 - t];
  __syncthreads();
  for (int w = 1; w < kLenBins; w <<= 1) {
    const uint32_t v = (t >= w) ? s[t - w] : 0u;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  off[kLenBins - 1 - t] = s[t] - hist[kLenBins - 1 - t];
}

__global__ void __launch_bounds__(256) kf_len_scatter(FitArgs a, uint32_t *off, int32_t *order) {
  __shared__ uint32_t cnt[kLenBins], base[kLenBins];
  cnt[threadIdx.x] = 0u;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int bin = (i < a.n) ? lengthBin(a, i) : 0;
  uint32_t local = 0;
  if (i < a.n) local = atomicAdd(&cnt[bin], 1u);
  __syncthreads();
  if (cnt[threadIdx.x]) base[threadIdx.x] = atomicAdd(&off[threadIdx.x], cnt[threadIdx.x]);
  __syncthreads();
  if (i < a.n) order[base[bin] + local] = (int32_t)i;
}

}  // namespace kfd

extern "C" hipError_t kfd_launch_fit(const kfd::FitArgs *args, int precision, hipStream_t stream,
                                     double *track_states_out, hipEvent_t between,
                                     const KfdLaunch *launch) {
  if (args->count <= 0) return hipSuccess;
  const bool pers = launch && launch->persistent_blocks > 0;
  const int lpt = (launch && !pers && launch->lanes_per_track > 1) ? launch->lanes_per_track : 1;
  if (lpt != 1 && lpt != 2 && lpt != 4) return hipErrorInvalidValue;
  const bool ps = launch && launch->store_by_plane != 0;
  if (ps && lpt != 1) return hipErrorInvalidValue;
  if (pers) hipLaunchKernelGGL(kfd::kf_queue_reset, dim3(1), dim3(64), 0, stream, args->queue);
  const int64_t blocks =
      pers ? launch->persistent_blocks : (args->count * lpt + KF_BLOCK - 1) / KF_BLOCK;
  const dim3 grid((unsigned)blocks), block(KF_BLOCK);
  const bool lds = args->n_planes <= KF_LDS_PLANES;
  const bool fp64 = precision == KF_PRECISION_FP64;
  const bool gridField = args->field.grid != 0;
  const bool rec = args->tsf != nullptr || args->tsc != nullptr;
#define KFD_LAUNCH2(S_, FWD_, LDS_, PERS_, L_, PS_, FK_, REC_)                             \
  hipLaunchKernelGGL((kfd::kf_fit_kernel<S_, FWD_, LDS_, PERS_, L_, PS_, FK_, REC_>), grid, \
                     block, 0, stream, *args)
#define KFD_LAUNCH1(S_, FWD_, LDS_, FK_, REC_)                       \
  do {                                                              \
    if (pers && ps)                                                 \
      KFD_LAUNCH2(S_, FWD_, LDS_, true, 1, true, FK_, REC_);        \
    else if (pers)                                                  \
      KFD_LAUNCH2(S_, FWD_, LDS_, true, 1, false, FK_, REC_);       \
    else if (lpt == 1 && ps)                                        \
      KFD_LAUNCH2(S_, FWD_, LDS_, false, 1, true, FK_, REC_);       \
    else if (lpt == 1)                                              \
      KFD_LAUNCH2(S_, FWD_, LDS_, false, 1, false, FK_, REC_);      \
    else if (lpt == 2)                                              \
      KFD_LAUNCH2(S_, FWD_, LDS_, false, 2, false, FK_, REC_);      \
    else                                                            \
      KFD_LAUNCH2(S_, FWD_, LDS_, false, 4, false, FK_, REC_);      \
  } while (0)
#define KFD_LAUNCH(S_, FWD_)                                 \
  do {                                                       \
    if (lds && !rec && gridField)                            \
      KFD_LAUNCH1(S_, FWD_, true, kfd::kFieldGrid, false);   \
    else if (lds && !rec)                                    \
      KFD_LAUNCH1(S_, FWD_, true, kfd::kFieldConst, false);  \
    else if (lds)                                            \
      KFD_LAUNCH1(S_, FWD_, true, kfd::kFieldAny, true);     \
    else                                                     \
      KFD_LAUNCH1(S_, FWD_, false, kfd::kFieldAny, true);    \
  } while (0)
  if (fp64)
    KFD_LAUNCH(double, true);
  else
    KFD_LAUNCH(float, true);
  if (between) {
    const hipError_t e = hipEventRecord(between, stream);
    if (e != hipSuccess) return e;
  }
  if (fp64)
    KFD_LAUNCH(double, false);
  else
    KFD_LAUNCH(float, false);
#undef KFD_LAUNCH
#undef KFD_LAUNCH1
#undef KFD_LAUNCH2
  if (track_states_out) {
    const int64_t eb = (args->count + 255) / 256;
    hipLaunchKernelGGL(kfd::kf_export_states, dim3((unsigned)eb), dim3(256), 0, stream, *args,
                       track_states_out);
  }
  return hipGetLastError();
}

extern "C" hipError_t kfd_launch_sort(const kfd::FitArgs *args, const KfdLaunch *launch,
                                      hipStream_t stream) {
  if (args->n <= 0) return hipSuccess;
  const unsigned sb = (unsigned)((args->n + 255) / 256);
  uint32_t *hist = launch->sort_scratch, *off = launch->sort_scratch + kfd::kLenBins;
  hipLaunchKernelGGL(kfd::kf_sort_reset, dim3(1), dim3(256), 0, stream, launch->sort_scratch);
  hipLaunchKernelGGL(kfd::kf_len_hist, dim3(sb), dim3(256), 0, stream, *args, hist);
  hipLaunchKernelGGL(kfd::kf_len_scan, dim3(1), dim3(256), 0, stream, hist, off);
  hipLaunchKernelGGL(kfd::kf_len_scatter, dim3(sb), dim3(256), 0, stream, *args, off,
                     launch->order);
  return hipGetLastError();
}
