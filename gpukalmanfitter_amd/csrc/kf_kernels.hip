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

template <typename S>
__global__ void __launch_bounds__(KF_BLOCK) kf_forward_kernel(FitArgs args) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const ChunkBases<S> cb(args, i);
  if (i >= args.n) return;
  TrackFit<S, true> tf(args, i, cb.d, cb.s, cb.f, cb.c, (uint32_t)(i & 63));
  tf.runForward();
}

template <typename S>
__global__ void __launch_bounds__(KF_BLOCK) kf_backward_kernel(FitArgs args) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const ChunkBases<S> cb(args, i);
  if (i >= args.n) return;
  TrackFit<S, false> tf(args, i, cb.d, cb.s, cb.f, cb.c, (uint32_t)(i & 63));
  tf.runBackward();
}

// Copies the public track-state records into kf_results.track_states
// ([state][KF_TS_STRIDE][n]).
__global__ void __launch_bounds__(256) kf_export_states(FitArgs args, double *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= args.n) return;
  const int nm = args.out_nmeas[i];
  const double *base = args.tsf + (size_t)(i >> 6) * (size_t)args.max_states * KF_TS_STRIDE * 64u;
  const uint32_t lane = (uint32_t)(i & 63);
  for (int k = 0; k < nm && k < args.max_states; ++k)
    for (int f = 0; f < KF_TS_STRIDE; ++f)
      out[((size_t)k * KF_TS_STRIDE + f) * (size_t)args.n + i] =
          base[((uint32_t)k * KF_TS_STRIDE + (uint32_t)f) * 64u + lane];
}

}  // namespace kfd

extern "C" hipError_t kfd_launch_fit(const kfd::FitArgs *args, int precision, hipStream_t stream,
                                     double *track_states_out, hipEvent_t between) {
  const int64_t blocks = (args->n + KF_BLOCK - 1) / KF_BLOCK;
  if (blocks <= 0) return hipSuccess;
  const dim3 grid((unsigned)blocks), block(KF_BLOCK);
  if (precision == KF_PRECISION_FP64)
    hipLaunchKernelGGL(kfd::kf_forward_kernel<double>, grid, block, 0, stream, *args);
  else
    hipLaunchKernelGGL(kfd::kf_forward_kernel<float>, grid, block, 0, stream, *args);
  if (between) {
    const hipError_t e = hipEventRecord(between, stream);
    if (e != hipSuccess) return e;
  }
  if (precision == KF_PRECISION_FP64)
    hipLaunchKernelGGL(kfd::kf_backward_kernel<double>, grid, block, 0, stream, *args);
  else
    hipLaunchKernelGGL(kfd::kf_backward_kernel<float>, grid, block, 0, stream, *args);
  if (track_states_out) {
    const int64_t eb = (args->n + 255) / 256;
    hipLaunchKernelGGL(kfd::kf_export_states, dim3((unsigned)eb), dim3(256), 0, stream, *args,
                       track_states_out);
  }
  return hipGetLastError();
}
