// kf_adjoint.hip -- kernels of the chi2 adjoint (kf_adjoint.hpp) and their
// launcher: the reverse sweep per track, a deterministic reduction over
// tracks, and the conversion to per-plane alignment parameters.
#include <hip/hip_runtime.h>

#include "kf_adjoint.hpp"

namespace kfd {

__global__ void __launch_bounds__(KF_BLOCK) kf_adjoint_kernel(FitArgs args, AdjArgs adj) {
  __shared__ double lds[kLDS_N * KF_BLOCK];
  __shared__ float ldf[kLDSF_N * KF_BLOCK];
  const int64_t local = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t chunk = __builtin_amdgcn_readfirstlane((uint32_t)(local >> 6));
  double *ck = args.tsc + chunk * (size_t)args.max_states * KF_CK_N * 64u;
  double *stk = adj.stack + chunk * (size_t)adj.max_seg_steps * kSR_N * 64u;
  if (local >= args.count) return;
  TrackAdjoint ta(args, adj, args.first + local, local, ck, stk, (uint32_t)(local & 63),
                  lds + threadIdx.x, ldf + threadIdx.x);
  ta.run();
}

// Two-stage sum over the tracks of a launch.  Stage 1: block (p, b) sums
// slice b of the positions for quantity p (p < n_planes: the 12 adjoints of
// plane p; p = n_planes: the field-scale adjoint; p = n_planes + 1: the chi2)
// into partial[p][b]; each thread sums a fixed stride, then a fixed-order
// tree.  Stage 2 adds the partials of each p in slice order, onto the sums of
// the previous launches unless this is the first.  The result does not depend
// on scheduling.
constexpr int kRedThreads = 256;
constexpr int kRedSlices = KF_ADJ_RED_SLICES;
__global__ void __launch_bounds__(kRedThreads) kf_adjoint_reduce(FitArgs args, AdjArgs adj,
                                                                 double *partial) {
  __shared__ double red[kRedThreads][kPG_N];
  const int p = blockIdx.x;
  const int b = blockIdx.y;
  const int64_t per = (args.count + kRedSlices - 1) / kRedSlices;
  const int64_t lo = (int64_t)b * per, hi = (lo + per < args.count) ? lo + per : args.count;
  double acc[kPG_N];
#pragma unroll
  for (int f = 0; f < kPG_N; ++f) acc[f] = 0.0;
  for (int64_t l = lo + threadIdx.x; l < hi; l += kRedThreads) {
    const int64_t i = args.first + l;
    if (p < args.n_planes) {
      const int q = p - args.seq_first[i];
      if (q >= 0 && q < args.seq_count[i] && q < adj.max_seq) {
#pragma unroll
        for (int f = 0; f < kPG_N; ++f)
          acc[f] += adj.plane_adj[((size_t)q * kPG_N + f) * args.count + l];
      }
    } else if (p == args.n_planes) {
      acc[0] += adj.scale_adj[l];
    } else {
      acc[0] += args.out_chi2[i];
    }
  }
#pragma unroll
  for (int f = 0; f < kPG_N; ++f) red[threadIdx.x][f] = acc[f];
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int f = 0; f < kPG_N; ++f) red[threadIdx.x][f] += red[threadIdx.x + w][f];
    }
    __syncthreads();
  }
  if (threadIdx.x < kPG_N)
    partial[((size_t)p * kRedSlices + b) * kPG_N + threadIdx.x] = red[0][threadIdx.x];
}

__global__ void __launch_bounds__(kPG_N) kf_adjoint_reduce2(const double *partial, double *sums,
                                                            int accumulate) {
  const int p = blockIdx.x;
  double s = 0.0;
  for (int b = 0; b < kRedSlices; ++b) s += partial[((size_t)p * kRedSlices + b) * kPG_N + threadIdx.x];
  double *o = sums + (size_t)p * kPG_N + threadIdx.x;
  *o = accumulate ? *o + s : s;
}

// dX/d(du, dv, dw) = R^T dX/dc; dX/d(a, b, c) = <dX/dR, R G> with G the
// generators of Rx, Ry, Rz at zero (R' = R Rx(a) Ry(b) Rz(c)).
__global__ void kf_adjoint_finish(FitArgs args, const double *sums, double *grad_planes,
                                  double *grad_scale, double *chi2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < args.n_planes) {
    const double *s = sums + (size_t)p * kPG_N;
    const double *r = args.planes[p].r;  // row-major R[k][j] = r[3k + j]
    double g[6];
#pragma unroll
    for (int j = 0; j < 3; ++j) g[j] = r[j] * s[0] + r[3 + j] * s[1] + r[6 + j] * s[2];
    const double *Rb = s + kPG_R;
    // <Rb, R Gx> = Rb[:,1] . R[:,2] - Rb[:,2] . R[:,1]
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      gx += Rb[3 * k + 1] * r[3 * k + 2] - Rb[3 * k + 2] * r[3 * k + 1];
      gy += Rb[3 * k + 2] * r[3 * k + 0] - Rb[3 * k + 0] * r[3 * k + 2];
      gz += Rb[3 * k + 0] * r[3 * k + 1] - Rb[3 * k + 1] * r[3 * k + 0];
    }
    g[3] = gx;
    g[4] = gy;
    g[5] = gz;
#pragma unroll
    for (int j = 0; j < 6; ++j) grad_planes[(size_t)p * 6 + j] = g[j];
  }
  if (p == 0) {
    if (grad_scale) grad_scale[0] = sums[(size_t)args.n_planes * kPG_N];
    if (chi2) chi2[0] = sums[(size_t)(args.n_planes + 1) * kPG_N];
  }
}

}  // namespace kfd

extern "C" hipError_t kfd_launch_adjoint(const kfd::FitArgs *args, const kfd::AdjArgs *adj,
                                         double *sums, int accumulate, hipStream_t stream) {
  const int64_t blocks = (args->count + KF_BLOCK - 1) / KF_BLOCK;
  if (blocks > 0)
    hipLaunchKernelGGL(kfd::kf_adjoint_kernel, dim3((unsigned)blocks), dim3(KF_BLOCK), 0, stream,
                       *args, *adj);
  double *partial = sums + (size_t)(args->n_planes + 2) * kfd::kPG_N;
  hipLaunchKernelGGL(kfd::kf_adjoint_reduce, dim3((unsigned)(args->n_planes + 2), kfd::kRedSlices),
                     dim3(kfd::kRedThreads), 0, stream, *args, *adj, partial);
  hipLaunchKernelGGL(kfd::kf_adjoint_reduce2, dim3((unsigned)(args->n_planes + 2)),
                     dim3(kfd::kPG_N), 0, stream, (const double *)partial, sums, accumulate);
  return hipGetLastError();
}

extern "C" hipError_t kfd_launch_adjoint_finish(const kfd::FitArgs *args, const double *sums,
                                                double *grad_planes, double *grad_scale,
                                                double *chi2, hipStream_t stream) {
  hipLaunchKernelGGL(kfd::kf_adjoint_finish, dim3((unsigned)((args->n_planes + 63) / 64)),
                     dim3(64), 0, stream, *args, sums, grad_planes, grad_scale, chi2);
  return hipGetLastError();
}
