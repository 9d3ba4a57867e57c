// kf_scan.hpp -- exclusive prefix sum of int32 counts into int64 offsets on
// the device (three kernels: tile scans, a scan of the tile totals, the
// add-back), used to turn per-item counts into CSR offsets.  Deterministic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kscan {

constexpr int kScanBlock = 256;
constexpr int kScanPerThread = 4;
constexpr int kScanTile = kScanBlock * kScanPerThread;

// Exclusive scan of one tile of kScanTile counts (block-local), tile total to
// sums[tile].
static __global__ void __launch_bounds__(kScanBlock)
    scan_tiles(int64_t n, const int32_t *counts, int64_t *out, int64_t *sums) {
  __shared__ int64_t part[kScanBlock];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
  int64_t v[kScanPerThread];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    v[k] = (base + k < n) ? (int64_t)counts[base + k] : 0;
    s += v[k];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan of the per-thread sums
  for (int w = 1; w < kScanBlock; w <<= 1) {
    const int64_t add = (threadIdx.x >= (unsigned)w) ? part[threadIdx.x - w] : 0;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  int64_t run = part[threadIdx.x] - s;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = part[threadIdx.x];
}

// Exclusive scan of the tile totals in place (one block, a sequential carry
// over chunks of kScanBlock); out[n] = the grand total.
static __global__ void __launch_bounds__(kScanBlock)
    scan_sums(int64_t tiles, int64_t *sums, int64_t n, int64_t *out) {
  __shared__ int64_t part[kScanBlock];
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < tiles; c0 += kScanBlock) {
    const int64_t i = c0 + threadIdx.x;
    const int64_t v = (i < tiles) ? sums[i] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int w = 1; w < kScanBlock; w <<= 1) {
      const int64_t add = (threadIdx.x >= (unsigned)w) ? part[threadIdx.x - w] : 0;
      __syncthreads();
      part[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < tiles) sums[i] = carry + part[threadIdx.x] - v;
    const int64_t chunk = part[kScanBlock - 1];
    __syncthreads();
    carry += chunk;
  }
  if (threadIdx.x == 0) out[n] = carry;
}

static __global__ void __launch_bounds__(kScanBlock)
    scan_add(int64_t n, const int64_t *sums, int64_t *out) {
  const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  if (i < n) out[i] += sums[i / kScanTile];
}


// Bytes of device scratch the scan of n counts needs.
inline int64_t scratchBytes(int64_t n) {
  return ((n + kScanTile - 1) / kScanTile + 1) * (int64_t)sizeof(int64_t);
}

// out[0..n] = exclusive prefix sum of counts[0..n-1], out[n] = total.
inline hipError_t exclusiveScan(int64_t n, const int32_t *counts, int64_t *out, void *scratch,
                                hipStream_t s) {
  int64_t *sums = (int64_t *)scratch;
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles > 0)
    hipLaunchKernelGGL(scan_tiles, dim3((unsigned)tiles), dim3(kScanBlock), 0, s, n, counts,
                       out, sums);
  hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanBlock), 0, s, tiles, sums, n, out);
  if (n > 0)
    hipLaunchKernelGGL(scan_add, dim3((unsigned)((n + kScanBlock - 1) / kScanBlock)),
                       dim3(kScanBlock), 0, s, n, (const int64_t *)sums, out);
  return hipGetLastError();
}

}  // namespace kscan
