// kf_intersect.hip -- batched intersection of rays with bounded planes
// (kf_intersect in include/kf_abi.h): the reference's intersection test
// (Run/GPU/IntersectGPUTest.cu:22-48, 116-172) over many rays.
//
// One thread per ray.  A block stages a tile of surfaces in LDS and every
// thread tests its ray against the tile (all lanes read the same surface: LDS
// broadcast).  Two passes over the same tests keep the output in (ray,
// surface) order without atomics: the first counts each ray's kept
// intersections, a prefix scan turns the counts into offsets, the second
// writes them.
#include <hip/hip_runtime.h>

#include "../../include/kf_abi.h"
#include "kf_device.hpp"
#include "kf_scan.hpp"

namespace kfd {

constexpr int kIxBlock = 256;  // rays per block
constexpr int kIxTile = 64;    // surfaces per LDS tile

// PlaneSurface::intersect (PlaneSurface.ipp:239-259) with
// PlanarHelper::intersectionEstimate (PlanarHelper.hpp:29-48) and the
// reference driver's selection: reachable (path^2 >= tolerance^2), path >= 0,
// inside the convex polygon when bcheck (BoundaryCheck::isInsidePolygon).
// This file is compiled without multiply-add contraction: every value is the
// correctly rounded result of the same operations, in the same order, as the
// reference formulation, so the kept pairs, paths and positions are
// bit-identical to an uncontracted FP64 build of it (double division on
// gfx950 is correctly rounded).
KFD_INLINE bool intersectKept(const kf_polygon_plane &s, const double p[3], const double d[3],
                              bool bcheck, double &path, double x[3]) {
  const double n[3] = {s.rot[2], s.rot[5], s.rot[8]};
  const double denom = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
  if (denom == 0.0) return false;  // unreachable
  const double cp[3] = {s.center[0] - p[0], s.center[1] - p[1], s.center[2] - p[2]};
  path = (n[0] * cp[0] + n[1] * cp[1] + n[2] * cp[2]) / denom;
  if (path * path < kOnSurfaceTolerance * kOnSurfaceTolerance) return false;  // on surface
  if (path < 0.0) return false;
#pragma unroll
  for (int k = 0; k < 3; ++k) x[k] = p[k] + path * d[k];
  if (!bcheck) return true;
  const double v[3] = {x[0] - s.center[0], x[1] - s.center[1], x[2] - s.center[2]};
  const double lu = s.rot[0] * v[0] + s.rot[3] * v[1] + s.rot[6] * v[2];
  const double lv = s.rot[1] * v[0] + s.rot[4] * v[1] + s.rot[7] * v[2];
  // the point is inside when it is on the same side of every edge
  // (signbit of the z component of edge x (point - edge start))
  const int nv =
      (s.n_vertices < 3) ? 3 : ((s.n_vertices > KF_POLY_MAX) ? KF_POLY_MAX : s.n_vertices);
  bool ref = false, inside = true;
  for (int i = 0; i < nv; ++i) {
    const int j = (i + 1 < nv) ? i + 1 : 0;
    const double e0 = s.vertices[2 * j] - s.vertices[2 * i];
    const double e1 = s.vertices[2 * j + 1] - s.vertices[2 * i + 1];
    const double q0 = lu - s.vertices[2 * i], q1 = lv - s.vertices[2 * i + 1];
    const bool side = signbit((e0 * q1) - (e1 * q0));
    if (i == 0) ref = side;
    inside = inside && (side == ref);
  }
  return inside;
}

// PASS 0: count per ray; PASS 1: write at offset[ray].
template <int PASS>
__global__ void __launch_bounds__(kIxBlock)
    kf_intersect_kernel(const kf_polygon_plane *surfaces, int64_t n_surfaces,
                        const double *ray_pos, const double *ray_dir, int64_t n_rays, int bcheck,
                        int32_t *counts, const int64_t *offsets, kf_intersection *hits,
                        int64_t capacity) {
  __shared__ kf_polygon_plane tile[kIxTile];
  const int64_t r = (int64_t)blockIdx.x * kIxBlock + threadIdx.x;
  const bool active = r < n_rays;
  double p[3] = {0, 0, 0}, d[3] = {0, 0, 0};
  if (active) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      p[k] = ray_pos[(size_t)k * n_rays + r];
      d[k] = ray_dir[(size_t)k * n_rays + r];
    }
  }
  int64_t out = (PASS == 1 && active) ? offsets[r] : 0;
  int32_t cnt = 0;
  constexpr int kWords = (int)(sizeof(kf_polygon_plane) / sizeof(double));
  for (int64_t t0 = 0; t0 < n_surfaces; t0 += kIxTile) {
    const int nt = (n_surfaces - t0 < kIxTile) ? (int)(n_surfaces - t0) : kIxTile;
    __syncthreads();
    const double *src = (const double *)(surfaces + t0);
    double *dst = (double *)tile;
    for (int w = threadIdx.x; w < nt * kWords; w += kIxBlock) dst[w] = src[w];
    __syncthreads();
    if (!active) continue;
    for (int i = 0; i < nt; ++i) {
      double path, x[3];
      if (intersectKept(tile[i], p, d, bcheck != 0, path, x)) {
        if (PASS == 1 && out < capacity) {
          kf_intersection &h = hits[out];
          h.ray = (int32_t)r;
          h.surface = (int32_t)(t0 + i);
          h.path = path;
          h.position[0] = x[0];
          h.position[1] = x[1];
          h.position[2] = x[2];
        }
        ++out;
        ++cnt;
      }
    }
  }
  if (PASS == 0 && active) counts[r] = cnt;
}

}  // namespace kfd

extern "C" int64_t kfd_intersect_scratch_bytes(int64_t n_rays) {
  const int64_t counts = (n_rays * (int64_t)sizeof(int32_t) + 255) / 256 * 256;
  const int64_t offsets = ((n_rays + 1) * (int64_t)sizeof(int64_t) + 255) / 256 * 256;
  return counts + offsets + kscan::scratchBytes(n_rays);
}

// Launches the two passes and the scan; n_hits[0] = the total.
extern "C" hipError_t kfd_launch_intersect(const kf_polygon_plane *surfaces, int64_t n_surfaces,
                                           const double *ray_pos, const double *ray_dir,
                                           int64_t n_rays, int bcheck, kf_intersection *hits,
                                           int64_t capacity, int64_t *n_hits, void *scratch,
                                           hipStream_t s) {
  char *base = (char *)scratch;
  int32_t *counts = (int32_t *)base;
  base += (n_rays * (int64_t)sizeof(int32_t) + 255) / 256 * 256;
  int64_t *offsets = (int64_t *)base;
  base += ((n_rays + 1) * (int64_t)sizeof(int64_t) + 255) / 256 * 256;
  const unsigned blocks = (unsigned)((n_rays + kfd::kIxBlock - 1) / kfd::kIxBlock);
  if (blocks > 0)
    hipLaunchKernelGGL(kfd::kf_intersect_kernel<0>, dim3(blocks), dim3(kfd::kIxBlock), 0, s,
                       surfaces, n_surfaces, ray_pos, ray_dir, n_rays, bcheck, counts,
                       (const int64_t *)nullptr, (kf_intersection *)nullptr, (int64_t)0);
  hipError_t e = kscan::exclusiveScan(n_rays, counts, offsets, base, s);
  if (e != hipSuccess) return e;
  if (blocks > 0)
    hipLaunchKernelGGL(kfd::kf_intersect_kernel<1>, dim3(blocks), dim3(kfd::kIxBlock), 0, s,
                       surfaces, n_surfaces, ray_pos, ray_dir, n_rays, bcheck, counts,
                       (const int64_t *)offsets, hits, capacity);
  e = hipMemcpyAsync(n_hits, offsets + n_rays, sizeof(int64_t), hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}
