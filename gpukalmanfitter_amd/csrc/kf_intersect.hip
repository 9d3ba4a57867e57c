// kf_intersect.hip -- batched intersection of rays with bounded planes
// (kf_intersect in include/kf_abi.h): the reference's intersection test
// (Run/GPU/IntersectGPUTest.cu:22-48, 116-172) over many rays.
//
// Work is split over (ray, surface chunk) pairs so that a launch fills the
// chip however few rays there are: a block takes 256 rays and one chunk of
// consecutive surfaces.  Every lane of a wave reads the same surface at the
// same time, so surface data comes through scalar loads into SGPRs (no LDS,
// no barriers).  Each pair first meets a bounding-sphere test (the ray's line
// against the sphere about the surface centre that holds the polygon, with a
// margin far above rounding): the full intersection runs only when a lane of
// the wave passes it, which for a detector mesh is a few percent of the
// surfaces.
//
// Output order is (ray, surface) without atomics: the count pass keeps, per
// (ray, chunk), the number of kept intersections and the first kIxSlots
// surface indices; an exclusive scan of the counts in (ray, chunk) order gives
// each (ray, chunk) its output offset; the write pass recomputes only the
// slotted surfaces (a (ray, chunk) with more than kIxSlots runs over its chunk
// again).
//
// This file is compiled without multiply-add contraction: every value is the
// correctly rounded result of the same operations, in the same order, as the
// reference formulation, so the kept pairs, paths and positions are
// bit-identical to an uncontracted FP64 build of it (double division on
// gfx950 is correctly rounded).
#include <hip/hip_runtime.h>

#include "../../include/kf_abi.h"
#include "kf_device.hpp"
#include "kf_scan.hpp"

namespace kfd {

constexpr int kIxBlock = 256;                // rays per block (count pass)
constexpr int kIxSlots = 8;                  // surface indices kept per (ray, chunk)
// sphere tests per group in the count pass: 4 keeps the kernel at 42 VGPRs
// (8 waves per SIMD hide the scalar-load latency; 8 tests per group need 98)
constexpr int kIxGroup = 4;
constexpr int64_t kIxTargetBlocks = 4096;    // count-pass blocks wanted (16 per CU)
constexpr int64_t kIxMinChunk = 256;         // surfaces per chunk, lower bound
constexpr int64_t kIxMaxChunk = 8192;        // and upper bound
// bounding-sphere margin: relative on the radius^2, and on |centre|^2 and
// |ray origin|^2 (rounding of the full test is ~1e-15 of these)
constexpr double kIxRelMargin = 1e-6;
constexpr double kIxAbsMargin = 3e-12;

struct IxPlan {
  int64_t chunk;     // surfaces per chunk
  int64_t n_chunks;  // chunks (>= 1)
  int64_t pairs;     // n_rays * n_chunks
};

inline IxPlan ixPlan(int64_t n_rays, int64_t n_surfaces) {
  const int64_t ray_blocks = (n_rays + kIxBlock - 1) / kIxBlock;
  const int64_t want = ray_blocks > 0 ? (kIxTargetBlocks + ray_blocks - 1) / ray_blocks : 1;
  int64_t chunk = (n_surfaces + want - 1) / want;
  chunk = chunk < kIxMinChunk ? kIxMinChunk : (chunk > kIxMaxChunk ? kIxMaxChunk : chunk);
  int64_t n_chunks = (n_surfaces + chunk - 1) / chunk;
  if (n_chunks < 1) n_chunks = 1;
  return {chunk, n_chunks, n_rays * n_chunks};
}

inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

// PlaneSurface::intersect (PlaneSurface.ipp:239-259) with
// PlanarHelper::intersectionEstimate (PlanarHelper.hpp:29-48) and the
// reference driver's selection: reachable (path^2 >= tolerance^2), path >= 0,
// inside the convex polygon when bcheck (BoundaryCheck::isInsidePolygon).
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

// Per surface: centre and the squared radius of the sphere about it that
// holds every vertex, with the margin (|centre|^2 part included).
// Entries n_surfaces .. n_surfaces + kIxGroup - 1 are the zero padding.
__global__ void __launch_bounds__(256)
    kf_ix_spheres(const kf_polygon_plane *__restrict__ surfaces, int64_t n_surfaces,
                  double4 *__restrict__ spheres) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_surfaces + kIxGroup) return;
  if (i >= n_surfaces) {
    spheres[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    return;
  }
  const kf_polygon_plane &s = surfaces[i];
  const int nv =
      (s.n_vertices < 3) ? 3 : ((s.n_vertices > KF_POLY_MAX) ? KF_POLY_MAX : s.n_vertices);
  double r2 = 0.0;
  for (int k = 0; k < nv; ++k) {
    const double u = s.vertices[2 * k], v = s.vertices[2 * k + 1];
    const double q = u * u + v * v;
    r2 = q > r2 ? q : r2;
  }
  const double cc =
      s.center[0] * s.center[0] + s.center[1] * s.center[1] + s.center[2] * s.center[2];
  spheres[i] = make_double4(s.center[0], s.center[1], s.center[2],
                            r2 * (1.0 + kIxRelMargin) + kIxAbsMargin * cc);
}

struct Ray {
  double p[3], d[3];
  double m[3];  // moment p x d
  double dd;    // |d|^2
  double pmdd;  // margin part of |p|^2, times |d|^2
};

KFD_INLINE Ray loadRay(const double *ray_pos, const double *ray_dir, int64_t n_rays, int64_t r) {
  Ray y;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    y.p[k] = ray_pos[(size_t)k * n_rays + r];
    y.d[k] = ray_dir[(size_t)k * n_rays + r];
  }
  y.m[0] = y.p[1] * y.d[2] - y.p[2] * y.d[1];
  y.m[1] = y.p[2] * y.d[0] - y.p[0] * y.d[2];
  y.m[2] = y.p[0] * y.d[1] - y.p[1] * y.d[0];
  y.dd = y.d[0] * y.d[0] + y.d[1] * y.d[1] + y.d[2] * y.d[2];
  y.pmdd = kIxAbsMargin * (y.p[0] * y.p[0] + y.p[1] * y.p[1] + y.p[2] * y.p[2]) * y.dd;
  return y;
}

// The ray's line passes within the sphere (or a component is not finite):
// |(c - p) x d|^2 <= R^2 |d|^2, with (c - p) x d = c x d - p x d.  Only a
// filter (the margin in R^2 is far above the rounding of either form), so
// it uses fused multiply-adds: 11 operations per pair.
KFD_INLINE bool sphereCandidate(const double4 &q, const Ray &y) {
  const double e0 = fma(q.y, y.d[2], fma(-q.z, y.d[1], -y.m[0]));
  const double e1 = fma(q.z, y.d[0], fma(-q.x, y.d[2], -y.m[1]));
  const double e2 = fma(q.x, y.d[1], fma(-q.y, y.d[0], -y.m[2]));
  const double lhs = fma(e0, e0, fma(e1, e1, e2 * e2));
  return !(fma(-q.w, y.dd, lhs) > y.pmdd);
}

// Count pass: block (ray block, chunk); per (ray, chunk) the number of kept
// intersections and the first kIxSlots surface indices.
__global__ void __launch_bounds__(kIxBlock)
    kf_ix_count(const kf_polygon_plane *__restrict__ surfaces, const double4 *__restrict__ spheres,
                int64_t n_surfaces, int64_t chunk, int64_t n_chunks,
                const double *__restrict__ ray_pos, const double *__restrict__ ray_dir,
                int64_t n_rays, int bcheck, int32_t *__restrict__ counts,
                int32_t *__restrict__ slots) {
  const int64_t c = blockIdx.x % n_chunks;
  const int64_t r = (blockIdx.x / n_chunks) * kIxBlock + threadIdx.x;
  if (r >= n_rays) return;
  const Ray y = loadRay(ray_pos, ray_dir, n_rays, r);
  const int64_t t = r * n_chunks + c, pairs = n_rays * n_chunks;
  // surface indices fit int32 (checked by the caller): 32-bit uniform loop
  const int s0 = (int)(c * chunk);
  const int s1 = (int)((c * chunk + chunk < n_surfaces) ? c * chunk + chunk : n_surfaces);
  int32_t cnt = 0;
  auto test = [&](int i) {
    double path, x[3];
    if (intersectKept(surfaces[i], y.p, y.d, bcheck != 0, path, x)) {
      if (cnt < kIxSlots) slots[(int64_t)cnt * pairs + t] = i;
      ++cnt;
    }
  };
  if (!bcheck) {
    for (int i = s0; i < s1; ++i) test(i);
    counts[t] = cnt;
    return;
  }
  // groups of kIxGroup surfaces: the sphere tests first (their loads issued
  // together), then the full test of the candidates in surface order
  // (the sphere array is padded by kIxGroup entries: the loads past s1 stay in
  // bounds and their bits are masked off)
  for (int i0 = s0; i0 < s1; i0 += kIxGroup) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < kIxGroup; ++j)
      if (sphereCandidate(spheres[i0 + j], y)) m |= 1u << j;
    if (s1 - i0 < kIxGroup) m &= (1u << (s1 - i0)) - 1u;
    if (m == 0) continue;
#pragma unroll 1
    for (int j = 0; j < kIxGroup; ++j)
      if ((m >> j) & 1u) test(i0 + j);
  }
  counts[t] = cnt;
}

KFD_INLINE void writeHit(kf_intersection *h, int64_t r, int64_t i, double path,
                         const double x[3]) {
  h->ray = (int32_t)r;
  h->surface = (int32_t)i;
  h->path = path;
  h->position[0] = x[0];
  h->position[1] = x[1];
  h->position[2] = x[2];
}

// Write pass: one thread per (ray, chunk), in (ray, chunk) order.
__global__ void __launch_bounds__(256)
    kf_ix_write(const kf_polygon_plane *__restrict__ surfaces, const double4 *__restrict__ spheres,
                int64_t n_surfaces, int64_t chunk, int64_t n_chunks,
                const double *__restrict__ ray_pos, const double *__restrict__ ray_dir,
                int64_t n_rays, int bcheck, const int32_t *__restrict__ counts,
                const int32_t *__restrict__ slots, const int64_t *__restrict__ offsets,
                kf_intersection *__restrict__ hits, int64_t capacity) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t pairs = n_rays * n_chunks;
  if (t >= pairs) return;
  const int32_t cnt = counts[t];
  int64_t out = offsets[t];
  if (cnt == 0 || out >= capacity) return;
  const int64_t r = t / n_chunks, c = t - r * n_chunks;
  const Ray y = loadRay(ray_pos, ray_dir, n_rays, r);
  double path, x[3];
  if (cnt <= kIxSlots) {
    for (int k = 0; k < cnt && out < capacity; ++k) {
      const int64_t i = slots[(int64_t)k * pairs + t];
      if (intersectKept(surfaces[i], y.p, y.d, bcheck != 0, path, x))
        writeHit(hits + out++, r, i, path, x);
    }
    return;
  }
  const int64_t s0 = c * chunk;
  const int64_t s1 = (s0 + chunk < n_surfaces) ? s0 + chunk : n_surfaces;
  for (int64_t i = s0; i < s1 && out < capacity; ++i) {
    if (bcheck && !sphereCandidate(spheres[i], y)) continue;
    if (intersectKept(surfaces[i], y.p, y.d, bcheck != 0, path, x))
      writeHit(hits + out++, r, i, path, x);
  }
}

struct IxScratch {
  double4 *spheres;
  int32_t *counts, *slots;
  int64_t *offsets;
  void *scan;
  int64_t bytes;
};

// Scratch layout from `base` (nullptr: sizes only).
inline IxScratch ixScratch(void *base, int64_t n_rays, int64_t n_surfaces) {
  const IxPlan pl = ixPlan(n_rays, n_surfaces);
  IxScratch s;
  char *b = (char *)base;
  int64_t at = 0;
  s.spheres = (double4 *)(b + at);
  at += align256((n_surfaces + kIxGroup) * (int64_t)sizeof(double4));
  s.counts = (int32_t *)(b + at);
  at += align256(pl.pairs * (int64_t)sizeof(int32_t));
  s.slots = (int32_t *)(b + at);
  at += align256(pl.pairs * kIxSlots * (int64_t)sizeof(int32_t));
  s.offsets = (int64_t *)(b + at);
  at += align256((pl.pairs + 1) * (int64_t)sizeof(int64_t));
  s.scan = b + at;
  at += align256(kscan::scratchBytes(pl.pairs));
  s.bytes = at;
  return s;
}

}  // namespace kfd

extern "C" int64_t kfd_intersect_scratch_bytes(int64_t n_rays, int64_t n_surfaces) {
  return kfd::ixScratch(nullptr, n_rays, n_surfaces).bytes;
}

// Launches the sphere set-up, the count pass, the scan and the write pass;
// n_hits[0] = the total.
extern "C" hipError_t kfd_launch_intersect(const kf_polygon_plane *surfaces, int64_t n_surfaces,
                                           const double *ray_pos, const double *ray_dir,
                                           int64_t n_rays, int bcheck, kf_intersection *hits,
                                           int64_t capacity, int64_t *n_hits, void *scratch,
                                           hipStream_t s) {
  using namespace kfd;
  const IxPlan pl = ixPlan(n_rays, n_surfaces);
  const IxScratch sc = ixScratch(scratch, n_rays, n_surfaces);
  hipLaunchKernelGGL(kf_ix_spheres, dim3((unsigned)((n_surfaces + kIxGroup + 255) / 256)),
                     dim3(256), 0, s, surfaces, n_surfaces, sc.spheres);
  const int64_t ray_blocks = (n_rays + kIxBlock - 1) / kIxBlock;
  if (ray_blocks > 0)
    hipLaunchKernelGGL(kf_ix_count, dim3((unsigned)(ray_blocks * pl.n_chunks)), dim3(kIxBlock), 0,
                       s, surfaces, (const double4 *)sc.spheres, n_surfaces, pl.chunk, pl.n_chunks,
                       ray_pos, ray_dir, n_rays, bcheck, sc.counts, sc.slots);
  hipError_t e = kscan::exclusiveScan(pl.pairs, sc.counts, sc.offsets, sc.scan, s);
  if (e != hipSuccess) return e;
  if (pl.pairs > 0 && capacity > 0)
    hipLaunchKernelGGL(kf_ix_write, dim3((unsigned)((pl.pairs + 255) / 256)), dim3(256), 0, s,
                       surfaces, (const double4 *)sc.spheres, n_surfaces, pl.chunk, pl.n_chunks,
                       ray_pos, ray_dir, n_rays, bcheck, (const int32_t *)sc.counts,
                       (const int32_t *)sc.slots, (const int64_t *)sc.offsets, hits, capacity);
  e = hipMemcpyAsync(n_hits, sc.offsets + pl.pairs, sizeof(int64_t), hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}
