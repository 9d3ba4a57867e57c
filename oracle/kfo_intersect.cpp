// kfo_intersect.cpp -- CPU restatement of the reference's ray x plane
// intersection with convex polygon bounds (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY: the checker of kf_intersect() in tests/ and the
// CPU-baseline leg of bench.py.
//
// Follows PlaneSurface::intersect (include/Surfaces/detail/PlaneSurface.ipp:
// 239-259), PlanarHelper::intersectionEstimate (include/Surfaces/detail/
// PlanarHelper.hpp:29-48), ConvexPolygonBounds::inside
// (include/Surfaces/ConvexPolygonBounds.ipp:99-102) and
// BoundaryCheck::isInside / isInsidePolygon with zero tolerance
// (include/Surfaces/BoundaryCheck.hpp), and the selection of the reference
// driver (Run/GPU/IntersectGPUTest.cu:22-36: status reachable, path >= 0).
#include <omp.h>

#include <chrono>
#include <cmath>
#include <vector>

#include "../include/kf_abi.h"

namespace {

constexpr double kOnSurfaceTolerance = 1e-4;  // s_onSurfaceTolerance, Definitions.hpp:41

enum Status { kUnreachable = 0, kReachable = 1, kOnSurface = 2, kMissed = 3 };

// Status, path and position of the intersection of (p, d) with surface s.
Status intersect(const kf_polygon_plane &s, const double p[3], const double d[3], bool bcheck,
                 double &path, double x[3]) {
  // PlanarHelper::intersectionEstimate: normal = column 2, centre = translation
  const double n[3] = {s.rot[2], s.rot[5], s.rot[8]};
  const double denom = d[0] * n[0] + d[1] * n[1] + d[2] * n[2];
  if (denom == 0.0) return kUnreachable;
  const double cp[3] = {s.center[0] - p[0], s.center[1] - p[1], s.center[2] - p[2]};
  path = (n[0] * cp[0] + n[1] * cp[1] + n[2] * cp[2]) / denom;
  for (int k = 0; k < 3; ++k) x[k] = p[k] + path * d[k];
  Status st = (path * path < kOnSurfaceTolerance * kOnSurfaceTolerance) ? kOnSurface : kReachable;
  if (!bcheck) return st;
  // local position: the first two rotation columns transposed times (x - c)
  const double v[3] = {x[0] - s.center[0], x[1] - s.center[1], x[2] - s.center[2]};
  const double lu = s.rot[0] * v[0] + s.rot[3] * v[1] + s.rot[6] * v[2];
  const double lv = s.rot[1] * v[0] + s.rot[4] * v[1] + s.rot[7] * v[2];
  // isInsidePolygon: side of every edge against the side of edge 0 -> 1
  auto lineSide = [&](int a, int b) {
    const double nx = s.vertices[2 * b] - s.vertices[2 * a];
    const double ny = s.vertices[2 * b + 1] - s.vertices[2 * a + 1];
    const double dx = lu - s.vertices[2 * a], dy = lv - s.vertices[2 * a + 1];
    return std::signbit((nx * dy) - (ny * dx));
  };
  const int nv = s.n_vertices;
  const bool reference = lineSide(0, 1);
  bool inside = true;
  for (int i = 2; i != nv && inside; ++i)
    if (lineSide(i - 1, i) != reference) inside = false;
  if (inside && lineSide(nv - 1, 0) != reference) inside = false;
  return inside ? st : kMissed;
}

}  // namespace

extern "C" {

// Every kept (ray, surface) intersection in (ray, surface) order, as
// kf_intersect(); host arrays; returns the number found (the first
// `capacity` are written).  n_threads <= 0: all threads.
int64_t kfo_intersect(const kf_polygon_plane *surfaces, int64_t n_surfaces,
                      const double *ray_pos, const double *ray_dir, int64_t n_rays,
                      int32_t bcheck, kf_intersection *hits, int64_t capacity,
                      int32_t n_threads) {
  if (n_threads <= 0) n_threads = omp_get_max_threads();
  std::vector<std::vector<kf_intersection>> per(n_rays);
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 16)
  for (int64_t r = 0; r < n_rays; ++r) {
    const double p[3] = {ray_pos[r], ray_pos[n_rays + r], ray_pos[2 * n_rays + r]};
    const double d[3] = {ray_dir[r], ray_dir[n_rays + r], ray_dir[2 * n_rays + r]};
    for (int64_t i = 0; i < n_surfaces; ++i) {
      double path = 0.0, x[3];
      const Status st = intersect(surfaces[i], p, d, bcheck != 0, path, x);
      if (st == kReachable && path >= 0) {
        kf_intersection h;
        h.ray = (int32_t)r;
        h.surface = (int32_t)i;
        h.path = path;
        for (int k = 0; k < 3; ++k) h.position[k] = x[k];
        per[r].push_back(h);
      }
    }
  }
  int64_t m = 0;
  for (int64_t r = 0; r < n_rays; ++r)
    for (const kf_intersection &h : per[r]) {
      if (m < capacity) hits[m] = h;
      ++m;
    }
  return m;
}

// Wall time of kfo_intersect on n_threads threads (seconds), for the CPU
// baseline.
double kfo_time_intersect(const kf_polygon_plane *surfaces, int64_t n_surfaces,
                          const double *ray_pos, const double *ray_dir, int64_t n_rays,
                          int32_t bcheck, int32_t n_threads) {
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t m = kfo_intersect(surfaces, n_surfaces, ray_pos, ray_dir, n_rays, bcheck,
                                  nullptr, 0, n_threads);
  const auto t1 = std::chrono::steady_clock::now();
  (void)m;
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
