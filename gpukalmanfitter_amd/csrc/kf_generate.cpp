// kf_generate.cpp -- seeded synthetic track generator (host, libkfgen.so).
//
// The per-track simulation (particle generation, truth propagation, material
// effects, smearing) is kf_gen_track.hpp, shared with the device generator
// (kf_generate_gpu.hip); this file runs it over tracks with OpenMP and holds
// the plane table, misalignment, hit packing and grid-field helpers.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#define KFG_HD
#include "kf_gen_track.hpp"

using kfg::Philox;
using kfg::TrackRng;

extern "C" {


// Fills the nominal plane table (n_planes entries).
void kfg_planes(const kfg_config *cfg, kf_plane *planes) {
  for (int i = 0; i < cfg->n_planes; ++i) {
    kf_plane &p = planes[i];
    std::memset(&p, 0, sizeof p);
    p.center[0] = cfg->first_x + i * cfg->pitch;
    // columns U = y, V = z, T = x (PlaneSurface.ipp:53-61 for normal x)
    const double R[9] = {0, 0, 1, 1, 0, 0, 0, 1, 0};
    std::memcpy(p.rot, R, sizeof R);
    p.half_u = cfg->half_u;
    p.half_v = cfg->half_v;
    p.geo_id = i;
    if (cfg->material) {
      // Test::makeSilicon, include/Test/Helper.hpp:25-29
      p.thickness = cfg->thickness;
      p.x0 = 9.370 * 10.0;
      p.l0 = 46.52 * 10.0;
      p.ar = 28.0855;
      p.z = 14.0;
      p.molar_rho = ((2.329 / 28.0855) * 1.0) / 1000.0;
    }
  }
}

// Applies a seeded misalignment to a plane table: shifts along the local
// axes and small rotations about them (R' = R * Rx(a) * Ry(b) * Rz(c) in the
// plane's local frame).  Used to build the true detector of config 4.
void kfg_misalign(const kfg_config *cfg, kf_plane *planes, uint64_t seed) {
  Philox ph{(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int i = 0; i < cfg->n_planes; ++i) {
    TrackRng r(ph, (uint32_t)i, 77u);
    double sh[3], an[3];
    for (int k = 0; k < 3; ++k) sh[k] = cfg->align_shift_sigma * r.normal();
    for (int k = 0; k < 3; ++k) an[k] = cfg->align_rot_sigma * r.normal();
    kf_plane &p = planes[i];
    double R[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) R[a][b] = p.rot[3 * a + b];
    for (int a = 0; a < 3; ++a)
      p.center[a] += R[a][0] * sh[0] + R[a][1] * sh[1] + R[a][2] * sh[2];
    const double ca = std::cos(an[0]), sa = std::sin(an[0]);
    const double cb = std::cos(an[1]), sb = std::sin(an[1]);
    const double cc = std::cos(an[2]), sc = std::sin(an[2]);
    const double Rx[3][3] = {{1, 0, 0}, {0, ca, -sa}, {0, sa, ca}};
    const double Ry[3][3] = {{cb, 0, sb}, {0, 1, 0}, {-sb, 0, cb}};
    const double Rz[3][3] = {{cc, -sc, 0}, {sc, cc, 0}, {0, 0, 1}};
    double T1[3][3], T2[3][3], T3[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        T1[a][b] = 0;
        for (int k = 0; k < 3; ++k) T1[a][b] += Rx[a][k] * Ry[k][b];
      }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        T2[a][b] = 0;
        for (int k = 0; k < 3; ++k) T2[a][b] += T1[a][k] * Rz[k][b];
      }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        T3[a][b] = 0;
        for (int k = 0; k < 3; ++k) T3[a][b] += R[a][k] * T2[k][b];
      }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) p.rot[3 * a + b] = T3[a][b];
  }
}

// Pass 1: simulates every track; writes the per-track hit count (n_hits_per
// track, used to build the CSR offsets) and all per-track outputs.  Hits are
// written into a dense scratch [n_tracks][max_planes] (u, v, geo id).
// Pass 2 (kfg_pack_hits) compacts them into the CSR arrays.
int kfg_simulate(const kfg_config *cfg, const kf_plane *true_planes,
                 const kfg_field *field, double *start_params, double *start_cov,
                 double *perigee, int32_t *seq_first, int32_t *seq_count,
                 int32_t *n_hits, double *scratch_u, double *scratch_v,
                 int32_t *scratch_geo, double *truth, int32_t n_threads) {
  const int64_t n = cfg->n_tracks;
  kfg::GenField fld;
  fld.grid = field->grid;
  for (int k = 0; k < 3; ++k) {
    fld.b[k] = field->b[k];
    fld.n[k] = field->n[k];
    fld.lo[k] = field->lo[k];
    fld.hi[k] = field->hi[k];
  }
  fld.v = field->values;
  fld.scale = cfg->true_field_scale;
  const Philox ph{(uint32_t)cfg->seed, (uint32_t)(cfg->seed >> 32)};
  const kfg::SimOut o{start_params, start_cov, perigee, seq_first, seq_count,
                      n_hits,       scratch_u, scratch_v, scratch_geo, truth};
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 256)
  for (int64_t it = 0; it < n; ++it) kfg::simulateTrack(*cfg, true_planes, fld, ph, it, o);
  return 0;
}

// Compacts the dense hit scratch into CSR arrays.  hit_offset must have
// n_tracks + 1 entries and be the exclusive prefix sum of n_hits.
int kfg_pack_hits(const kfg_config *cfg, const int32_t *n_hits, const int64_t *hit_offset,
                  const double *scratch_u, const double *scratch_v,
                  const int32_t *scratch_geo, double sigma_u, double sigma_v,
                  int32_t *hit_geo, double *hit_meas, double *hit_cov, int32_t n_threads) {
  const int64_t n = cfg->n_tracks;
  const int64_t m = hit_offset[n];
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#pragma omp parallel for num_threads(n_threads) schedule(static)
  for (int64_t it = 0; it < n; ++it)
    kfg::packTrack(*cfg, n_hits, hit_offset, m, scratch_u, scratch_v, scratch_geo, sigma_u,
                   sigma_v, hit_geo, hit_meas, hit_cov, it);
  return 0;
}

// Fills the config-3 analytic field on a grid: Bz = b0 (1 - 0.1 (r/1 m)^2),
// Br = br (z / 2 m), (Bx, By) = Br (x, y) / r, values in GeV/(e mm).
void kfg_fill_grid(float *values, int32_t nx, int32_t ny, int32_t nz, const double lo[3],
                   const double hi[3], double b0, double br) {
  for (int ix = 0; ix < nx; ++ix)
    for (int iy = 0; iy < ny; ++iy)
      for (int iz = 0; iz < nz; ++iz) {
        const double x = lo[0] + ix * (hi[0] - lo[0]) / (nx - 1);
        const double y = lo[1] + iy * (hi[1] - lo[1]) / (ny - 1);
        const double z = lo[2] + iz * (hi[2] - lo[2]) / (nz - 1);
        const double r = std::sqrt(x * x + y * y);
        const double bz = b0 * (1.0 - 0.1 * (r / 1000.0) * (r / 1000.0));
        const double brad = br * (z / 2000.0);
        const double bx = (r > 0) ? brad * x / r : 0.0;
        const double by = (r > 0) ? brad * y / r : 0.0;
        const size_t o = (((size_t)ix * ny + iy) * nz + iz) * 3;
        values[o] = (float)bx;
        values[o + 1] = (float)by;
        values[o + 2] = (float)bz;
      }
}

// Philox4x32-10 of (counter[4], key[2]) in place, for the known-answer test.
void kfg_philox(uint32_t *counter, const uint32_t *key) {
  Philox ph{key[0], key[1]};
  ph.gen(counter);
}

void kfg_gamma(double k, uint64_t seed, int64_t n, double *out) {
  const Philox ph{(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t i = 0; i < n; ++i) {
    TrackRng r(ph, (uint32_t)i, 8u);
    out[i] = r.gamma(k);
  }
}

void kfg_landau(uint64_t seed, int64_t n, double *out) {
  const Philox ph{(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t i = 0; i < n; ++i) {
    TrackRng r(ph, (uint32_t)i, 7u);
    out[i] = r.landau();
  }
}

}  // extern "C"
