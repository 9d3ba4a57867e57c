// kf_generate.h -- C interface of the seeded synthetic track generator
// (libkfgen.so, kf_generate.cpp): the workloads of the benchmark, the tests
// and the kf_run driver.
#pragma once

#include <stdint.h>

#include "../../include/kf_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

// Material simulation of kfg_config.material.  REFERENCE: the reference's
// simulator (Highland space-angle deflection, Landau Bethe-Bloch loss at the
// most probable value, Bethe-Heitler loss).  FIT_MODEL: exactly what the fit's
// material update assumes (Gaussian projected angles of width theta0 and
// theta0 / sin(theta), Gaussian loss of the mean Bethe value and the Landau
// FWHM width, nothing else), for pull tests of the material update;
// FIT_MODEL_ELOSS: its energy loss alone (for fits without multiple scattering).
#define KFG_MATERIAL_REFERENCE 1
#define KFG_MATERIAL_FIT_MODEL 2
#define KFG_MATERIAL_FIT_MODEL_ELOSS 3

// Generator configuration.  Planes: plane i of the detector sits at
// x = first_x + i * pitch with normal +x (local u = y, v = z).
typedef struct kfg_config {
  int64_t n_tracks;
  uint64_t seed;
  int32_t n_planes;      // planes in the detector
  int32_t min_planes;    // per-track sequence length ~ U{min_planes..max_planes}
  int32_t max_planes;
  int32_t material;      // 0: none; planes carry silicon of `thickness` (in the table and
                         // simulated) with KFG_MATERIAL_REFERENCE or KFG_MATERIAL_FIT_MODEL
  double first_x, pitch, half_u, half_v;
  double thickness;      // mm
  double hole_prob;      // probability to drop a hit
  double p_min, p_max, phi_min, phi_max, eta_min, eta_max;
  double vertex_sigma;   // per axis, mm
  double hit_sigma_u, hit_sigma_v;
  double sigma_d0, sigma_z0, sigma_phi, sigma_theta, sigma_p_rel, sigma_t;
  double mass;
  // misalignment of the simulated (true) planes, config 4
  double align_shift_sigma;  // mm, per axis
  double align_rot_sigma;    // rad, per local axis
  double true_field_scale;   // field scale used by the simulation
} kfg_config;

typedef struct kfg_field {
  int32_t grid;
  int32_t n[3];
  double b[3];
  const float *values;
  double lo[3];
  double hi[3];
} kfg_field;

// Fills the nominal plane table (n_planes entries).
void kfg_planes(const kfg_config *cfg, kf_plane *planes);
// Misaligns the planes (shifts and small rotations, seeded).
void kfg_misalign(const kfg_config *cfg, kf_plane *planes, uint64_t seed);
// Simulates every track; hits go to dense scratch [n_tracks][max_planes].
int kfg_simulate(const kfg_config *cfg, const kf_plane *true_planes, const kfg_field *field,
                 double *start_params, double *start_cov, double *perigee, int32_t *seq_first,
                 int32_t *seq_count, int32_t *n_hits, double *scratch_u, double *scratch_v,
                 int32_t *scratch_geo, double *truth, int32_t n_threads);
// Compacts the dense hit scratch into the CSR hit arrays of kf_tracks.
int kfg_pack_hits(const kfg_config *cfg, const int32_t *n_hits, const int64_t *hit_offset,
                  const double *scratch_u, const double *scratch_v, const int32_t *scratch_geo,
                  double sigma_u, double sigma_v, int32_t *hit_geo, double *hit_meas,
                  double *hit_cov, int32_t n_threads);
// The analytic grid field (Bz = b0 (1 - 0.1 (r/1 m)^2), Br = br z / 2 m).
void kfg_fill_grid(float *values, int32_t nx, int32_t ny, int32_t nz, const double lo[3],
                   const double hi[3], double b0, double br);
// Philox4x32-10 of (counter[4], key[2]) in place.
void kfg_philox(uint32_t *counter, const uint32_t *key);
// n standard Landau variates of the generator's sampler (stream `seed`).
void kfg_landau(uint64_t seed, int64_t n, double *out);
// n Gamma(k, 1) variates of the generator's sampler (stream `seed`).
void kfg_gamma(double k, uint64_t seed, int64_t n, double *out);

// The same generator on the GPU (libkfgen_gpu.so, kf_generate_gpu.hip): one
// thread per track, asynchronous on `stream` (a hipStream_t); every pointer is
// device memory, including true_planes [n_planes] and field->values.  Return
// 0 or a negative code.
int kfg_simulate_device(const kfg_config *cfg, const kf_plane *true_planes, const kfg_field *field,
                        double *start_params, double *start_cov, double *perigee,
                        int32_t *seq_first, int32_t *seq_count, int32_t *n_hits,
                        double *scratch_u, double *scratch_v, int32_t *scratch_geo,
                        double *truth, void *stream);
// hit_offset[0..n] = exclusive prefix sum of n_hits[0..n-1] (hit_offset[n] =
// the number of hits); scratch: kfg_offsets_scratch_bytes(n) bytes of device
// memory.
int64_t kfg_offsets_scratch_bytes(int64_t n);
int kfg_offsets_device(int64_t n, const int32_t *n_hits, int64_t *hit_offset, void *scratch,
                       void *stream);
// kfg_pack_hits on the GPU; n_hits_total = hit_offset[n].
int kfg_pack_hits_device(const kfg_config *cfg, const int32_t *n_hits, const int64_t *hit_offset,
                         int64_t n_hits_total, const double *scratch_u, const double *scratch_v,
                         const int32_t *scratch_geo, double sigma_u, double sigma_v,
                         int32_t *hit_geo, double *hit_meas, double *hit_cov, void *stream);

#ifdef __cplusplus
}
#endif
