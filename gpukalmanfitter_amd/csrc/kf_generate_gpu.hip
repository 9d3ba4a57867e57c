// kf_generate_gpu.hip -- the synthetic workload generator on the GPU
// (libkfgen_gpu.so): the per-track simulation of kf_gen_track.hpp, one thread
// per track, then the hit offsets (exclusive prefix sum of the per-track hit
// counts) and the packing into the CSR hit arrays, all in device memory.  The
// same seeds give the same tracks as the host generator (kf_generate.cpp): the
// Philox streams are identical, floating-point values agree to the rounding of
// the device math library.
#include <hip/hip_runtime.h>

#define KFG_HD __host__ __device__
#include "kf_gen_track.hpp"
#include "kf_scan.hpp"

namespace {

constexpr int kSimBlock = 128;

__global__ void __launch_bounds__(kSimBlock)
    kfg_sim_kernel(kfg_config cfg, const kf_plane *planes, kfg::GenField fld, kfg::Philox ph,
                   kfg::SimOut o) {
  const int64_t it = (int64_t)blockIdx.x * kSimBlock + threadIdx.x;
  if (it >= cfg.n_tracks) return;
  kfg::simulateTrack(cfg, planes, fld, ph, it, o);
}

__global__ void __launch_bounds__(kSimBlock)
    kfg_pack_kernel(kfg_config cfg, const int32_t *n_hits, const int64_t *hit_offset, int64_t m,
                    const double *su, const double *sv, const int32_t *sg, double sigma_u,
                    double sigma_v, int32_t *hit_geo, double *hit_meas, double *hit_cov) {
  const int64_t it = (int64_t)blockIdx.x * kSimBlock + threadIdx.x;
  if (it >= cfg.n_tracks) return;
  kfg::packTrack(cfg, n_hits, hit_offset, m, su, sv, sg, sigma_u, sigma_v, hit_geo, hit_meas,
                 hit_cov, it);
}

kfg::GenField genField(const kfg_config *cfg, const kfg_field *field) {
  kfg::GenField f;
  f.grid = field->grid;
  for (int k = 0; k < 3; ++k) {
    f.b[k] = field->b[k];
    f.n[k] = field->n[k];
    f.lo[k] = field->lo[k];
    f.hi[k] = field->hi[k];
  }
  f.v = field->values;
  f.scale = cfg->true_field_scale;
  return f;
}

}  // namespace

extern "C" {

int kfg_simulate_device(const kfg_config *cfg, const kf_plane *true_planes,
                        const kfg_field *field, double *start_params, double *start_cov,
                        double *perigee, int32_t *seq_first, int32_t *seq_count, int32_t *n_hits,
                        double *scratch_u, double *scratch_v, int32_t *scratch_geo, double *truth,
                        void *stream) {
  if (!cfg || !true_planes || !field || cfg->n_tracks < 0 || cfg->max_planes < 1 ||
      cfg->max_planes > cfg->n_planes || (field->grid && !field->values))
    return -1;
  if (cfg->n_tracks == 0) return 0;
  const kfg::Philox ph{(uint32_t)cfg->seed, (uint32_t)(cfg->seed >> 32)};
  const kfg::SimOut o{start_params, start_cov, perigee, seq_first, seq_count,
                      n_hits,       scratch_u, scratch_v, scratch_geo, truth};
  const unsigned blocks = (unsigned)((cfg->n_tracks + kSimBlock - 1) / kSimBlock);
  hipLaunchKernelGGL(kfg_sim_kernel, dim3(blocks), dim3(kSimBlock), 0, (hipStream_t)stream, *cfg,
                     true_planes, genField(cfg, field), ph, o);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int64_t kfg_offsets_scratch_bytes(int64_t n) { return kscan::scratchBytes(n); }

int kfg_offsets_device(int64_t n, const int32_t *n_hits, int64_t *hit_offset, void *scratch,
                       void *stream) {
  if (n < 0 || !hit_offset || (n > 0 && (!n_hits || !scratch))) return -1;
  return kscan::exclusiveScan(n, n_hits, hit_offset, scratch, (hipStream_t)stream) == hipSuccess
             ? 0
             : -2;
}

int kfg_pack_hits_device(const kfg_config *cfg, const int32_t *n_hits, const int64_t *hit_offset,
                         int64_t n_hits_total, const double *scratch_u, const double *scratch_v,
                         const int32_t *scratch_geo, double sigma_u, double sigma_v,
                         int32_t *hit_geo, double *hit_meas, double *hit_cov, void *stream) {
  if (!cfg || cfg->n_tracks < 0 || n_hits_total < 0) return -1;
  if (cfg->n_tracks == 0) return 0;
  const unsigned blocks = (unsigned)((cfg->n_tracks + kSimBlock - 1) / kSimBlock);
  hipLaunchKernelGGL(kfg_pack_kernel, dim3(blocks), dim3(kSimBlock), 0, (hipStream_t)stream,
                     *cfg, n_hits, hit_offset, n_hits_total, scratch_u, scratch_v, scratch_geo,
                     sigma_u, sigma_v, hit_geo, hit_meas, hit_cov);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
