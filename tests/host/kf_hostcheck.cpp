// kf_hostcheck.cpp -- CPU build of the fitter's per-track code (TrackFit)
// and of the chi2 adjoint (TrackAdjoint), for the tests: it lets the CPU
// suite check the hand-written adjoint against central finite differences of
// the same forward code without a GPU.  Test infrastructure only: the
// library never loads it.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../gpukalmanfitter_amd/csrc/kf_adjoint.hpp"
#include "../../gpukalmanfitter_amd/csrc/kf_setup.hpp"

using namespace kfd;

extern "C" {

typedef struct kfh_field {
  int32_t grid;
  int32_t n[3];
  double b[3];
  const float *values;  // grid [nx][ny][nz][3]
  double lo[3];
  double hi[3];
  double scale;
} kfh_field;

}  // extern "C"

namespace {

struct Host {
  FitArgs a;
  std::vector<DevPlane> planes;
  std::vector<double> out_params, out_cov, out_chi2, sd, ss, ck, carry_d;
  std::vector<int32_t> out_ndf, out_nmeas, carry_i;
  std::vector<uint32_t> out_status;
  size_t per = 0;

  std::vector<double> tsf;
  std::vector<float> ssf;

  Host(const kf_plane *pl, int32_t np, const kfh_field *f, const kf_tracks *tr,
       const kf_options *opt, bool smoothing = false, bool track_states = false,
       bool mixed = false) {
    std::memset(&a, 0, sizeof a);
    for (int32_t p = 0; p < np; ++p) planes.push_back(devPlane(pl[p]));
    const int64_t n = tr->n_tracks;
    a.n = n;
    a.first = 0;
    a.count = n;
    a.n_hits = tr->n_hits;
    a.start_params = tr->start_params;
    a.start_cov = tr->start_cov;
    a.perigee = tr->perigee;
    a.seq_first = tr->seq_first;
    a.seq_count = tr->seq_count;
    a.hit_offset = tr->hit_offset;
    a.hit_geo_id = tr->hit_geo_id;
    a.hit_meas = tr->hit_meas;
    a.hit_cov = tr->hit_cov;
    a.planes = planes.data();
    a.n_planes = np;
    a.field.grid = f->grid;
    a.field.scale = f->scale;
    for (int k = 0; k < 3; ++k) {
      a.field.b[k] = f->b[k];
      a.field.n[k] = f->n[k];
      a.field.lo[k] = f->lo[k];
      a.field.hi[k] = f->hi[k];
      a.field.w[k] = f->grid ? (f->hi[k] - f->lo[k]) / (f->n[k] - 1) : 0.0;
      a.field.iw[k] = f->grid ? 1.0 / a.field.w[k] : 0.0;
    }
    a.field.values = f->values;
    a.smoothing = smoothing ? opt->smoothing : 0;
    a.outlier_chi2_cut = opt->outlier_chi2_cut;
    a.unique_geo = 1;
    for (int32_t p = 0; p < np && a.unique_geo; ++p)
      for (int32_t q = p + 1; q < np; ++q)
        if (pl[p].geo_id == pl[q].geo_id) {
          a.unique_geo = 0;
          break;
        }
    a.ms = opt->multiple_scattering;
    a.eloss = opt->energy_loss;
    a.max_steps = opt->max_steps;
    a.max_rk_trials = (uint32_t)opt->max_rk_trials;
    a.max_states = opt->max_states > 0 ? opt->max_states : np;
    a.mass = opt->mass;
    a.tolerance = opt->tolerance;
    a.step_size_cutoff = opt->step_size_cutoff;
    a.max_step_size = opt->max_step_size;
    a.target_tolerance = opt->target_tolerance;
    a.overstep_limit = opt->overstep_limit;
    physicsConstants(a);
    out_params.resize(6 * n);
    out_cov.resize(21 * n);
    out_chi2.resize(n);
    out_ndf.resize(n);
    out_nmeas.resize(n);
    out_status.resize(n);
    a.out_params = out_params.data();
    a.out_cov = out_cov.data();
    a.out_chi2 = out_chi2.data();
    a.out_ndf = out_ndf.data();
    a.out_nmeas = out_nmeas.data();
    a.out_status = out_status.data();
    const size_t chunks = (size_t)(n + 63) / 64;
    per = (size_t)a.max_states * 64;
    sd.resize(chunks * per * KF_SD_N);
    ss.resize(chunks * per * KF_SS_N);
    ck.resize(chunks * per * KF_CK_N);
    carry_i.resize(KF_CARRY_I * n);
    carry_d.resize(KF_CARRY_D * n);
    if (mixed) {
      ssf.resize(chunks * per * KF_SS_N);
      a.tss = ssf.data();
    } else {
      a.tss = ss.data();
    }
    if (track_states) {
      tsf.resize(chunks * per * KF_TS_STRIDE);
      a.tsf = tsf.data();
    }
    a.tsd = sd.data();
    a.tsc = ck.data();
    a.carry_i = carry_i.data();
    a.carry_d = carry_d.data();
  }

  void forward() {
    for (int64_t i = 0; i < a.n; ++i) {
      const size_t ch = (size_t)(i >> 6);
      TrackFit<double, true> tf(a, i, a.tsd + ch * per * KF_SD_N,
                                (double *)a.tss + ch * per * KF_SS_N, nullptr,
                                a.tsc + ch * per * KF_CK_N, (uint32_t)(i & 63), i);
      tf.runForward();
    }
  }

  // The whole fit, both kernels' programs one track at a time.
  template <typename S, bool PS>
  void fit() {
    a.tsc = nullptr;
    for (int pass = 0; pass < 2; ++pass)
      for (int64_t i = 0; i < a.n; ++i) {
        const size_t ch = (size_t)(i >> 6);
        double *d = a.tsd + ch * per * KF_SD_N;
        S *s = (S *)a.tss + ch * per * KF_SS_N;
        double *f = a.tsf ? a.tsf + ch * per * KF_TS_STRIDE : nullptr;
        if (pass == 0) {
          TrackFit<S, true, 1, PS> tf(a, i, d, s, f, nullptr, (uint32_t)(i & 63), i);
          tf.pls = planes.data();
          tf.runForward();
        } else {
          TrackFit<S, false, 1, PS> tf(a, i, d, s, f, nullptr, (uint32_t)(i & 63), i);
          tf.pls = planes.data();
          tf.runBackward();
        }
      }
  }
};

}  // namespace

extern "C" {

// The fit of every track by the kernels' per-track program (FP64 or mixed;
// store slots by measurement or, with by_plane, by plane position):
// params [6][n], cov [21][n], chi2, ndf, n_measured, status [n], and, when
// track_states is not NULL, the track-state records [max_states][KF_TS_STRIDE][n].
int kfh_fit(const kf_plane *planes, int32_t n_planes, const kfh_field *f, const kf_tracks *tr,
            const kf_options *opt, int32_t by_plane, double *params, double *cov, double *chi2,
            int32_t *ndf, int32_t *n_measured, uint32_t *status, double *track_states) {
  const bool mixed = opt->precision == KF_PRECISION_MIXED;
  Host h(planes, n_planes, f, tr, opt, true, track_states != nullptr, mixed);
  if (mixed)
    by_plane ? h.fit<float, true>() : h.fit<float, false>();
  else
    by_plane ? h.fit<double, true>() : h.fit<double, false>();
  const int64_t n = tr->n_tracks;
  std::memcpy(params, h.out_params.data(), 6 * n * sizeof(double));
  std::memcpy(cov, h.out_cov.data(), 21 * n * sizeof(double));
  std::memcpy(chi2, h.out_chi2.data(), n * sizeof(double));
  std::memcpy(ndf, h.out_ndf.data(), n * sizeof(int32_t));
  std::memcpy(n_measured, h.out_nmeas.data(), n * sizeof(int32_t));
  std::memcpy(status, h.out_status.data(), n * sizeof(uint32_t));
  if (track_states) {
    const int S = h.a.max_states;
    for (int64_t i = 0; i < n; ++i) {
      const double *base = h.tsf.data() + (size_t)(i >> 6) * h.per * KF_TS_STRIDE;
      for (int k = 0; k < h.out_nmeas[i] && k < S; ++k)
        for (int c = 0; c < KF_TS_STRIDE; ++c)
          track_states[((size_t)k * KF_TS_STRIDE + c) * n + i] =
              base[((size_t)k * KF_TS_STRIDE + c) * 64 + (i & 63)];
    }
  }
  return 0;
}

// Summed-chi2 objective per track (forward filter, FP64, smoothing off).
int kfh_chi2(const kf_plane *planes, int32_t n_planes, const kfh_field *f, const kf_tracks *tr,
             const kf_options *opt, double *chi2) {
  Host h(planes, n_planes, f, tr, opt);
  h.forward();
  for (int64_t i = 0; i < tr->n_tracks; ++i) chi2[i] = h.out_chi2[i];
  return 0;
}

// The adjoint: grad_planes [n_planes][6] (du, dv, dw, a, b, c), the field
// scale, the chi2 sum and per-track KF_ADJ_* flags.
int kfh_chi2_gradient(const kf_plane *planes, int32_t n_planes, const kfh_field *f,
                      const kf_tracks *tr, const kf_options *opt, double *grad_planes,
                      double *grad_scale, double *chi2_sum, uint32_t *flags) {
  Host h(planes, n_planes, f, tr, opt);
  h.forward();
  const int64_t n = tr->n_tracks;
  const int cap = KF_ADJ_MAX_SEGMENT_STEPS;
  const size_t chunks = (size_t)(n + 63) / 64;
  std::vector<double> stack(chunks * 64 * (size_t)cap * kSR_N), pa((size_t)n_planes * kPG_N * n),
      sa(n);
  AdjArgs g;
  g.stack = stack.data();
  g.max_seg_steps = cap;
  g.max_seq = n_planes;
  g.plane_adj = pa.data();
  g.scale_adj = sa.data();
  g.status = flags;
  std::vector<double> lds((size_t)kLDS_N * KF_BLOCK);  // one lane's slots at a time
  std::vector<float> ldf((size_t)kLDSF_N * KF_BLOCK);
  for (int64_t i = 0; i < n; ++i) {
    const size_t ch = (size_t)(i >> 6);
    TrackAdjoint ta(h.a, g, i, i, h.a.tsc + ch * h.per * KF_CK_N,
                    stack.data() + ch * (size_t)cap * kSR_N * 64, (uint32_t)(i & 63), lds.data(),
                    ldf.data());
    ta.run();
  }
  double chi = 0.0, s = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    chi += h.out_chi2[i];
    s += sa[i];
  }
  for (int p = 0; p < n_planes; ++p) {
    double acc[kPG_N] = {0};
    for (int64_t i = 0; i < n; ++i) {
      const int q = p - tr->seq_first[i];
      if (q >= 0 && q < tr->seq_count[i] && q < n_planes)
        for (int k = 0; k < kPG_N; ++k) acc[k] += pa[((size_t)q * kPG_N + k) * n + i];
    }
    const double *r = planes[p].rot;
    for (int j = 0; j < 3; ++j) grad_planes[6 * p + j] = r[j] * acc[0] + r[3 + j] * acc[1] + r[6 + j] * acc[2];
    const double *Rb = acc + kPG_R;
    double gx = 0, gy = 0, gz = 0;
    for (int k = 0; k < 3; ++k) {
      gx += Rb[3 * k + 1] * r[3 * k + 2] - Rb[3 * k + 2] * r[3 * k + 1];
      gy += Rb[3 * k + 2] * r[3 * k + 0] - Rb[3 * k + 0] * r[3 * k + 2];
      gz += Rb[3 * k + 0] * r[3 * k + 1] - Rb[3 * k + 1] * r[3 * k + 0];
    }
    grad_planes[6 * p + 3] = gx;
    grad_planes[6 * p + 4] = gy;
    grad_planes[6 * p + 5] = gz;
  }
  *grad_scale = s;
  *chi2_sum = chi;
  return 0;
}

}  // extern "C"
