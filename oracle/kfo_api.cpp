// kfo_api.cpp -- C entry points of the oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// CPU-baseline leg of bench.py.  The batch entry point takes the same
// kf_tracks / kf_results structs as the product's kf_fit(), but with host
// pointers, and fits the tracks with OpenMP over tracks like the reference's
// CPU driver (Run/CPU/KalmanFitterCPUTest.cpp:193-214).
#include <omp.h>

#include <chrono>
#include <cstring>
#include <vector>

#include "../include/kf_abi.h"
#include "kfo_fitter.hpp"

using namespace kfo;

#ifdef KFO_COUNTING
namespace kfo {
thread_local uint64_t g_flops = 0;
thread_local uint64_t g_flops32 = 0;
thread_local uint64_t g_trans = 0;
thread_local int g_in32 = 0;
thread_local uint64_t g_flops_bwd = 0;
thread_local uint64_t g_flops32_bwd = 0;
thread_local int g_bwd = 0;
}  // namespace kfo
#endif

extern "C" {

typedef struct kfo_field {
  int32_t grid;
  int32_t n[3];
  double b[3];
  const float *values;
  double lo[3];
  double hi[3];
  double scale;
} kfo_field;

int kfo_is_counting(void) {
#ifdef KFO_COUNTING
  return 1;
#else
  return 0;
#endif
}

}  // extern "C"

namespace {

Field makeField(const kfo_field *f) {
  Field fl;
  fl.grid = f->grid != 0;
  fl.b = v3(f->b[0], f->b[1], f->b[2]);
  fl.values = f->values;
  for (int d = 0; d < 3; ++d) {
    fl.n[d] = f->n[d];
    fl.lo[d] = f->lo[d];
    fl.hi[d] = f->hi[d];
  }
  fl.scale = f->scale;
  return fl;
}

Surface makePlane(const kf_plane &p) {
  Surface s;
  s.kind = SurfaceKind::Plane;
  for (int i = 0; i < 3; ++i) {
    s.center[i] = p.center[i];
    for (int j = 0; j < 3; ++j) s.rot(i, j) = p.rot[3 * i + j];
  }
  if (p.thickness > 0.0) {
    Material m;
    m.x0 = p.x0;
    m.l0 = p.l0;
    m.ar = p.ar;
    m.z = p.z;
    m.molarRho = p.molar_rho;
    s.slab = MaterialSlab(m, p.thickness);
  }
  s.geoId = p.geo_id;
  return s;
}

Surface makeLine(const double c[3]) {
  Surface s;
  s.kind = SurfaceKind::Line;
  s.rot = M33::identity();
  s.center = v3(c[0], c[1], c[2]);
  return s;
}

M66 unpackSym(const double *v, int64_t stride) {
  M66 m;
  int k = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) {
      m(i, j) = v[(int64_t)k * stride];
      m(j, i) = v[(int64_t)k * stride];
      ++k;
    }
  return m;
}

void packUpper(const M66 &m, double *out, int64_t stride) {
  int k = 0;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) out[(int64_t)(k++) * stride] = kval(m(i, j));
}

PropOptions makePropOptions(const kf_options *o) {
  PropOptions po;
  po.absPdgCode = o->abs_pdg;
  po.mass = o->mass;
  po.maxSteps = (unsigned)o->max_steps;
  po.maxRungeKuttaStepTrials = (unsigned)o->max_rk_trials;
  po.maxStepSize = o->max_step_size;
  po.targetTolerance = o->target_tolerance;
  po.tolerance = o->tolerance;
  po.stepSizeCutOff = o->step_size_cutoff;
  po.overstepLimit = o->overstep_limit;
  return po;
}

void writeBound(const BoundParams &b, double *rec, int off, int64_t stride) {
  for (int k = 0; k < 6; ++k) rec[(int64_t)(off + k) * stride] = kval(b.par[k]);
  packUpper(b.cov, rec + (int64_t)(off + 6) * stride, stride);
}

}  // namespace

extern "C" {

// Fits n tracks (host memory).  track_states, when not NULL, receives
// [max_states][KF_TS_STRIDE][n] like kf_results.track_states; steps (may be
// NULL) the number of propagation steps of each track; flops / trans (may be
// NULL; counting build only) the counted flops and transcendentals per track,
// flop_classes ([3][n]) the part of the flops in the float class of the
// mixed mode, the part from the smoother on, and the float-class part of
// the latter.
int kfo_fit_batch(const kf_plane *planes, int32_t n_planes, const kfo_field *field,
                  const kf_tracks *tr, const kf_options *opt, const kf_results *res,
                  int32_t max_states, int32_t n_threads, int32_t *steps,
                  uint64_t *flops, uint64_t *trans, uint64_t *flop_classes) {
  if (!planes || !field || !tr || !opt || !res) return KF_E_INVALID;
  const int64_t n = tr->n_tracks;
  std::vector<Surface> surf(n_planes);
  for (int i = 0; i < n_planes; ++i) surf[i] = makePlane(planes[i]);
  const Field fl = makeField(field);
  const KalmanFitter kf(&fl);
  const PropOptions po = makePropOptions(opt);
  if (n_threads <= 0) n_threads = omp_get_max_threads();

#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 64)
  for (int64_t it = 0; it < n; ++it) {
#ifdef KFO_COUNTING
    const uint64_t f0 = g_flops, t0 = g_trans, g0 = g_flops32, b0 = g_flops_bwd,
                   c0 = g_flops32_bwd;
    g_bwd = 0;
#endif
    const int first = tr->seq_first[it], count = tr->seq_count[it];
    std::vector<const Surface *> seq(count > 0 ? count : 0);
    for (int k = 0; k < count; ++k) seq[k] = &surf[first + k];
    const double pc[3] = {tr->perigee[it], tr->perigee[n + it], tr->perigee[2 * n + it]};
    Surface line = makeLine(pc);
    V6 sp;
    for (int k = 0; k < 6; ++k) sp[k] = tr->start_params[k * n + it];
    M66 sc = unpackSym(tr->start_cov + it, n);
    std::vector<SourceLink> sls;
    const int64_t h0 = tr->hit_offset[it], h1 = tr->hit_offset[it + 1];
    for (int64_t h = h0; h < h1; ++h) {
      SourceLink sl;
      sl.meas[0] = tr->hit_meas[h];
      sl.meas[1] = tr->hit_meas[tr->n_hits + h];
      sl.cov(0, 0) = tr->hit_cov[h];
      sl.cov(0, 1) = tr->hit_cov[tr->n_hits + h];
      sl.cov(1, 0) = tr->hit_cov[tr->n_hits + h];
      sl.cov(1, 1) = tr->hit_cov[2 * tr->n_hits + h];
      sl.geoId = tr->hit_geo_id[h];
      sls.push_back(sl);
    }
    BoundParams start = makeBound(sc, sp, &line);
    FitOptions fo;
    fo.smoothing = opt->smoothing != 0;
    fo.multipleScattering = opt->multiple_scattering != 0;
    fo.energyLoss = opt->energy_loss != 0;
    fo.referenceSurface = &line;
    fo.outlierChi2Cut = opt->outlier_chi2_cut;
    FitResult r;
    const int cap = max_states > 0 ? max_states : count;
    r.states.resize(cap > 0 ? cap : 1);
    uint32_t status = 0;
    bool ok = false;
    if (count <= 0 || sp[4] == Real(0.0) || cap < 1 || sls.empty()) {
      status |= KF_STATUS_BAD_INPUT;
    } else {
      ok = kf.fit(sls, start, fo, po, r, seq.data(), (unsigned)count);
    }
    if (ok) status |= KF_STATUS_OK;
    if (r.smoothed) status |= KF_STATUS_SMOOTHED;
    if (r.finished) status |= KF_STATUS_FINISHED;
    if (!r.result) status |= KF_STATUS_RESULT_FAIL;
    if (r.stepFailed) status |= KF_STATUS_STEP_FAIL;
    if (r.maxStepsReached) status |= KF_STATUS_MAX_STEPS;
    if (r.navBreak && r.measurementStates != sls.size()) status |= KF_STATUS_NAV_BREAK;
    if (r.smootherNaN) status |= KF_STATUS_SMOOTHER_NAN;
    if (res->status) res->status[it] = status;
    if (res->n_measured) res->n_measured[it] = (int32_t)r.measurementStates;
    if (res->ndf) {
      int32_t used = 0;
      for (unsigned k = 0; k < r.measurementStates; ++k) used += r.states[k].outlier ? 0 : 1;
      res->ndf[it] = 2 * used;
    }
    if (res->chi2) {
      Real c = 0.0;
      for (unsigned k = 0; k < r.measurementStates; ++k)
        if (!r.states[k].outlier) c = c + r.states[k].chi2;
      res->chi2[it] = kval(c);
    }
    if (res->params)
      for (int k = 0; k < 6; ++k) res->params[k * n + it] = kval(r.fittedParameters.par[k]);
    if (res->cov) packUpper(r.fittedParameters.cov, res->cov + it, n);
    if (res->track_states) {
      const int64_t stride = n;
      for (unsigned k = 0; k < r.measurementStates && (int)k < cap; ++k) {
        double *rec = res->track_states + (int64_t)k * KF_TS_STRIDE * n + it;
        const TrackStateRec &ts = r.states[k];
        writeBound(ts.predicted, rec, KF_TS_PREDICTED, stride);
        writeBound(ts.filtered, rec, KF_TS_FILTERED, stride);
        if (r.smoothed && !r.smootherNaN)
          writeBound(ts.smoothed, rec, KF_TS_SMOOTHED, stride);
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < 6; ++b)
            rec[(int64_t)(KF_TS_JACOBIAN + 6 * a + b) * stride] = kval(ts.jacobian(a, b));
        rec[(int64_t)KF_TS_PATH * stride] = kval(ts.pathLength);
        rec[(int64_t)KF_TS_CHI2 * stride] = kval(ts.chi2);
        rec[(int64_t)KF_TS_GEO_ID * stride] = (double)ts.sl.geoId;
      }
    }
    if (steps) steps[it] = (int32_t)r.steps;
#ifdef KFO_COUNTING
    if (flops) flops[it] = g_flops - f0;
    if (trans) trans[it] = g_trans - t0;
    if (flop_classes) {
      flop_classes[it] = g_flops32 - g0;
      flop_classes[n + it] = g_flops_bwd - b0;
      flop_classes[2 * n + it] = g_flops32_bwd - c0;
    }
#else
    (void)flops;
    (void)trans;
    (void)flop_classes;
#endif
  }
  return KF_OK;
}

// Wall-clock seconds of kfo_fit_batch over the same arguments (timer around
// the fit loop only, as Run/CPU/KalmanFitterCPUTest.cpp:192-215).
double kfo_time_fit_batch(const kf_plane *planes, int32_t n_planes,
                          const kfo_field *field, const kf_tracks *tr,
                          const kf_options *opt, const kf_results *res,
                          int32_t n_threads) {
  auto t0 = std::chrono::steady_clock::now();
  int rc = kfo_fit_batch(planes, n_planes, field, tr, opt, res, 0, n_threads,
                         nullptr, nullptr, nullptr, nullptr);
  auto t1 = std::chrono::steady_clock::now();
  if (rc != KF_OK) return -1.0;
  return std::chrono::duration<double>(t1 - t0).count();
}

// Propagates start parameters (bound, on the z-parallel line through
// `perigee`) through the planes in order without fitting; at every plane that
// is reached the bound state is taken (covariance transport included).
// Outputs per reached plane k: params[k*6..], cov[k*36..] (full),
// jac[k*36..] (bound-to-bound from the previous bound state), pos[k*3..],
// dir[k*3..], path[k].  Returns the number of planes reached.
int kfo_propagate(const kf_plane *planes, int32_t n_planes, const kfo_field *field,
                  const double start_par[6], const double start_cov[36],
                  const double perigee[3], const kf_options *opt, double *params,
                  double *cov, double *jac, double *pos, double *dir, double *path) {
  std::vector<Surface> surf(n_planes);
  std::vector<const Surface *> seq(n_planes);
  for (int i = 0; i < n_planes; ++i) {
    surf[i] = makePlane(planes[i]);
    seq[i] = &surf[i];
  }
  const Field fl = makeField(field);
  Stepper stepper(&fl);
  Navigator nav;
  Surface line = makeLine(perigee);
  V6 sp;
  M66 sc;
  for (int k = 0; k < 6; ++k) sp[k] = start_par[k];
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) sc(a, b) = start_cov[6 * a + b];
  BoundParams start = makeBound(sc, sp, &line);
  PropState st;
  st.options = makePropOptions(opt);
  st.stepping = stepper.makeState(start, 1, st.options.maxStepSize, st.options.tolerance);
  st.navigation.startSurface = &line;
  st.navigation.sequence = seq.data();
  st.navigation.size = (unsigned)n_planes;
  st.navigation.targetSurface = nullptr;
  int reached = 0;
  auto onSurface = [&]() {
    const Surface *s = st.navigation.currentSurface;
    if (!s) return;
    BoundParams b;
    M66 j;
    Real pl;
    stepper.boundState(st.stepping, *s, b, j, pl);
    for (int k = 0; k < 6; ++k) params[reached * 6 + k] = kval(b.par[k]);
    for (int a = 0; a < 6; ++a)
      for (int c = 0; c < 6; ++c) {
        cov[reached * 36 + 6 * a + c] = kval(b.cov(a, c));
        jac[reached * 36 + 6 * a + c] = kval(j(a, c));
      }
    for (int k = 0; k < 3; ++k) {
      pos[reached * 3 + k] = kval(st.stepping.pos[k]);
      dir[reached * 3 + k] = kval(st.stepping.dir[k]);
    }
    path[reached] = kval(pl);
    ++reached;
  };
  nav.status(st, stepper);
  onSurface();
  nav.target(st, stepper);
  for (unsigned steps = 0; steps < st.options.maxSteps; ++steps) {
    if (st.navigation.navigationBreak) break;
    stepper.step(st);
    nav.status(st, stepper);
    onSurface();
    nav.target(st, stepper);
  }
  return reached;
}

// One GainMatrixUpdater call on a plane (full 6x6 covariances, row-major).
void kfo_gain_update(const double pred[6], const double pcov[36], const double meas[2],
                     const double V[4], double filt[6], double fcov[36], double *chi2) {
  Surface plane;  // x-normal plane at the origin: rot columns (y, z, x)
  plane.rot = M33::zero();
  plane.rot(1, 0) = 1.0;
  plane.rot(2, 1) = 1.0;
  plane.rot(0, 2) = 1.0;
  KalmanFitter kf(nullptr);
  TrackStateRec ts;
  V6 p;
  M66 c;
  for (int k = 0; k < 6; ++k) p[k] = pred[k];
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) c(a, b) = pcov[6 * a + b];
  ts.predicted = makeBound(c, p, &plane);
  ts.sl.meas[0] = meas[0];
  ts.sl.meas[1] = meas[1];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) ts.sl.cov(a, b) = V[2 * a + b];
  kf.gainMatrixUpdate(ts, &plane, 0.0);
  for (int k = 0; k < 6; ++k) filt[k] = kval(ts.filtered.par[k]);
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) fcov[6 * a + b] = kval(ts.filtered.cov(a, b));
  *chi2 = kval(ts.chi2);
}

// GainMatrixSmoother over n states given as arrays (full 6x6, row-major).
// Returns 0 on success, 1 if the gain had a NaN.
int kfo_rts_smooth(int32_t n, const double *fpar, const double *fcov,
                   const double *ppar, const double *pcov, const double *jac,
                   double *spar, double *scov) {
  Surface plane;
  plane.rot = M33::zero();
  plane.rot(1, 0) = 1.0;
  plane.rot(2, 1) = 1.0;
  plane.rot(0, 2) = 1.0;
  KalmanFitter kf(nullptr);
  std::vector<TrackStateRec> st(n);
  for (int i = 0; i < n; ++i) {
    V6 a, b;
    M66 ca, cb, j;
    for (int k = 0; k < 6; ++k) {
      a[k] = fpar[6 * i + k];
      b[k] = ppar[6 * i + k];
    }
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) {
        ca(r, c) = fcov[36 * i + 6 * r + c];
        cb(r, c) = pcov[36 * i + 6 * r + c];
        j(r, c) = jac[36 * i + 6 * r + c];
      }
    st[i].filtered = makeBound(ca, a, &plane);
    st[i].predicted = makeBound(cb, b, &plane);
    st[i].jacobian = j;
  }
  bool nanFound = false;
  kf.smooth(st, (unsigned)n, nanFound);
  if (nanFound) return 1;
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < 6; ++k) spar[6 * i + k] = kval(st[i].smoothed.par[k]);
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) scov[36 * i + 6 * r + c] = kval(st[i].smoothed.cov(r, c));
  }
  return 0;
}

// Material formulas (RPP) for a slab of `thickness` of the given material.
// which: 0 = Bethe mean energy loss, 1 = Highland theta0, 2 = Landau sigma(q/p)
double kfo_material(int32_t which, double thickness, double x0, double l0, double ar,
                    double z, double molar_rho, double mass, double qop, double q) {
  Material m;
  m.x0 = x0;
  m.l0 = l0;
  m.ar = ar;
  m.z = z;
  m.molarRho = molar_rho;
  MaterialSlab s(m, thickness);
  if (which == 0) return kval(energyLossBethe(s, mass, qop, q));
  if (which == 1) return kval(multipleScatteringTheta0(s, mass, qop, q));
  return kval(landauSigmaQOverP(s, mass, qop, q));
}

// One Actor::materialInteractor call (KalmanFitter.hpp:405-449) on the plane
// `plane` for a stepper state at pos/dir with momentum p, charge q and bound
// covariance cov[36] (row-major), navigation direction nav (+1 / -1).  stage:
// 0 = the plane is an intermediate surface, 1 = it is the navigation's start
// surface, -1 = its target surface (the split-factor cases of
// PointwiseMaterialInteraction.hpp:87-108).  Writes the momentum and the bound
// covariance after the call.
void kfo_material_interact(const kf_plane *plane, const double pos[3], const double dir[3],
                           double p, int32_t q, double mass, const double cov[36], int32_t nav,
                           int32_t stage, int32_t multiple_scattering, int32_t energy_loss,
                           double *p_out, double cov_out[36]) {
  Surface surf = makePlane(*plane);
  KalmanFitter kf(nullptr);
  PropState st;
  st.options.mass = mass;
  StepperState &s = st.stepping;
  s.pos = v3(pos[0], pos[1], pos[2]);
  s.dir = v3(dir[0], dir[1], dir[2]);
  s.p = p;
  s.q = q;
  s.navDir = nav;
  s.covTransport = true;
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) s.cov(a, b) = cov[6 * a + b];
  st.navigation.currentSurface = &surf;
  st.navigation.startSurface = (stage == 1) ? &surf : nullptr;
  st.navigation.targetSurface = (stage == -1) ? &surf : nullptr;
  FitOptions fo;
  fo.multipleScattering = multiple_scattering != 0;
  fo.energyLoss = energy_loss != 0;
  kf.materialInteractor(&surf, st, fo);
  *p_out = kval(s.p);
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) cov_out[6 * a + b] = kval(s.cov(a, b));
}

// Field lookup (for the grid trilinear tests).
void kfo_field_get(const kfo_field *field, const double pos[3], double b[3]) {
  Field fl = makeField(field);
  V3 r = fl.get(v3(pos[0], pos[1], pos[2]));
  for (int k = 0; k < 3; ++k) b[k] = kval(r[k]);
}

}  // extern "C"
