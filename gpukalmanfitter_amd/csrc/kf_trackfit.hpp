// kf_trackfit.hpp -- the per-track Kalman fit program (one track per lane),
// shared by the fit kernels (kf_kernels.hip) and the chi2 adjoint
// (kf_adjoint.hip).
//
// The per-lane program is the reference's propagation loop (Propagator.ipp:
// 3-96): step -> navigator status -> Kalman actor -> aborters -> navigator
// target, until the Kalman aborter fires.  The loop is arranged so lanes
// re-converge before the expensive actor work: the inner loop runs the
// iterations in which the actor has nothing to do (no surface reached, no
// finalisation, target not reached), and only exits when a lane has a surface
// update, the smoother or the final perigee transport due.  Every lane still
// executes exactly the reference's per-iteration sequence.
#pragma once

#include <hip/hip_runtime.h>

#include "kf_device.hpp"
#include "kf_kernels.hpp"

namespace kfd {

// FWD = true: the forward kernel (start, steps, filter at every surface) that
// stops where the reference's actor would call finalize() and saves a carry
// record; FWD = false: the backward kernel (smoother, then propagation back to
// the perigee) that resumes from that record.
enum { kResumeNone = 0, kResumePreLoop = 1, kResumeInLoop = 2, kResumePostLoop = 3 };

// L: lanes per track (1, 2 or 4; kf_device.hpp, lane groups).  The lanes of a
// group run this program on the same data; the free transport matrix of each
// RK step is built column by column across them, and only the group's first
// lane writes to memory.
template <typename S, bool FWD, int L = 1, bool PS = false, bool OOL = false, int FK = kFieldAny,
          bool REC = true>
struct TrackFit {
  static constexpr bool kMixed = sizeof(S) < sizeof(double);
  const FitArgs &a;
  const int64_t i;
  // index of this lane in its group, and whether it writes the track's
  // records and outputs
  uint32_t rank = 0;
  bool lead = true;
  KFD_INLINE void setRank(uint32_t r) {
    rank = r;
    lead = (r == 0);
  }

  // ---- stepper state (EigenStepper::State, EigenStepper.hpp:27-120) ----
  double pos[3], dir[3], p, t, q;
  int navDir;
  CStep ss;
  double pathAcc;
  // bound covariance (packed upper triangle).  S: in the mixed mode every
  // consumer of it between surfaces (the next sandwich, the stores) takes it
  // in float, so it is kept in float (half the registers)
  S cov[21];
  // free-from-bound Jacobian J (8x6).  Non-constant entries only:
  // A = J[0:3, 0:5], b = J[3, 4], G = J[4:7, 2:5]; J[3,5] = J[7,4] = 1, the
  // rest is zero between surfaces.
  S A[3][5];
  S b;
  S G[3][3];
  // propagation derivative (dir, dt/ds, k4, 0) of the last step; zero after
  // a Jacobian re-initialisation
  S k4d[3];
  S dtds;
  bool derivSet;
  bool stepFailed;

  // ---- navigation (DirectNavigator::State) ----
  int next, size, seqFirst;
  int cur;  // -1: none, -2: the perigee target, >= 0: plane-table index
  bool targetReached, navBreak;

  // ---- KalmanFitterResult ----
  int measured;
  int outliers;    // measured states the outlier finder left without an update
  int firstPlane;  // plane-table index of the first measured state
  // store slots (KF_SD_* / KF_SS_*).  PS = false: the slot of measured state
  // k is k.  PS = true (kf_schedule.store_by_plane): k + min(holes before
  // it, max_states - hits), i.e. its plane's position in the sequence
  // whenever the sequence fits max_states slots.  holes counts the planes
  // passed without a hit; slotLast / slotPrev are the slots of the last two
  // measured states.
  int holes, slotLast, slotPrev;
  bool smoothed, finished, result, smootherNaN;
  int64_t h0, h1;
  // hit search: while every match so far was the next hit in order and the
  // plane geo ids are distinct (inOrder), no hit before hitNext can match a
  // later plane, so the search starts at hitNext, whose data is loaded ahead
  // (pf*) while the track steps to the next surface
  int64_t hitNext;
  bool inOrder, pfValid;
  int32_t pfGeo;
  double pfM0, pfM1, pfV00, pfV01, pfV11;
  KFD_INLINE void prefetchHit() {
    pfValid = hitNext < h1;
    if (pfValid) {
      pfGeo = a.hit_geo_id[hitNext];
      pfM0 = a.hit_meas[hitNext];
      pfM1 = a.hit_meas[a.n_hits + hitNext];
      pfV00 = a.hit_cov[hitNext];
      pfV01 = a.hit_cov[a.n_hits + hitNext];
      pfV11 = a.hit_cov[2 * a.n_hits + hitNext];
    }
  }
  double chi2sum;
  double lc[3];  // perigee line point
  double fpar[6];
  double fcov[21];

  // quantities of the momentum used by every step, recomputed when p changes:
  // q/p, hypot(1, m/p) (dt/ds) and p hypot(1, m/p)
  double cacheP = -1.0, cacheQ = 0.0, cacheQop = 0.0, cacheDtds = 0.0, cacheDen = 0.0;
  KFD_INLINE void momentumCache() {
    if (p != cacheP || q != cacheQ) {
      cacheP = p;
      cacheQ = q;
      cacheQop = q / p;
      cacheDtds = hypot(1.0, a.mass / p);
      cacheDen = p * cacheDtds;
    }
  }
  // the last plane intersection computed by surfaceStatus (path length along
  // the direction, denominator n.d): navStatus and navTarget of one
  // iteration usually ask for the same plane; valid until pos, dir or navDir
  // change (moved())
  int pathPlane = -1;
  double pathLen = 0.0, pathDenom = 0.0;
  KFD_INLINE void moved() { pathPlane = -1; }

  // per-state stores of this wave's 64-track chunk (wave-uniform bases) and
  // this lane's slot in them
  double *tsd;
  S *tss;
  double *tsf;
  double *tsc;
  // the plane table: a copy in LDS (set by the kernel) or a.planes
  const DevPlane *pls;
  uint32_t lane;
  // slot of the forward -> backward hand-over record: the position within
  // the launch, so consecutive lanes use consecutive slots
  int64_t cpos;

  // REC = false: a build for fits that record neither track states nor
  // checkpoints (both stores are then null at compile time, and their
  // writes are not in the code)
  KFD_INLINE TrackFit(const FitArgs &args, int64_t idx, double *d_base, S *s_base, double *f_base,
                      double *c_base, uint32_t ln, int64_t carry_slot = 0)
      : a(args), i(idx), tsd(d_base), tss(s_base), tsf(REC ? f_base : nullptr),
        tsc(REC ? c_base : nullptr), pls(args.planes), lane(ln), cpos(carry_slot) {}

  // ------------------------------------------------------------ storage ----
  // Element f of state k: base[(k * F + f) * 64 + lane] -- a 32-bit lane
  // offset from a wave-uniform base, so each access is one saddr load/store.
  KFD_INLINE static uint32_t slot(int k, int F, int f, uint32_t ln) {
    return (uint32_t)(((uint32_t)k * (uint32_t)F + (uint32_t)f) * 64u + ln);
  }
  KFD_INLINE double *recF(int k, int f) const { return tsf + slot(k, KF_TS_STRIDE, f, lane); }
  // The smoother stores tsd / tss in 16-byte groups per lane: the group of
  // element f (f a multiple of W = 16 / sizeof(element)) of state k at
  // base[(k * F + f) * 64 + lane * W].
  static constexpr int kW = 16 / (int)sizeof(S);
  KFD_INLINE static uint32_t gslot(int k, int F, int f, int W, uint32_t ln) {
    return (uint32_t)(((uint32_t)k * (uint32_t)F + (uint32_t)f) * 64u + ln * (uint32_t)W);
  }
  KFD_INLINE double *grpD(int k, int f) const { return tsd + gslot(k, KF_SD_N, f, 2, lane); }
  KFD_INLINE S *grpS(int k, int f) const { return tss + gslot(k, KF_SS_N, f, kW, lane); }
  // The forward kernel's store writes are non-temporal: the backward kernel
  // reads them back only after the whole forward launch (gigabytes later),
  // so they need not displace the hits and plane data in the caches
  // (measured: forward 2-3 %, backward 3 % faster on 1,048,576 tracks).
  typedef float v4f __attribute__((ext_vector_type(4)));
  typedef double v2d __attribute__((ext_vector_type(2)));
  KFD_INLINE static void st16(float *p, const float *v) {
    const v4f x = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(x, (v4f *)p);
  }
  KFD_INLINE static void st16(double *p, const double *v) {
    const v2d x = {v[0], v[1]};
    __builtin_nontemporal_store(x, (v2d *)p);
  }
  // a store slot index kept in a padding element of the S store, bit for bit
  KFD_INLINE static float linkBits(int s, float) { return __builtin_bit_cast(float, s); }
  KFD_INLINE static double linkBits(int s, double) {
    return __builtin_bit_cast(double, (int64_t)s);
  }
  KFD_INLINE static int linkOf(float v) { return __builtin_bit_cast(int, v); }
  KFD_INLINE static int linkOf(double v) { return (int)__builtin_bit_cast(int64_t, v); }
  KFD_INLINE static void ld16(const float *p, float *v) {
    const float4 x = *(const float4 *)p;
    v[0] = x.x;
    v[1] = x.y;
    v[2] = x.z;
    v[3] = x.w;
  }
  KFD_INLINE static void ld16(const double *p, double *v) {
    const double2 x = *(const double2 *)p;
    v[0] = x.x;
    v[1] = x.y;
  }
  // six parameters at tsd element f (0 or 6) of state k
  KFD_INLINE void storePar(int k, int f, const double v[6]) const {
    if (!lead) return;
#pragma unroll
    for (int g = 0; g < 6; g += 2) st16(grpD(k, f + g), v + g);
  }
  KFD_INLINE void loadPar(int k, int f, double v[6]) const {
#pragma unroll
    for (int g = 0; g < 6; g += 2) ld16(grpD(k, f + g), v + g);
  }
  // n elements (a multiple of kW) at tss element f of state k
  template <int N>
  KFD_INLINE void storeS(int k, int f, const S v[N]) const {
    if (!lead) return;
#pragma unroll
    for (int g = 0; g < N; g += kW) st16(grpS(k, f + g), v + g);
  }
  template <int N>
  KFD_INLINE void loadS(int k, int f, S v[N]) const {
#pragma unroll
    for (int g = 0; g < N; g += kW) ld16(grpS(k, f + g), v + g);
  }
  KFD_INLINE double *recC(int k, int f) const { return tsc + slot(k, KF_CK_N, f, lane); }

  // Checkpoint k: the state right after the update and material of measured
  // surface k, from which the adjoint recomputes the next segment.
  KFD_INLINE void saveCheckpoint(int k) const {
    if (!lead) return;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      *recC(k, KF_CK_POS + m) = pos[m];
      *recC(k, KF_CK_DIR + m) = dir[m];
    }
    *recC(k, KF_CK_P) = p;
    *recC(k, KF_CK_T) = t;
    *recC(k, KF_CK_PATH) = pathAcc;
#pragma unroll
    for (int m = 0; m < 4; ++m) *recC(k, KF_CK_SS + m) = ss.v[m];
#pragma unroll
    for (int m = 0; m < 21; ++m) *recC(k, KF_CK_COV + m) = cov[m];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 5; ++c) *recC(k, KF_CK_J + 5 * r + c) = (double)A[r][c];
    *recC(k, KF_CK_J + 15) = (double)b;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) *recC(k, KF_CK_J + 16 + 3 * r + c) = (double)G[r][c];
    *recC(k, KF_CK_NEXT) = (double)next;
  }

  KFD_INLINE void loadCheckpoint(int k) {
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      pos[m] = *recC(k, KF_CK_POS + m);
      dir[m] = *recC(k, KF_CK_DIR + m);
    }
    p = *recC(k, KF_CK_P);
    t = *recC(k, KF_CK_T);
    pathAcc = *recC(k, KF_CK_PATH);
#pragma unroll
    for (int m = 0; m < 4; ++m) ss.v[m] = *recC(k, KF_CK_SS + m);
    ss.dir = 1;
    navDir = 1;
    moved();
#pragma unroll
    for (int m = 0; m < 21; ++m) cov[m] = *recC(k, KF_CK_COV + m);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 5; ++c) A[r][c] = (S)*recC(k, KF_CK_J + 5 * r + c);
    b = (S)*recC(k, KF_CK_J + 15);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) G[r][c] = (S)*recC(k, KF_CK_J + 16 + 3 * r + c);
    next = (int)*recC(k, KF_CK_NEXT);
    derivSet = false;
    k4d[0] = k4d[1] = k4d[2] = (S)0;
    dtds = (S)0;
  }
  // public record: parameters + covariance at offset off
  template <typename T>
  KFD_INLINE void storeBoundF(int k, int off, const double par[6], const T c[21]) const {
    if (!lead) return;
#pragma unroll
    for (int m = 0; m < 6; ++m) *recF(k, off + m) = par[m];
#pragma unroll
    for (int m = 0; m < 21; ++m) *recF(k, off + 6 + m) = (double)c[m];
  }
  // packed index of Jacobian entry (r, c), r in {0, 1, 2, 3, 5}, c < 5
  KFD_INLINE static constexpr int jidx(int r, int c) { return ((r < 4) ? r : 4) * 5 + c; }

  // -------------------------------------------------------------- field ----
  // FK: the field kind the kernel was built for (kFieldConst and
  // kFieldGrid leave the other kind's code out of the step loop)
  KFD_INLINE void field(const double x[3], double B[3]) const {
    if constexpr (FK == kFieldConst)
      fieldConst(a.field, B);
    else if constexpr (FK == kFieldGrid)
      fieldGrid(a.field, x[0], x[1], x[2], B);
    else
      fieldAt(a.field, x[0], x[1], x[2], B);
  }

  // ------------------------------------------------ Jacobian init (plane) ----
  // PlaneSurface::initJacobianToGlobal, PlaneSurface.ipp:122-156
  KFD_INLINE void initJacobianPlane(const DevPlane &pl, double cosT, double sinT, double cosP,
                                    double sinP) {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      A[r][0] = (S)pl.r[3 * r + 0];
      A[r][1] = (S)pl.r[3 * r + 1];
      A[r][2] = (S)0;
      A[r][3] = (S)0;
      A[r][4] = (S)0;
    }
    b = (S)0;
    G[0][0] = (S)((-sinT) * sinP);
    G[0][1] = (S)(cosT * cosP);
    G[1][0] = (S)(sinT * cosP);
    G[1][1] = (S)(cosT * sinP);
    G[2][0] = (S)0;
    G[2][1] = (S)(-sinT);
    G[0][2] = G[1][2] = G[2][2] = (S)0;
    derivSet = false;
    k4d[0] = k4d[1] = k4d[2] = (S)0;
    dtds = (S)0;
  }

  // ------------------------------------------------------------- start ----
  // BoundParameters<LineSurface>(cov, params, &perigee) (SingleBound-
  // TrackParameters.hpp:48-60), EigenStepper::State ctor (EigenStepper.hpp:
  // 41-61) and LineSurface::initJacobianToGlobal (LineSurface.ipp:63-112).
  KFD_INLINE void start(const double pv[6]) {
    const double pabs = fabs(1.0 / pv[4]);
    const SinCos s3 = sinCosD(pv[3]), s2 = sinCosD(pv[2]);
    const double mom[3] = {pabs * s3.s * s2.c, pabs * s3.s * s2.s, pabs * s3.c};
    // LineSurface::localToGlobal: centre + (0,0,loc1) + loc0 * |z x mom|
    const double ra[3] = {0.0 - mom[1], mom[0] - 0.0, 0.0};
    const double ran = sqrt(ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2]);
    const double raN[3] = {ra[0] / ran, ra[1] / ran, ra[2] / ran};
    const double locz[3] = {0.0 + lc[0], 0.0 + lc[1], pv[1] + lc[2]};
#pragma unroll
    for (int k = 0; k < 3; ++k) pos[k] = locz[k] + pv[0] * raN[k];
    const double charge = (pv[4] > 0.0) ? 1.0 : -1.0;
    double pc[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) pc[k] = pv[k];
    correctParams(pc);
    const double mn = sqrt(mom[0] * mom[0] + mom[1] * mom[1] + mom[2] * mom[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) dir[k] = mom[k] / mn;
    p = mn;
    q = charge;
    t = pc[5];
    navDir = 1;
    moved();
    ss.init(1.0 * fabs(a.max_step_size));
    pathAcc = 0.0;
    // line initJacobianToGlobal at (pos, dir) with the corrected parameters
    const double cosT = dir[2];
    const double sinT = sqrt(dir[0] * dir[0] + dir[1] * dir[1]);
    const double invSinT = 1.0 / sinT;
    const double cosP = dir[0] * invSinT, sinP = dir[1] * invSinT;
    const double mxr[3] = {0.0 * dir[2] - 1.0 * dir[1], 1.0 * dir[0] - 0.0 * dir[2],
                           0.0 * dir[1] - 0.0 * dir[0]};
    const double mxn = sqrt(mxr[0] * mxr[0] + mxr[1] * mxr[1] + mxr[2] * mxr[2]);
    const double mX[3] = {mxr[0] / mxn, mxr[1] / mxn, mxr[2] / mxn};
    const double mY[3] = {0.0, 0.0, 1.0};
    double mD[3];
    cross(mX, mY, mD);
    const double j42 = (-sinT) * sinP, j43 = cosT * cosP, j52 = sinT * cosP, j53 = cosT * sinP;
    const double j63 = (-sinT);
    const double ipdn = 1.0 / dot3(dir, mD);
    const double dPhi[3] = {j42, j52, 0.0};
    const double dTheta[3] = {j43, j53, j63};
    double dDPhiY[3], dDThetaY[3];
    cross(mY, dPhi, dDPhiY);
    cross(mY, dTheta, dDThetaY);
    const double xp = dot3(mX, dDPhiY), xt = dot3(mX, dDThetaY);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      dDPhiY[r] = dDPhiY[r] - xp * mX[r];
      dDThetaY[r] = dDThetaY[r] - xt * mX[r];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      A[r][0] = (S)mX[r];
      A[r][1] = (S)mY[r];
      A[r][2] = (S)(dDPhiY[r] * pc[0] * ipdn);
      A[r][3] = (S)(dDThetaY[r] * pc[0] * ipdn);
      A[r][4] = (S)0;
    }
    b = (S)0;
    G[0][0] = (S)j42;
    G[0][1] = (S)j43;
    G[1][0] = (S)j52;
    G[1][1] = (S)j53;
    G[2][0] = (S)0;
    G[2][1] = (S)j63;
    G[0][2] = G[1][2] = G[2][2] = (S)0;
    derivSet = false;
    k4d[0] = k4d[1] = k4d[2] = (S)0;
    dtds = (S)0;
    stepFailed = false;
  }

  // Library sin/cos, atan2 and log: out of line with OOL (the chi2 adjoint's
  // segment recompute, see oolSinCos), inline in the fit kernels (measured
  // faster there: the calls on every surface cost more than the registers
  // they free).
  KFD_INLINE static SinCos sinCosD(double x) {
    if constexpr (!OOL) {
      return SinCos{sin(x), cos(x)};
    } else {
      return oolSinCos(x);
    }
  }
  KFD_INLINE static double atan2D(double y, double x) {
    if constexpr (!OOL) {
      return atan2(y, x);
    } else {
      return oolAtan2(y, x);
    }
  }
  KFD_INLINE static S logS(S x) {
    if constexpr (!OOL || kMixed) {
      return log(x);
    } else {
      return oolLog(x);
    }
  }

  // The step-size factor after a rejected trial (EigenStepper.ipp:274-283):
  // pow(tolerance / |2 err|, 0.25) clamped to [0.25, 4].  The fourth root is
  // taken as two square roots (each correctly rounded: within an ulp of the
  // pow, the same NaN, 0 and infinity cases); the inlined library pow was
  // ~200 instructions in the step loop and its temporaries raised the loop's
  // register peak (an out-of-line call of it, or of this, measured within
  // 1 % in the FP64 kernels and slower in the mixed ones).
  KFD_INLINE static double stepScale(double tol, double err) {
    double sc = sqrt(sqrt(tol / fabs(2.0 * err)));
    sc = (0.25 < sc) ? sc : 0.25;
    return (4.0 < sc) ? 4.0 : sc;
  }

  // -------------------------------------------------------------- step ----
  // EigenStepper::step, EigenStepper.ipp:226-332, with the transport of
  // transportMatrix (:50-124) applied to J directly.
  KFD_INLINE bool step() {
    double B1[3], Bm[3], Bl[3];
    field(pos, B1);
    moved();
    momentumCache();
    const double qop = cacheQop;
    double k1[3], k2[3], k3[3], k4[3];
    cross(dir, B1, k1);
#pragma unroll
    for (int k = 0; k < 3; ++k) k1[k] = qop * k1[k];
    double err = 0.0;
    unsigned nTrials = 0;
    for (;;) {
      const double h = ss.get();
      const double h2 = h * h;
      const double half_h = h * 0.5;
      double p1[3], tv[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) p1[k] = pos[k] + half_h * dir[k] + (h2 * 0.125) * k1[k];
      field(p1, Bm);
#pragma unroll
      for (int k = 0; k < 3; ++k) tv[k] = dir[k] + half_h * k1[k];
      cross(tv, Bm, k2);
#pragma unroll
      for (int k = 0; k < 3; ++k) k2[k] = qop * k2[k];
#pragma unroll
      for (int k = 0; k < 3; ++k) tv[k] = dir[k] + half_h * k2[k];
      cross(tv, Bm, k3);
#pragma unroll
      for (int k = 0; k < 3; ++k) k3[k] = qop * k3[k];
#pragma unroll
      for (int k = 0; k < 3; ++k) p1[k] = pos[k] + h * dir[k] + (h2 * 0.5) * k3[k];
      field(p1, Bl);
#pragma unroll
      for (int k = 0; k < 3; ++k) tv[k] = dir[k] + h * k3[k];
      cross(tv, Bl, k4);
#pragma unroll
      for (int k = 0; k < 3; ++k) k4[k] = qop * k4[k];
      double e1 = 0.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double e = k1[k] - k2[k] - k3[k] + k4[k];
        e1 = (k == 0) ? fabs(e) : e1 + fabs(e);
      }
      err = h2 * e1;
      err = (err < 1e-20) ? 1e-20 : err;
      if (err <= a.tolerance) break;
      // std::min(std::max(0.25, x), 4.) with its NaN behaviour: a NaN error
      // estimate (h = DBL_MAX after an actor release) scales by 0.25
      ss.assign(ss.get() * stepScale(a.tolerance, err));
      const double hn = ss.get();
      if (hn * hn < a.step_size_cutoff * a.step_size_cutoff || nTrials > a.max_rk_trials) {
        stepFailed = true;
        return false;
      }
      nTrials++;
    }
    const double h = ss.get();
    // propagationTime, EigenStepper.ipp:36-47
    const double dt = cacheDtds;
    t = t + h * dt;
    transport(B1, Bm, Bl, k1, k2, k3, h, qop);
    // h^2/6 and h/6 as products with 1/6, one division for the
    // normalisation (within an ulp of the quotients)
    const double h6 = h * (1.0 / 6.0);
    const double hh6 = h * h6;
#pragma unroll
    for (int k = 0; k < 3; ++k) pos[k] = pos[k] + h * dir[k] + hh6 * (k1[k] + k2[k] + k3[k]);
#pragma unroll
    for (int k = 0; k < 3; ++k)
      dir[k] = dir[k] + h6 * (k1[k] + 2.0 * (k2[k] + k3[k]) + k4[k]);
    const double idn = 1.0 / sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) dir[k] = dir[k] * idn;
    derivSet = true;
    dtds = (S)dt;
#pragma unroll
    for (int k = 0; k < 3; ++k) k4d[k] = (S)k4[k];
    pathAcc = pathAcc + h;
    return true;
  }

  // J <- D J for the free transport matrix D of one RK step
  // (EigenStepper.ipp:50-124 builds D; its identity / zero blocks are used
  // implicitly).
  KFD_INLINE void transport(const double Bfd[3], const double Bmd[3], const double Bld[3],
                            const double k1d[3], const double k2d[3], const double k3d[3],
                            double hd, double qopd) {
    const S h = (S)hd, qop = (S)qopd, half_h = (S)(hd * 0.5);
    S d[3], Bf[3], Bm[3], Bl[3], k1[3], k2[3], k3[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      d[k] = (S)dir[k];
      Bf[k] = (S)Bfd[k];
      Bm[k] = (S)Bmd[k];
      Bl[k] = (S)Bld[k];
      k1[k] = (S)k1d[k];
      k2[k] = (S)k2d[k];
      k3[k] = (S)k3d[k];
    }
    S dk1dL[3], dk2dL[3], dk3dL[3], dk4dL[3], tv[3], c1[3], c2[3];
    cross(d, Bf, dk1dL);
#pragma unroll
    for (int k = 0; k < 3; ++k) tv[k] = d[k] + half_h * k1[k];
    cross(tv, Bm, c1);
    cross(dk1dL, Bm, c2);
#pragma unroll
    for (int k = 0; k < 3; ++k) dk2dL[k] = c1[k] + (qop * half_h) * c2[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) tv[k] = d[k] + half_h * k2[k];
    cross(tv, Bm, c1);
    cross(dk2dL, Bm, c2);
#pragma unroll
    for (int k = 0; k < 3; ++k) dk3dL[k] = c1[k] + (qop * half_h) * c2[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) tv[k] = d[k] + h * k3[k];
    cross(tv, Bl, c1);
    cross(dk3dL, Bl, c2);
#pragma unroll
    for (int k = 0; k < 3; ++k) dk4dL[k] = c1[k] + (qop * h) * c2[k];

    // divisions: a float 1/6 constant is biased by 3e-8 relative, which
    // accumulates over the steps of a track
    const S h6 = h / (S)6.0;
    const S hh6 = (h * h) / (S)6.0;
    S dFdT[3][3], dGdT[3][3], dFdL[3], dGdL[3];
    if constexpr (L == 1) {
      // dk/dT matrices (row r, column c)
      S dk1dT[3][3] = {{(S)0, Bf[2], -Bf[1]}, {-Bf[2], (S)0, Bf[0]}, {Bf[1], -Bf[0], (S)0}};
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) dk1dT[r][c] = qop * dk1dT[r][c];
      S dk2dT[3][3], dk3dT[3][3], dk4dT[3][3];
      crossStep(dk1dT, half_h, Bm, qop, dk2dT);
      crossStep(dk2dT, half_h, Bm, qop, dk3dT);
      crossStep(dk3dT, h, Bl, qop, dk4dT);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const S id = (r == c) ? (S)1 : (S)0;
          dFdT[r][c] = h * (id + h6 * (dk1dT[r][c] + dk2dT[r][c] + dk3dT[r][c]));
          dGdT[r][c] =
              id + h6 * (dk1dT[r][c] + (S)2.0 * (dk2dT[r][c] + dk3dT[r][c]) + dk4dT[r][c]);
        }
    } else {
      // lane groups: the columns of dF/dT and dG/dT are independent (column
      // c follows the chain dk1..dk4 of the unit vector e_c); each lane
      // builds the columns of its slots, c = rank + L * slot (the last
      // column again where that runs past it), and the group gathers them
      constexpr int kSlots = (3 + L - 1) / L;
      S colF[kSlots][3], colG[kSlots][3];
#pragma unroll
      for (int sl = 0; sl < kSlots; ++sl) {
        const uint32_t c = (rank + (uint32_t)(L * sl) < 3u) ? rank + (uint32_t)(L * sl) : 2u;
        const S e[3] = {(c == 0u) ? (S)1 : (S)0, (c == 1u) ? (S)1 : (S)0,
                        (c == 2u) ? (S)1 : (S)0};
        S k1c[3], k2c[3], k3c[3], k4c[3], col[3];
        // column c of dk1/dT = qop [B]x e_c: the unit-vector cross product
        // spelled out as the matrix of the single-lane path
        const S m1[3] = {e[1] * Bf[2] - e[2] * Bf[1], e[2] * Bf[0] - e[0] * Bf[2],
                         e[0] * Bf[1] - e[1] * Bf[0]};
#pragma unroll
        for (int k = 0; k < 3; ++k) k1c[k] = qop * m1[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) col[k] = e[k] + half_h * k1c[k];
        cross(col, Bm, k2c);
#pragma unroll
        for (int k = 0; k < 3; ++k) k2c[k] = qop * k2c[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) col[k] = e[k] + half_h * k2c[k];
        cross(col, Bm, k3c);
#pragma unroll
        for (int k = 0; k < 3; ++k) k3c[k] = qop * k3c[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) col[k] = e[k] + h * k3c[k];
        cross(col, Bl, k4c);
#pragma unroll
        for (int k = 0; k < 3; ++k) k4c[k] = qop * k4c[k];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          colF[sl][r] = h * (e[r] + h6 * (k1c[r] + k2c[r] + k3c[r]));
          colG[sl][r] = e[r] + h6 * (k1c[r] + (S)2.0 * (k2c[r] + k3c[r]) + k4c[r]);
        }
      }
      // gather: column c from rank c % L, slot c / L
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        S f[kSlots], g[kSlots], fo[3], go[3];
#pragma unroll
        for (int sl = 0; sl < kSlots; ++sl) {
          f[sl] = colF[sl][r];
          g[sl] = colG[sl][r];
        }
        groupGather<L, 3>(f, fo);
        groupGather<L, 3>(g, go);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          dFdT[r][c] = fo[c];
          dGdT[r][c] = go[c];
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      dFdL[r] = hh6 * (dk1dL[r] + dk2dL[r] + dk3dL[r]);
      dGdL[r] = h6 * (dk1dL[r] + (S)2.0 * (dk2dL[r] + dk3dL[r]) + dk4dL[r]);
    }
    const double m = a.mass;
    const S d37 = (S)(hd * m * m * q / cacheDen);

    // J <- D J on the non-constant entries
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      S g[3] = {G[0][c], G[1][c], G[2][c]};
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        S fa = dFdT[r][0] * g[0] + dFdT[r][1] * g[1] + dFdT[r][2] * g[2];
        S ga = dGdT[r][0] * g[0] + dGdT[r][1] * g[1] + dGdT[r][2] * g[2];
        if (c == 2) {
          fa = fa + dFdL[r];
          ga = ga + dGdL[r];
        }
        A[r][2 + c] = A[r][2 + c] + fa;
        G[r][c] = ga;
      }
    }
    b = b + d37;
  }

  // qop * cross_cols(I + hh * M, B)  (Helpers.hpp:168-176)
  KFD_INLINE static void crossStep(const S M[3][3], S hh, const S B[3], S qop, S out[3][3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      S col[3], r[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) col[k] = ((k == c) ? (S)1 : (S)0) + hh * M[k][c];
      cross(col, B, r);
#pragma unroll
      for (int k = 0; k < 3; ++k) out[k][c] = qop * r[k];
    }
  }

  // ------------------------------------------------- bound-state transport ----
  // detail::covarianceTransport (CovarianceEngine.hpp:269-290): derivative
  // correction J' = J - deriv sVec, bound Jacobian jb = jacToLocal J',
  // cov <- jb cov jb^T.  sVec is given; jacToLocal rows 0/1 are (u0, u1).
  KFD_INLINE void covarianceTransport(const double sVec[6], const double u0[3],
                                      const double u1[3], double cosT, double sinT,
                                      double invSinT, double cosP, double sinP, S jb[6][6]) {
    S dr[8];
#pragma unroll
    for (int k = 0; k < 3; ++k) dr[k] = derivSet ? (S)dir[k] : (S)0;
    dr[3] = derivSet ? dtds : (S)0;
#pragma unroll
    for (int k = 0; k < 3; ++k) dr[4 + k] = derivSet ? k4d[k] : (S)0;
    const S l20 = (S)(-sinP * invSinT), l21 = (S)(cosP * invSinT);
    const S l30 = (S)(cosP * cosT), l31 = (S)(sinP * cosT), l32 = (S)(-sinT);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const S sv = (S)sVec[c];
      S jp[8];
#pragma unroll
      for (int r = 0; r < 3; ++r) jp[r] = ((c < 5) ? A[r][c] : (S)0) - dr[r] * sv;
      jp[3] = ((c == 4) ? b : ((c == 5) ? (S)1 : (S)0)) - dr[3] * sv;
#pragma unroll
      for (int r = 0; r < 3; ++r) jp[4 + r] = ((c >= 2 && c < 5) ? G[r][c - 2] : (S)0) - dr[4 + r] * sv;
      jp[7] = (c == 4) ? (S)1 : (S)0;
      jb[0][c] = (S)u0[0] * jp[0] + (S)u0[1] * jp[1] + (S)u0[2] * jp[2];
      jb[1][c] = (S)u1[0] * jp[0] + (S)u1[1] * jp[1] + (S)u1[2] * jp[2];
      jb[2][c] = l20 * jp[4] + l21 * jp[5];
      jb[3][c] = l30 * jp[4] + l31 * jp[5] + l32 * jp[6];
      jb[4][c] = jp[7];
      jb[5][c] = jp[3];
    }
    S cs[21];
#pragma unroll
    for (int m = 0; m < 21; ++m) cs[m] = cov[m];
    sandwich(jb, cs, cov);
  }

  // cn = jb c jb^T in S.  Row 4 of jb is e4 and column 5 is e5 exactly (the
  // q/p row and the time column of the transport), so those terms are
  // skipped; row r of jb c is formed and consumed at once.  The smoother
  // calls this with the stored operands to recompute the predicted
  // covariance bit for bit.
  template <typename Out>
  KFD_INLINE static void sandwich(const S jb[6][6], const S c[21], Out cn[21]) {
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      S T[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        if (r == 4) {
          T[k] = c[sidx(4, k)];
        } else {
          S s = jb[r][0] * c[sidx(0, k)];
#pragma unroll
          for (int m = 1; m < 5; ++m) s = s + jb[r][m] * c[sidx(m, k)];
          if (r == 5) s = s + c[sidx(5, k)];
          T[k] = s;
        }
      }
#pragma unroll
      for (int cc = r; cc < 6; ++cc) {
        S s;
        if (cc == 4) {
          s = T[4];
        } else {
          s = T[0] * jb[cc][0];
#pragma unroll
          for (int m = 1; m < 5; ++m) s = s + T[m] * jb[cc][m];
          if (cc == 5) s = s + T[5];
        }
        cn[sidx(r, cc)] = (Out)s;
      }
    }
  }

  // boundState on a plane (EigenStepper.ipp:516-540, CovarianceEngine.hpp:
  // 345-371) including the Jacobian re-initialisation on the plane.
  KFD_INLINE void boundStatePlane(const DevPlane &pl, double par[6], S jb[6][6]) {
    const double cosT = dir[2];
    const double sinT = sqrt(dir[0] * dir[0] + dir[1] * dir[1]);
    const double invSinT = 1.0 / sinT;
    const double cosP = dir[0] * invSinT, sinP = dir[1] * invSinT;
    const double u0[3] = {pl.r[0], pl.r[3], pl.r[6]};
    const double u1[3] = {pl.r[1], pl.r[4], pl.r[7]};
    const double nr[3] = {pl.r[2], pl.r[5], pl.r[8]};
    // PlaneSurface::derivativeFactors, PlaneSurface.ipp:191-204
    const double dp = nr[0] * dir[0] + nr[1] * dir[1] + nr[2] * dir[2];
    const double idp = 1.0 / dp;
    const double nv[3] = {nr[0] * idp, nr[1] * idp, nr[2] * idp};
    double sVec[6];
#pragma unroll
    for (int c = 0; c < 5; ++c)
      sVec[c] = nv[0] * (double)A[0][c] + nv[1] * (double)A[1][c] + nv[2] * (double)A[2][c];
    sVec[5] = 0.0;
    covarianceTransport(sVec, u0, u1, cosT, sinT, invSinT, cosP, sinP, jb);
    // bound parameters: global2parameters on the plane
    const double qop = q / p;
    const double pabs = fabs(1.0 / qop);
    const double mom[3] = {pabs * dir[0], pabs * dir[1], pabs * dir[2]};
    const double charge = copysign(1.0, qop);
    const double dd[3] = {pos[0] - pl.c[0], pos[1] - pl.c[1], pos[2] - pl.c[2]};
    par[0] = pl.r[0] * dd[0] + pl.r[3] * dd[1] + pl.r[6] * dd[2];
    par[1] = pl.r[1] * dd[0] + pl.r[4] * dd[1] + pl.r[7] * dd[2];
    par[2] = atan2D(mom[1], mom[0]);
    par[3] = atan2D(sqrt(mom[0] * mom[0] + mom[1] * mom[1]), mom[2]);
    const double mn = sqrt(mom[0] * mom[0] + mom[1] * mom[1] + mom[2] * mom[2]);
    par[4] = ((fabs(charge) < 1e-4) ? 1.0 : charge) / mn;
    par[5] = t;
    correctParams(par);
    // reinitializeJacobians on the plane (PlaneSurface.ipp:122-156)
    initJacobianPlane(pl, cosT, sinT, cosP, sinP);
  }

  // boundState on the perigee line (LineSurface.ipp:114-220); the Jacobians
  // are not re-initialised because propagation ends here.
  KFD_INLINE void boundStateLine(double par[6]) {
    const double cosT = dir[2];
    const double sinT = sqrt(dir[0] * dir[0] + dir[1] * dir[1]);
    const double invSinT = 1.0 / sinT;
    const double cosP = dir[0] * invSinT, sinP = dir[1] * invSinT;
    // reference frame: measY = z, measX = |z x dir|, measDepth = measX x measY
    const double mxr[3] = {0.0 * dir[2] - 1.0 * dir[1], 1.0 * dir[0] - 0.0 * dir[2],
                           0.0 * dir[1] - 0.0 * dir[0]};
    const double mxn = sqrt(mxr[0] * mxr[0] + mxr[1] * mxr[1] + mxr[2] * mxr[2]);
    const double mX[3] = {mxr[0] / mxn, mxr[1] / mxn, mxr[2] / mxn};
    const double mY[3] = {0.0, 0.0, 1.0};
    // LineSurface::derivativeFactors, LineSurface.ipp:145-182
    const double pc[3] = {pos[0] - lc[0], pos[1] - lc[1], pos[2] - lc[2]};
    const double long_c = mY[0] * dir[0] + mY[1] * dir[1] + mY[2] * dir[2];
    const double nvec[3] = {dir[0] - long_c * mY[0], dir[1] - long_c * mY[1], dir[2] - long_c * mY[2]};
    const double nrm = 1.0 / (1.0 - long_c * long_c);
    double sVec[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double jc[3], jd[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        jc[r] = (c < 5) ? (double)A[r][c] : 0.0;
        jd[r] = (c >= 2 && c < 5) ? (double)G[r][c - 2] : 0.0;
      }
      const double s_vec = nvec[0] * jc[0] + nvec[1] * jc[1] + nvec[2] * jc[2];
      double diff[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) diff[r] = mY[r] * s_vec - jd[r];
      const double row = pc[0] * diff[0] + pc[1] * diff[1] + pc[2] * diff[2];
      sVec[c] = nrm * (s_vec - row);
    }
    S jb[6][6];
    covarianceTransport(sVec, mX, mY, cosT, sinT, invSinT, cosP, sinP, jb);
    const double qop = q / p;
    const double pabs = fabs(1.0 / qop);
    const double mom[3] = {pabs * dir[0], pabs * dir[1], pabs * dir[2]};
    const double charge = copysign(1.0, qop);
    // LineSurface::globalToLocal, LineSurface.ipp:200-220
    const double l3[3] = {pos[0] - lc[0], pos[1] - lc[1], pos[2] - lc[2]};
    double l0 = sqrt(l3[0] * l3[0] + l3[1] * l3[1]);
    const double cr[3] = {0.0 * mom[2] - 1.0 * mom[1], 1.0 * mom[0] - 0.0 * mom[2],
                          0.0 * mom[1] - 0.0 * mom[0]};
    const double sgn = (cr[0] * pc[0] + cr[1] * pc[1] + cr[2] * pc[2] < 0.0) ? -1.0 : 1.0;
    par[0] = l0 * sgn;
    par[1] = l3[2];
    par[2] = atan2D(mom[1], mom[0]);
    par[3] = atan2D(sqrt(mom[0] * mom[0] + mom[1] * mom[1]), mom[2]);
    const double mn = sqrt(mom[0] * mom[0] + mom[1] * mom[1] + mom[2] * mom[2]);
    par[4] = ((fabs(charge) < 1e-4) ? 1.0 : charge) / mn;
    par[5] = t;
    correctParams(par);
  }

  // EigenStepper::update from bound parameters on a plane
  // (boundParameters2freeParameters, coordinate_transformations.hpp:145-158,
  // then EigenStepper.ipp:554-564)
  KFD_INLINE void updateFromBound(const DevPlane &pl, const double par[6]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) pos[k] = (pl.r[3 * k] * par[0] + pl.r[3 * k + 1] * par[1]) + pl.c[k];
    const SinCos s3 = sinCosD(par[3]), s2 = sinCosD(par[2]);
    const double cT = s3.c, sT = s3.s;
    const double d0 = s2.c * sT, d1 = s2.s * sT, d2 = cT;
    const double idn = 1.0 / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    dir[0] = d0 * idn;
    dir[1] = d1 * idn;
    dir[2] = d2 * idn;
    moved();
    p = fabs(1.0 / par[4]);
    t = par[5];
  }

  // sin(d), cos(d) for |d| < 1/16 by their Taylor series to d^9 / d^10 (the
  // first omitted terms are below 2e-21)
  KFD_INLINE static void sinCosSmall(double d, double &s, double &c) {
    const double d2 = d * d;
    s = d * (1.0 - d2 * (1.0 / 6.0) *
                       (1.0 - d2 * (1.0 / 20.0) * (1.0 - d2 * (1.0 / 42.0) * (1.0 - d2 * (1.0 / 72.0)))));
    c = 1.0 - d2 * 0.5 *
                  (1.0 - d2 * (1.0 / 12.0) *
                             (1.0 - d2 * (1.0 / 30.0) *
                                        (1.0 - d2 * (1.0 / 56.0) * (1.0 - d2 * (1.0 / 90.0)))));
  }

  // updateFromBound after the gain-matrix update on a plane, while dir still
  // holds the predicted direction: cos and sin of the filtered angles follow
  // from those of the predicted ones (dir's components) by the angle-addition
  // theorems with the update d = filtered - predicted, in place of four
  // library calls; an update of 1/16 rad or more, or a direction within 1e-3
  // of the z axis, takes updateFromBound.  Agrees with it to a few ulp.
  KFD_INLINE void updateFromFiltered(const DevPlane &pl, const double xf[6], const double xp[6]) {
    double dphi = xf[2] - xp[2];  // modulo 2 pi (xf[2] is wrapped)
    if (dphi > kPi)
      dphi -= 2.0 * kPi;
    else if (dphi < -kPi)
      dphi += 2.0 * kPi;
    const double dth = xf[3] - xp[3];
    const double sinT = sqrt(dir[0] * dir[0] + dir[1] * dir[1]), cosT = dir[2];
    if (!(fabs(dphi) < 0.0625 && fabs(dth) < 0.0625 && sinT > 1e-3)) {
      updateFromBound(pl, xf);
      return;
    }
    const double iS = 1.0 / sinT;
    const double cosP = dir[0] * iS, sinP = dir[1] * iS;
    double sp, cp, st, ct;
    sinCosSmall(dphi, sp, cp);
    sinCosSmall(dth, st, ct);
    const double cPf = cosP * cp - sinP * sp, sPf = sinP * cp + cosP * sp;
    const double cTf = cosT * ct - sinT * st, sTf = sinT * ct + cosT * st;
#pragma unroll
    for (int k = 0; k < 3; ++k) pos[k] = (pl.r[3 * k] * xf[0] + pl.r[3 * k + 1] * xf[1]) + pl.c[k];
    const double d0 = cPf * sTf, d1 = sPf * sTf, d2 = cTf;
    const double idn = 1.0 / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    dir[0] = d0 * idn;
    dir[1] = d1 * idn;
    dir[2] = d2 * idn;
    moved();
    p = fabs(1.0 / xf[4]);
    t = xf[5];
  }

  // ----------------------------------------------------------- navigation ----
  KFD_INLINE const DevPlane &plane(int idx) const { return pls[idx]; }

  // updateSingleSurfaceStatus on a plane, SteppingHelper.hpp:31-65 (+
  // PlanarHelper.hpp:28-49); returns 0 unreachable, 1 reachable, 2 on surface
  KFD_INLINE int surfaceStatus(int idx) {
    const double nd = (double)navDir;
    double denom, path;
    if (idx == pathPlane) {
      denom = pathDenom;
      path = pathLen;
    } else {
      const DevPlane &pl = plane(idx);
      const double d[3] = {nd * dir[0], nd * dir[1], nd * dir[2]};
      const double nr[3] = {pl.r[2], pl.r[5], pl.r[8]};
      denom = dot3(d, nr);
      const double cm[3] = {pl.c[0] - pos[0], pl.c[1] - pos[1], pl.c[2] - pos[2]};
      path = (denom != 0.0) ? dot3(nr, cm) / denom : 0.0;
      pathPlane = idx;
      pathDenom = denom;
      pathLen = path;
    }
    if (denom != 0.0) {
      if (path * path < kOnSurfaceTolerance * kOnSurfaceTolerance) {
        ss.release(kActor);
        return 2;
      }
      const double pLimit = ss.v[kAborter];
      if (path > -a.overstep_limit && path * path < pLimit * pLimit) {
        ss.release(kActor);
        ss.update(nd * path, kActor);
        return 1;
      }
    }
    return 0;
  }

  // DirectNavigator::status, DirectNavigator.hpp:158-188
  KFD_INLINE void navStatus() {
    cur = -1;
    if (next != size) {
      const int sp = seqFirst + next;
      if (surfaceStatus(sp) == 2) {
        cur = sp;
        ++next;
        if (next != size) ss.release(kActor);
      }
    }
  }

  // DirectNavigator::target, DirectNavigator.hpp:197-225
  KFD_INLINE void navTarget() {
    cur = -1;
    if (next != size) {
      if (surfaceStatus(seqFirst + next) == 0) ++next;
    } else {
      navBreak = true;  // the perigee target is always set
    }
  }

  // PathLimitReached, StandardAborters.hpp:24-57
  KFD_INLINE bool pathLimit() {
    if (targetReached) return true;
    const double distance = (double)navDir * kDblMax - pathAcc;
    ss.update(distance, kAborter);
    const bool reached = distance * distance < a.target_tolerance * a.target_tolerance;
    if (reached) targetReached = true;
    return reached;
  }

  // SurfaceReached on the perigee line, StandardAborters.hpp:86-132
  // (closest approach, LineSurface.ipp:229-262)
  KFD_INLINE bool surfaceReached() {
    if (targetReached) return true;
    if (cur == -2) {
      targetReached = true;
      return true;
    }
    const double nd = (double)navDir;
    const double d[3] = {nd * dir[0], nd * dir[1], nd * dir[2]};
    const double mab[3] = {lc[0] - pos[0], lc[1] - pos[1], lc[2] - pos[2]};
    const double eb[3] = {0.0, 0.0, 1.0};
    const double eaTeb = dot3(d, eb);
    const double denom = 1.0 - eaTeb * eaTeb;
    double distance = kDblMax;
    bool onSurface = false;
    if (denom * denom > kOnSurfaceTolerance * kOnSurfaceTolerance) {
      const double u = (dot3(mab, d) - dot3(mab, eb) * eaTeb) / denom;
      onSurface = u * u < kOnSurfaceTolerance * kOnSurfaceTolerance;
      distance = u;
    }
    if (onSurface) {
      cur = -2;
      targetReached = true;
    } else {
      ss.update(nd * distance, kAborter);
    }
    return onSurface;
  }

  KFD_INLINE bool kalmanAbort() const { return !result || finished; }
  KFD_INLINE bool finalizeCond() const {
    return (measured == (int)(h1 - h0) || (measured > 0 && navBreak)) && !smoothed && !finished;
  }

  // ------------------------------------------------------------- material ----
  // Actor::materialInteractor (KalmanFitter.hpp:405-449) with
  // PointwiseMaterialInteraction (.hpp:66-148, .ipp:14-53) and the RPP
  // formulas of Interactions.ipp (Bethe :150-176, Landau sigma :316-342,
  // Highland :501-546).  Only plane surfaces carry material (full update).
  // The formulas run in S: in the mixed mode the energy loss (~1e-4 of p)
  // and the noise variances need no more than float's 1e-7 relative; the
  // momentum itself is updated in double.
  KFD_INLINE void material(const DevPlane &pl) {
    if (!pl.has_material) return;
    const double momentum = p;
    const S qq = (S)q;
    const S qOverP = (S)(q / momentum);
    const S mass = (S)a.mass;
    const S nr[3] = {(S)pl.r[2], (S)pl.r[5], (S)pl.r[8]};
    const S d[3] = {(S)dir[0], (S)dir[1], (S)dir[2]};
    const S pcor = (S)1 / fabs(dot3(nr, d));
    const S thick = (S)pl.thickness * pcor;
    const S xOverX0 = (S)pl.thicknessInX0 * pcor;
    if (!((S)0 < thick)) return;
    // RelativisticQuantities, Interactions.ipp:27-52
    const S q2OverBeta2 = qq * qq + (mass * qOverP) * (mass * qOverP);
    const S mOverP = mass * fabs(qOverP / qq);
    const S pOverM = (S)1 / mOverP;
    const S beta2 = (S)1 / ((S)1 + mOverP * mOverP);
    const S betaGamma = pOverM;
    const S gamma = sqrt((S)1 + pOverM * pOverM);
    const S Ne = (S)pl.molarElectronDensity;
    const S eps = (S)0.5 * (S)a.kK * Ne * thick * q2OverBeta2;
    S Eloss = (S)0, vPhi = (S)0, vTheta = (S)0, vQop = (S)0;
    if (a.eloss) {
      const S I = (S)pl.meanExcitation;
      const S kMe = (S)a.kMe;
      S dhalf = (S)0;
      if (!(betaGamma < (S)10)) {
        const S plasma = (S)a.kPlasma * sqrt(Ne);
        dhalf = logS(betaGamma) + logS(plasma / I) - (S)0.5;
      }
      const S u = (S)2 * kMe * betaGamma * betaGamma;
      const S mfrac = kMe / mass;
      const S nom = (S)2 * kMe * betaGamma * betaGamma;
      const S den = (S)1 + (S)2 * gamma * mfrac + mfrac * mfrac;
      const S wmax = nom / den;
      const S running = (S)0.5 * logS(u / I) + (S)0.5 * logS(wmax / I) - beta2 - dhalf;
      Eloss = eps * running;
    }
    if (a.ms) {
      const S tt = sqrt(xOverX0 * q2OverBeta2);
      const S theta0 =
          (S)a.kHighland * fabs(qOverP / qq) * tt * ((S)1 + (S)a.kHighlandLog * logS(tt));
      const S dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const S sigmaPhi = theta0 * (dn / sqrt(d[0] * d[0] + d[1] * d[1]));
      vPhi = sigmaPhi * sigmaPhi;
      vTheta = theta0 * theta0;
    }
    if (a.eloss) {
      const S fwhm = (S)4 * eps;
      const S sigmaE = fwhm / (S)a.kLandauConv;
      const S pInv = qOverP / qq;
      const S sq = sqrt(q2OverBeta2) * pInv * pInv * sigmaE;
      vQop = sq * sq;
    }
    const double nextE =
        sqrt(a.mass * a.mass + momentum * momentum) - copysign((double)Eloss, (double)navDir);
    p = (a.mass < nextE) ? sqrt(nextE * nextE - a.mass * a.mass) : 0.0;
    const double mode = (navDir == 1) ? 1.0 : -1.0;
    double v;  // std::max(0, v) of updateVariance
    v = (double)cov[sidx(2, 2)] + copysign((double)vPhi, mode);
    cov[sidx(2, 2)] = (S)((0.0 < v) ? v : 0.0);
    v = (double)cov[sidx(3, 3)] + copysign((double)vTheta, mode);
    cov[sidx(3, 3)] = (S)((0.0 < v) ? v : 0.0);
    v = (double)cov[sidx(4, 4)] + copysign((double)vQop, mode);
    cov[sidx(4, 4)] = (S)((0.0 < v) ? v : 0.0);
  }

  // -------------------------------------------------------------- filter ----
  // Actor::filter (KalmanFitter.hpp:284-350) with GainMatrixUpdater
  // (GainMatrixUpdater.hpp:34-106).  FP64 mode: the reference's (I - K H) P
  // update.  Mixed mode: the Joseph form (I - K H) P (I - K H)^T + K V K^T,
  // which stays positive definite under the float-rounded predicted
  // covariance; each (i, j) pair is formed once, so the result is symmetric
  // by construction.  The 2x2 inversion and the update are double in both.
  KFD_INLINE bool filter(int sp) {
    const DevPlane &pl = plane(sp);
    // the track's hit on this plane: first match in hit order (SurfaceFinder,
    // KalmanFitter.hpp:31-38); the loaded-ahead next hit, else a scan with
    // four independent geo-id loads per round trip
    int64_t hit = -1;
    double m0, m1, V00, V01, V11;
    if (inOrder && pfValid && pfGeo == pl.geo_id) {
      hit = hitNext;
      m0 = pfM0;
      m1 = pfM1;
      V00 = pfV00;
      V01 = pfV01;
      V11 = pfV11;
    } else {
      for (int64_t h = inOrder ? hitNext : h0; h < h1 && hit < 0; h += 4) {
        int32_t g[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) g[q] = (h + q < h1) ? a.hit_geo_id[h + q] : 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (hit < 0 && h + q < h1 && g[q] == pl.geo_id) hit = h + q;
      }
      if (hit < 0) {  // hole
        if constexpr (PS) ++holes;
        return true;
      }
      m0 = a.hit_meas[hit];
      m1 = a.hit_meas[a.n_hits + hit];
      V00 = a.hit_cov[hit];
      V01 = a.hit_cov[a.n_hits + hit];
      V11 = a.hit_cov[2 * a.n_hits + hit];
      inOrder = inOrder && hit == hitNext;
    }
    if (inOrder) {
      hitNext = hit + 1;
      prefetchHit();
    }
    if (measured >= a.max_states) return false;
    const int k = measured;
    int s = k;  // store slot
    if constexpr (PS) {
      const int slack = a.max_states - (int)(h1 - h0);
      s = k + ((holes < slack) ? holes : (slack > 0 ? slack : 0));
    }
    if (k == 0) firstPlane = sp;
    double xp[6];
    S jb[6][6];
    boundStatePlane(pl, xp, jb);
    // predicted parameters and the transport Jacobian to the store, with the
    // slot of the state two measurements back (the smoother's look-ahead)
    storePar(s, KF_SD_XP, xp);
    {
      S jv[KF_SS_CF - KF_SS_J];
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 5; ++c)
          if (r != 4) jv[jidx(r, c)] = jb[r][c];
      jv[KF_SS_LINK] = PS ? linkBits(slotPrev, (S)0) : (S)0;
#pragma unroll
      for (int m = KF_SS_LINK + 1; m < KF_SS_CF - KF_SS_J; ++m) jv[m] = (S)0;
      storeS<KF_SS_CF - KF_SS_J>(s, KF_SS_J, jv);
    }
    if (tsf && lead) {
      storeBoundF(k, KF_TS_PREDICTED, xp, cov);
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) *recF(k, KF_TS_JACOBIAN + 6 * r + c) = (double)jb[r][c];
      *recF(k, KF_TS_PATH) = pathAcc;
      *recF(k, KF_TS_GEO_ID) = (double)pl.geo_id;
    }
    // gain-matrix update
    const double S00 = cov[sidx(0, 0)] + V00, S01 = cov[sidx(0, 1)] + V01;
    const double S10 = cov[sidx(1, 0)] + V01, S11 = cov[sidx(1, 1)] + V11;
    const double det = S00 * S11 - S01 * S10;
    const double idet = 1.0 / det;
    const double i00 = S11 * idet, i01 = -S01 * idet, i10 = -S10 * idet, i11 = S00 * idet;
    const double r0 = m0 - xp[0], r1 = m1 - xp[1];
    const double chi2 = r0 * (i00 * r0 + i01 * r1) + r1 * (i10 * r0 + i11 * r1);
    double xf[6];
    // outlier finder: above the cut the state keeps its prediction (no
    // update, the covariance stays the predicted one) and stays out of the
    // chi2 sum and ndf; 0 disables it (the reference's void finder)
    if (a.outlier_chi2_cut > 0.0 && chi2 > a.outlier_chi2_cut) {
#pragma unroll
      for (int r = 0; r < 6; ++r) xf[r] = xp[r];
      ++outliers;
    } else {
    double K0[6], K1[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      K0[r] = cov[sidx(r, 0)] * i00 + cov[sidx(r, 1)] * i10;
      K1[r] = cov[sidx(r, 0)] * i01 + cov[sidx(r, 1)] * i11;
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) xf[r] = xp[r] + (K0[r] * r0 + K1[r] * r1);
    // the update in place on cov: rows 0 and 1 of the result go to t01 first
    // (every entry reads the old rows 0 and 1), then rows 2..5 overwrite
    // cov, then rows 0 and 1 are written back
    double t01[11];
    if constexpr (kMixed) {
      // B = (I - K H) P: columns 0 and 1 here, the other (r <= c) entries
      // where they are consumed
      double B0[6], B1[6];
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        B0[r] = cov[sidx(r, 0)] - K0[r] * cov[sidx(0, 0)] - K1[r] * cov[sidx(1, 0)];
        B1[r] = cov[sidx(r, 1)] - K0[r] * cov[sidx(0, 1)] - K1[r] * cov[sidx(1, 1)];
      }
      auto joseph = [&](int r, int c) {
        const double brc = (c == 0) ? B0[r] : (c == 1) ? B1[r]
                                    : cov[sidx(r, c)] - K0[r] * cov[sidx(0, c)] - K1[r] * cov[sidx(1, c)];
        const double kv0 = V00 * K0[c] + V01 * K1[c];
        const double kv1 = V01 * K0[c] + V11 * K1[c];
        return (brc - B0[r] * K0[c] - B1[r] * K1[c]) + (K0[r] * kv0 + K1[r] * kv1);
      };
#pragma unroll
      for (int c = 0; c < 6; ++c) t01[c] = joseph(0, c);
#pragma unroll
      for (int c = 1; c < 6; ++c) t01[5 + c] = joseph(1, c);
#pragma unroll
      for (int r = 2; r < 6; ++r)
#pragma unroll
        for (int c = r; c < 6; ++c) cov[sidx(r, c)] = (S)joseph(r, c);
    } else {
      auto standard = [&](int r, int c) {
        const double ca = ((r == 0) ? 1.0 : 0.0) - K0[r];
        const double cb = ((r == 1) ? 1.0 : 0.0) - K1[r];
        double s = ca * cov[sidx(0, c)] + cb * cov[sidx(1, c)];
        if (r >= 2) s = s + cov[sidx(r, c)];
        return s;
      };
#pragma unroll
      for (int c = 0; c < 6; ++c) t01[c] = standard(0, c);
#pragma unroll
      for (int c = 1; c < 6; ++c) t01[5 + c] = standard(1, c);
#pragma unroll
      for (int r = 2; r < 6; ++r)
#pragma unroll
        for (int c = r; c < 6; ++c) cov[sidx(r, c)] = (S)standard(r, c);
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) cov[sidx(0, c)] = (S)t01[c];
#pragma unroll
    for (int c = 1; c < 6; ++c) cov[sidx(1, c)] = (S)t01[5 + c];
    chi2sum = chi2sum + chi2;
    }
    correctParams(xf);
    storePar(s, KF_SD_XF, xf);
    if (tsf && lead) {
      storeBoundF(k, KF_TS_FILTERED, xf, cov);
      *recF(k, KF_TS_CHI2) = chi2;
    }
    // the filtered covariance and, after the material update, the three
    // diagonal entries in which the covariance the next transport starts
    // from differs from it
    S cm[KF_SS_N - KF_SS_CF];
#pragma unroll
    for (int m = 0; m < 21; ++m) cm[m] = cov[m];
    updateFromFiltered(pl, xf, xp);
    material(pl);
    cm[KF_SS_MD - KF_SS_CF + 0] = cov[sidx(2, 2)];
    cm[KF_SS_MD - KF_SS_CF + 1] = cov[sidx(3, 3)];
    cm[KF_SS_MD - KF_SS_CF + 2] = cov[sidx(4, 4)];
    storeS<KF_SS_N - KF_SS_CF>(s, KF_SS_CF, cm);
    if (tsc) saveCheckpoint(k);
    if constexpr (PS) {
      slotPrev = slotLast;
      slotLast = s;
    }
    ++measured;
    return true;
  }

  // -------------------------------------------------------------- smoother ----
  // The diagonal scale of spdSolve, ~1 / sqrt(c).  Any positive scale gives
  // the same solution in exact arithmetic (the factorisation works on the
  // scaled diagonal as computed, not on an assumed 1), so the GPU takes the
  // hardware's reciprocal square root without refinement (one instruction
  // instead of a correctly rounded square root and division); NaN for c < 0
  // and infinity for c = 0 as 1 / sqrt(c).
  KFD_INLINE static double scaleOf(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsq(c);
#else
    return 1.0 / sqrt(c);  // host build of the fit program
#endif
  }

  // X = C^-1 M for a symmetric positive-definite 6x6 C (packed) and six
  // right-hand sides, in place on M: LDL^T factorisation of the
  // diagonally scaled matrix S C S (S = diag(C)^-1/2, unit diagonal, so no
  // pivoting is needed), then forward substitution, scaling by D^-1 and back
  // substitution per column.  The reference inverts C explicitly (Eigen's
  // partial-pivoting LU on the host, cofactors on the device,
  // include/Utilities/Math.hpp:74-112); solving is ~3x cheaper and agrees to
  // rounding.  A non-positive diagonal entry gives NaNs (caught by the
  // smoother's NaN check, as the reference's).
  KFD_INLINE static void spdSolve(const double C[21], double M[6][6]) {
    double sc[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) sc[k] = scaleOf(C[sidx(k, k)]);
    // W[r][k] = Lo[r][k] D[k] (r > k), Lo below the diagonal, Dinv = 1 / D
    double Lo[6][6], W[6][6], Dinv[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double dj = (sc[j] * C[sidx(j, j)]) * sc[j];
#pragma unroll
      for (int k = 0; k < j; ++k) dj = dj - W[j][k] * Lo[j][k];
      Dinv[j] = 1.0 / dj;
#pragma unroll
      for (int r = j + 1; r < 6; ++r) {
        double w = (sc[r] * C[sidx(r, j)]) * sc[j];
#pragma unroll
        for (int k = 0; k < j; ++k) w = w - W[r][k] * Lo[j][k];
        W[r][j] = w;
        Lo[r][j] = w * Dinv[j];
      }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double y[6];
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        double v = sc[r] * M[r][c];
#pragma unroll
        for (int k = 0; k < r; ++k) v = v - Lo[r][k] * y[k];
        y[r] = v;
      }
#pragma unroll
      for (int r = 0; r < 6; ++r) y[r] = y[r] * Dinv[r];
#pragma unroll
      for (int r = 5; r >= 0; --r) {
        double v = y[r];
#pragma unroll
        for (int k = r + 1; k < 6; ++k) v = v - Lo[k][r] * y[k];
        y[r] = v;
      }
#pragma unroll
      for (int r = 0; r < 6; ++r) M[r][c] = sc[r] * y[r];
    }
  }

  // Filtered covariance of state k, in double.
  KFD_INLINE void loadFilteredCov(int k, double c[21]) const {
    S cm[KF_SS_N - KF_SS_CF];
    loadS<KF_SS_N - KF_SS_CF>(k, KF_SS_CF, cm);
#pragma unroll
    for (int m = 0; m < 21; ++m) c[m] = (double)cm[m];
  }

  // Stored operands of smoother step k: the Jacobian, predicted parameters
  // and link of state k + 1 (slot sn), the filtered covariance,
  // post-material diagonal and filtered parameters of state k (slot sk).
  struct SmoothIn {
    S j[25];
    S cf[21];
    S md[3];
    double xf[6];
    double xpn[6];
    int link;  // slot of state k - 1 (-1: none)
  };
  KFD_INLINE void loadSmoothIn(int sk, int sn, SmoothIn &in) const {
    S jv[KF_SS_CF - KF_SS_J], cm[KF_SS_N - KF_SS_CF];
    loadS<KF_SS_CF - KF_SS_J>(sn, KF_SS_J, jv);
    loadS<KF_SS_N - KF_SS_CF>(sk, KF_SS_CF, cm);
#pragma unroll
    for (int m = 0; m < 25; ++m) in.j[m] = jv[m];
    in.link = linkOf(jv[KF_SS_LINK]);
#pragma unroll
    for (int m = 0; m < 21; ++m) in.cf[m] = cm[m];
#pragma unroll
    for (int m = 0; m < 3; ++m) in.md[m] = cm[KF_SS_MD - KF_SS_CF + m];
    loadPar(sk, KF_SD_XF, in.xf);
    loadPar(sn, KF_SD_XP, in.xpn);
  }

  // One smoother step (GainMatrixSmoother.hpp:21-94): state k from its
  // stored operands `in` and the smoothed state k + 1 (xsn, csn), which it
  // overwrites with the smoothed state k.  False on a NaN gain.
  // G = (C_f J^T) C_p^-1, x_s = x_f + G (x_s' - x_p'), C_s = C_f - G (C_p - C_s') G^T,
  // in double in both modes.
  KFD_INLINE bool smoothStep(int k, const SmoothIn &in, double xsn[6], double csn[21]) {
    S J[6][6];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c)
        J[r][c] = (r == 4 || c == 5) ? ((r == c) ? (S)1 : (S)0) : in.j[jidx(r, c)];
    // predicted covariance of state k + 1: the forward pass's transport of
    // the post-material covariance of state k, from the same operands
    double Cp[21];
    {
      S cm[21];
#pragma unroll
      for (int m = 0; m < 21; ++m) cm[m] = in.cf[m];
      cm[sidx(2, 2)] = in.md[0];
      cm[sidx(3, 3)] = in.md[1];
      cm[sidx(4, 4)] = in.md[2];
      sandwich(J, cm, Cp);
    }
    double cf[21];
#pragma unroll
    for (int m = 0; m < 21; ++m) cf[m] = (double)in.cf[m];
    // G^T = C_p^-1 (J C_f) (C_f, C_p symmetric); row 4 of J is e4 and
    // column 5 is e5
    double G[6][6];
    {
      double X[6][6];
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          double s;
          if (r == 4) {
            s = cf[sidx(4, c)];
          } else {
            s = (double)J[r][0] * cf[sidx(0, c)];
#pragma unroll
            for (int m = 1; m < 5; ++m) s = s + (double)J[r][m] * cf[sidx(m, c)];
            if (r == 5) s = s + cf[sidx(5, c)];
          }
          X[r][c] = s;
        }
      spdSolve(Cp, X);
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 6; ++c) G[r][c] = X[c][r];
    }
    bool nan = false;  // GainMatrixSmoother.hpp:63-66
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c) nan = nan || (G[r][c] != G[r][c]);
    if (nan) {
      smootherNaN = true;
      return false;
    }
    double dx[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) dx[r] = xsn[r] - in.xpn[r];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      double s = G[r][0] * dx[0];
#pragma unroll
      for (int m = 1; m < 6; ++m) s = s + G[r][m] * dx[m];
      xsn[r] = in.xf[r] + s;
    }
    double Dm[21];
#pragma unroll
    for (int m = 0; m < 21; ++m) Dm[m] = Cp[m] - csn[m];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      double B1[6];  // row r of G (C_p - C_s')
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        double s = G[r][0] * Dm[sidx(0, c)];
#pragma unroll
        for (int m = 1; m < 6; ++m) s = s + G[r][m] * Dm[sidx(m, c)];
        B1[c] = s;
      }
#pragma unroll
      for (int c = r; c < 6; ++c) {
        double s = B1[0] * G[c][0];
#pragma unroll
        for (int m = 1; m < 6; ++m) s = s + B1[m] * G[c][m];
        csn[sidx(r, c)] = cf[sidx(r, c)] - s;
      }
    }
    correctParams(xsn);
    if (tsf) storeBoundF(k, KF_TS_SMOOTHED, xsn, csn);
    return true;
  }

  // The smoother over the stored states.  The operands of step k - 1 are
  // loaded while step k computes.  PS = false: the loop runs over k (state k
  // in slot k).  PS = true: it walks the store slots downwards from the last
  // one the sequence can use and smooths a state when it reaches its slot,
  // so the lanes whose tracks cross the same planes read the same store
  // chunk in the same iteration (a lane idles at the slots of its holes);
  // the walk ends after at most max_states iterations.
  KFD_INLINE bool smooth(double xs[6], double cs[21]) {
    const int n = measured;
    double xsn[6], csn[21];
    // last state: smoothed = filtered
    const int sl = PS ? slotLast : n - 1;  // its slot
    loadPar(sl, KF_SD_XF, xsn);
    loadFilteredCov(sl, csn);
    if (tsf) storeBoundF(n - 1, KF_TS_SMOOTHED, xsn, csn);
    SmoothIn cur;
    if constexpr (!PS) {
      if (n >= 2) loadSmoothIn(n - 2, n - 1, cur);
      for (int k = n - 2; k >= 0; --k) {
        SmoothIn nxt;
        loadSmoothIn((k > 0) ? k - 1 : 0, k, nxt);
        if (!smoothStep(k, cur, xsn, csn)) return false;
        cur = nxt;
      }
    } else {
      int sk = slotPrev;  // slot of state k
      if (n >= 2) loadSmoothIn(sk, sl, cur);
      int slot = ((size < a.max_states) ? size : a.max_states) - 1;
      int k = n - 2;
      for (; k >= 0 && slot >= 0; --slot) {
        if (slot != sk) continue;  // a hole of this track
        const int link = cur.link;
        SmoothIn nxt;
        loadSmoothIn((k > 0 && link >= 0 && link < sk) ? link : sk, sk, nxt);
        if (!smoothStep(k, cur, xsn, csn)) return false;
        cur = nxt;
        sk = link;
        --k;
      }
      if (k >= 0) {  // slots out of order: the store is inconsistent
        smootherNaN = true;
        return false;
      }
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) xs[m] = xsn[m];
#pragma unroll
    for (int m = 0; m < 21; ++m) cs[m] = csn[m];
    return true;
  }

  // Actor::operator(), KalmanFitter.hpp:216-273.  Forward kernel: filter and
  // the finalisation check; returns true where finalize() must run (the
  // backward kernel does it).  Backward kernel: only the post-finalisation
  // part can act (smoothed is set).
  KFD_INLINE bool actor() {
    if constexpr (FWD) {
      if (cur >= 0 && !smoothed && !finished) {
        if (!filter(cur)) result = false;
      }
      if (finalizeCond()) {
        if (!a.smoothing) {
          finished = true;
        } else {
          return true;
        }
      }
      return false;
    } else {
      actorPost();
      return false;
    }
  }

  KFD_INLINE void actorPost() {
    if (smoothed && !finished && surfaceReached()) {
      boundStateLine(fpar);
#pragma unroll
      for (int m = 0; m < 21; ++m) fcov[m] = cov[m];
      finished = true;
    }
  }

  // The propagation loop (Propagator.ipp:39-69) with the cheap iterations in
  // an inner loop.  Returns true if the forward kernel must hand over to the
  // backward kernel (finalize due inside the actor).
  KFD_INLINE bool loop(int &steps, bool &terminated) {
    while (steps < a.max_steps) {
      bool event = false;
      while (steps < a.max_steps) {
        step();
        navStatus();
        if (FWD && cur >= 0 && !smoothed && !finished) {
          event = true;
          break;
        }
        if (FWD && finalizeCond()) {
          event = true;
          break;
        }
        if (!FWD && smoothed && !finished && surfaceReached()) {
          event = true;
          break;
        }
        if (kalmanAbort() || pathLimit()) {
          terminated = true;
          break;
        }
        navTarget();
        ++steps;
      }
      if (!event) break;
      if (actor()) return true;
      if (kalmanAbort() || pathLimit()) {
        terminated = true;
        break;
      }
      navTarget();
      ++steps;
    }
    return false;
  }

  // ------------------------------------------------------------- carry ----
  // Per-track hand-over record from the forward to the backward kernel.
  enum {
    kCI_resume = 0, kCI_steps, kCI_measured, kCI_first, kCI_next, kCI_flags, kCI_slotLast,
    kCI_slotPrev, kCI_N
  };
  enum { kCD_q = 0, kCD_J = 1, kCD_k4 = 26, kCD_dtds = 29, kCD_N = 30 };
  static_assert(kCI_N == KF_CARRY_I && kCD_N == KF_CARRY_D, "hand-over record layout");
  KFD_INLINE int32_t &ci(int f) const { return a.carry_i[(size_t)f * a.count + cpos]; }
  KFD_INLINE double &cd(int f) const { return a.carry_d[(size_t)f * a.count + cpos]; }

  KFD_INLINE void saveCarry(int resume, int steps, bool terminated) const {
    if (!lead) return;
    ci(kCI_resume) = resume;
    ci(kCI_steps) = steps;
    ci(kCI_measured) = measured;
    ci(kCI_first) = firstPlane;
    ci(kCI_next) = next;
    if constexpr (PS) {
      ci(kCI_slotLast) = slotLast;
      ci(kCI_slotPrev) = slotPrev;
    }
    ci(kCI_flags) = (result ? 1 : 0) | (stepFailed ? 2 : 0) | (navBreak ? 4 : 0) |
                    (terminated ? 8 : 0) | (derivSet ? 16 : 0);
    cd(kCD_q) = q;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 5; ++c) cd(kCD_J + 5 * r + c) = (double)A[r][c];
    cd(kCD_J + 15) = (double)b;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) cd(kCD_J + 16 + 3 * r + c) = (double)G[r][c];
#pragma unroll
    for (int k = 0; k < 3; ++k) cd(kCD_k4 + k) = (double)k4d[k];
    cd(kCD_dtds) = (double)dtds;
  }

  KFD_INLINE int loadCarry(int &steps, bool &terminated) {
    const int resume = ci(kCI_resume);
    steps = ci(kCI_steps);
    measured = ci(kCI_measured);
    firstPlane = ci(kCI_first);
    next = ci(kCI_next);
    if constexpr (PS) {
      slotLast = ci(kCI_slotLast);
      slotPrev = ci(kCI_slotPrev);
    }
    const int fl = ci(kCI_flags);
    result = (fl & 1) != 0;
    stepFailed = (fl & 2) != 0;
    navBreak = (fl & 4) != 0;
    terminated = (fl & 8) != 0;
    derivSet = (fl & 16) != 0;
    q = cd(kCD_q);
    return resume;
  }

  KFD_INLINE void loadCarryJacobian() {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 5; ++c) A[r][c] = (S)cd(kCD_J + 5 * r + c);
    b = (S)cd(kCD_J + 15);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) G[r][c] = (S)cd(kCD_J + 16 + 3 * r + c);
#pragma unroll
    for (int k = 0; k < 3; ++k) k4d[k] = (S)cd(kCD_k4 + k);
    dtds = (S)cd(kCD_dtds);
  }

  // ----------------------------------------------------------------- fit ----
  KFD_INLINE uint32_t statusBits(int steps, bool terminated) const {
    uint32_t status = 0;
    if (result && measured == (int)(h1 - h0)) status |= KF_STATUS_OK;
    if (smoothed) status |= KF_STATUS_SMOOTHED;
    if (finished) status |= KF_STATUS_FINISHED;
    if (!result) status |= KF_STATUS_RESULT_FAIL;
    if (stepFailed) status |= KF_STATUS_STEP_FAIL;
    if (!terminated && steps >= a.max_steps) status |= KF_STATUS_MAX_STEPS;
    if (navBreak && measured != (int)(h1 - h0)) status |= KF_STATUS_NAV_BREAK;
    if (smootherNaN) status |= KF_STATUS_SMOOTHER_NAN;
    return status;
  }

  KFD_INLINE void writeFinal(uint32_t status) const {
    if (!lead) return;
    if (a.out_params)
#pragma unroll
      for (int k = 0; k < 6; ++k) a.out_params[(size_t)k * a.n + i] = fpar[k];
    if (a.out_cov)
#pragma unroll
      for (int k = 0; k < 21; ++k) a.out_cov[(size_t)k * a.n + i] = fcov[k];
    if (a.out_status) a.out_status[i] = status;
  }

  KFD_INLINE void loadTrack() {
    h0 = a.hit_offset[i];
    h1 = a.hit_offset[i + 1];
    hitNext = h0;
    inOrder = a.unique_geo != 0;
    pfValid = false;
    seqFirst = a.seq_first[i];
    size = a.seq_count[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) lc[k] = a.perigee[(size_t)k * a.n + i];
#pragma unroll
    for (int k = 0; k < 6; ++k) fpar[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 21; ++k) fcov[k] = 0.0;
    smoothed = finished = smootherNaN = false;
    cur = -1;
    targetReached = false;
  }

  // Forward kernel body: Propagator::propagate (Propagator.ipp:3-96) up to
  // the finalisation.
  KFD_INLINE void runForward() {
    loadTrack();
    double pv[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) pv[k] = a.start_params[(size_t)k * a.n + i];
#pragma unroll
    for (int k = 0; k < 21; ++k) cov[k] = a.start_cov[(size_t)k * a.n + i];
    measured = 0;
    outliers = 0;
    firstPlane = 0;
    holes = 0;
    slotLast = slotPrev = -1;
    result = true;
    chi2sum = 0.0;
    next = 0;
    navBreak = false;
    stepFailed = false;
    bool terminated = false;
    int steps = 0;
    int resume = kResumeNone;
    const bool bad = size <= 0 || seqFirst < 0 || seqFirst + size > a.n_planes || h0 < 0 ||
                     h1 <= h0 || h1 > a.n_hits || pv[4] == 0.0 || a.max_states < 1;
    if (!bad) {
      prefetchHit();
      start(pv);
      navStatus();
      if (actor()) {
        resume = kResumePreLoop;
      } else {
        if (!(kalmanAbort() || pathLimit())) {
          navTarget();
          if (loop(steps, terminated)) resume = kResumeInLoop;
        }
        if (resume == kResumeNone) {
          if (!terminated) navBreak = true;
          if (actor()) resume = kResumePostLoop;
        }
      }
    }
    if (lead) {
      if (a.out_chi2) a.out_chi2[i] = chi2sum;
      if (a.out_ndf) a.out_ndf[i] = 2 * (measured - outliers);
      if (a.out_nmeas) a.out_nmeas[i] = measured;
    }
    if (resume != kResumeNone) {
      saveCarry(resume, steps, terminated);
    } else {
      if (lead) ci(kCI_resume) = kResumeNone;
      writeFinal(bad ? KF_STATUS_BAD_INPUT : statusBits(steps, terminated));
    }
  }

  // Backward kernel body: Actor::finalize (KalmanFitter.hpp:360-393), the
  // rest of that actor call, and the propagation back to the perigee.
  KFD_INLINE void runBackward() {
    int steps = 0;
    bool terminated = false;
    loadTrack();
    const int resume = loadCarry(steps, terminated);
    if (resume == kResumeNone) return;
    // finalize
    smoothed = true;
    double xs[6], cs[21];
    if (smooth(xs, cs)) {
      updateFromBound(plane(firstPlane), xs);
#pragma unroll
      for (int m = 0; m < 21; ++m) cov[m] = (S)cs[m];
      navDir = -1;
      moved();
      ss.init(-1.0 * fabs(a.max_step_size));
      pathAcc = 0.0;
    } else {
      result = false;
    }
    loadCarryJacobian();
    actorPost();
    if (resume == kResumePreLoop) {
      if (!(kalmanAbort() || pathLimit())) {
        navTarget();
        loop(steps, terminated);
      }
      if (!terminated) navBreak = true;
      actor();
    } else if (resume == kResumeInLoop) {
      if (kalmanAbort() || pathLimit()) {
        terminated = true;
      } else {
        navTarget();
        ++steps;
        loop(steps, terminated);
      }
      if (!terminated) navBreak = true;
      actor();
    }
    writeFinal(statusBits(steps, terminated));
  }
};

}  // namespace kfd
