// kf_adjoint.hpp -- hand-written adjoint of the summed fit chi2.
//
// Objective: X = sum over tracks and measured surfaces of the predicted-
// residual chi2 r^T S^-1 r of the forward filter (the fit's results.chi2).
// Parameters: per plane, the centre c and the rotation R (columns U, V,
// normal), and the field scale s.  The reduction kernel maps (dX/dc, dX/dR)
// to the six alignment parameters of a plane (local shifts and small local
// rotations, R' = R Rx(a) Ry(b) Rz(c), c' = c + R (du, dv, dw)).
//
// One track per lane, reverse mode by hand.  The forward filter (TrackFit)
// runs first and writes a checkpoint of the state after the update at every
// measured surface.  The reverse sweep then walks the measured surfaces from
// the last to the first; for each it reloads the checkpoint of the surface
// before, recomputes that segment's RK steps (their inputs go to a per-lane
// stack in global memory; the RK stages themselves are recomputed again in
// the backward step), recomputes the bound state, chi2 and gain update at the
// surface, and back-propagates:
//   update + chi2 -> bound state (parameters, covariance transport
//   jb C jb^T with the path-length correction) -> RK steps in reverse (stage
//   vectors, transport matrix D and J <- D J) -> the checkpoint state.
// Only the spatial 5x5 blocks enter: the time row and column of the
// covariance and Jacobians never reach the chi2.
//
// Derivative conventions: discrete decisions (step accept/reject, surface
// reached, hole) are held fixed; a step whose length was set by the
// navigator (distance to the target plane) is differentiated through that
// distance, a step set by the error control is held fixed.  The material
// update (energy loss and the scattering / straggling variances) is
// differentiated with respect to the momentum, the direction and the plane
// normal (path length through the slab).
#pragma once

#include "kf_trackfit.hpp"

namespace kfd {

// per-lane stack of the RK steps of the segment being reversed
enum {
  kSR_POS = 0,   // position before the step (3)
  kSR_DIR = 3,   // direction before the step (3)
  kSR_H = 6,     // accepted step length
  kSR_G = 7,     // J direction rows (9) before the step
  kSR_TGT = 16,  // target plane index + 1 if the step length was the navigator's, else 0
  kSR_N = 17
};

// per-track plane adjoints: slot q (sequence position) holds dX/dc (3) and
// dX/dR (9, row-major) of plane seq_first + q
enum { kPG_C = 0, kPG_R = 3, kPG_N = 12 };

// per-lane LDS slots (doubles, stride KF_BLOCK) for adjoint state of the
// reverse step loop, so that it does not occupy registers there: the
// covariance adjoint (carried across the loop, unused in it), the adjoint of
// the Jacobian's position rows A (constant over a segment's steps), of its
// direction rows G, and the next G adjoint while a step builds it
enum { kLDS_CB = 0, kLDS_AB = 25, kLDS_GB = 40, kLDS_GT = 49, kLDS_N = 58 };
// float slots: the grid-field gradient at a reversed step's three stage
// points (27); with the double slots 73.2 KB per 128-lane block, two blocks
// (four waves) per CU
enum { kLDSF_N = 27 };

struct AdjArgs {
  double *stack;          // [chunk][max_seg_steps][kSR_N][64]
  int32_t max_seg_steps;  // stack depth per lane
  int32_t max_seq;        // plane-adjoint slots per track (longest sequence)
  // per-track adjoints, indexed by the position within the launch
  // (FitArgs::first + local = track)
  double *plane_adj;      // [max_seq][kPG_N][count]
  double *scale_adj;      // [count]
  uint32_t *status;       // [n]: 0 ok, else KF_ADJ_* bits
};


// a x b
KFD_INLINE void crs(const double a[3], const double b[3], double r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
// c = a x b:  abar += b x cbar,  bbar += cbar x a
KFD_INLINE void crsVJP(const double a[3], const double b[3], const double cb[3], double ab[3],
                       double bb[3]) {
  double t[3];
  crs(b, cb, t);
#pragma unroll
  for (int k = 0; k < 3; ++k) ab[k] += t[k];
  crs(cb, a, t);
#pragma unroll
  for (int k = 0; k < 3; ++k) bb[k] += t[k];
}
// B = s F(x): sb += Bb . F(x) = Bb . B / s; xb += (dB/dx)^T Bb on a grid.
KFD_INLINE void fieldVJP(const DevField &fld, const double x[3], const double B[3],
                         const double Bb[3], double xb[3], double &sb) {
  sb += (Bb[0] * B[0] + Bb[1] * B[1] + Bb[2] * B[2]) / fld.scale;
  if (fld.grid) {
    double Bg[3], dB[3][3];
    fieldAtGrad(fld, x, Bg, dB);
#pragma unroll
    for (int a = 0; a < 3; ++a) xb[a] += Bb[0] * dB[0][a] + Bb[1] * dB[1][a] + Bb[2] * dB[2][a];
  }
}

// Grid field: the backward of a lookup whose gradient the forward recompute
// stored (dB[k][a] of stage point `pt` at dBs[(9 pt + 3 k + a) * KF_BLOCK],
// float), instead of reading the corners again.
KFD_INLINE void fieldVJPStored(const DevField &fld, const float *dBs, int pt, const double B[3],
                               const double Bb[3], double xb[3], double &sb) {
  sb += (Bb[0] * B[0] + Bb[1] * B[1] + Bb[2] * B[2]) / fld.scale;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    xb[a] += Bb[0] * (double)dBs[(9 * pt + a) * KF_BLOCK] +
             Bb[1] * (double)dBs[(9 * pt + 3 + a) * KF_BLOCK] +
             Bb[2] * (double)dBs[(9 * pt + 6 + a) * KF_BLOCK];
}

// One RK4 step (EigenStepper.ipp:226-332 at its accepted length) with every
// intermediate the backward step needs.
struct StepFwd {
  double B1[3], Bm[3], Bl[3];
  double k1[3], k2[3], k3[3], k4[3];
  double t2[3], t3[3], t4[3];
  double p1[3], p2[3];
  double L1[3], L2[3], L3[3], L4[3];
  double du[3], dun;

  // dBs: for a grid field, the field gradient at the three stage points is
  // stored there (one lookup gives the value and the gradient); nullptr: the
  // backward reads the corners again
  KFD_INLINE void lookup(const DevField &fld, const double x[3], double B[3], float *dBs,
                         int pt) {
    if (dBs && fld.grid) {
      double dB[3][3];
      fieldAtGrad(fld, x, B, dB);
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) dBs[(9 * pt + 3 * k + a) * KF_BLOCK] = (float)dB[k][a];
    } else {
      fieldAt(fld, x[0], x[1], x[2], B);
    }
  }

  KFD_INLINE void run(const DevField &fld, const double pos[3], const double d[3], double h,
                      double qop, float *dBs = nullptr) {
    const double hh = h * 0.5;
    lookup(fld, pos, B1, dBs, 0);
    crs(d, B1, k1);
#pragma unroll
    for (int k = 0; k < 3; ++k) k1[k] *= qop;
#pragma unroll
    for (int k = 0; k < 3; ++k) p1[k] = pos[k] + hh * d[k] + (h * h * 0.125) * k1[k];
    lookup(fld, p1, Bm, dBs, 1);
#pragma unroll
    for (int k = 0; k < 3; ++k) t2[k] = d[k] + hh * k1[k];
    crs(t2, Bm, k2);
#pragma unroll
    for (int k = 0; k < 3; ++k) k2[k] *= qop;
#pragma unroll
    for (int k = 0; k < 3; ++k) t3[k] = d[k] + hh * k2[k];
    crs(t3, Bm, k3);
#pragma unroll
    for (int k = 0; k < 3; ++k) k3[k] *= qop;
#pragma unroll
    for (int k = 0; k < 3; ++k) p2[k] = pos[k] + h * d[k] + (h * h * 0.5) * k3[k];
    lookup(fld, p2, Bl, dBs, 2);
#pragma unroll
    for (int k = 0; k < 3; ++k) t4[k] = d[k] + h * k3[k];
    crs(t4, Bl, k4);
#pragma unroll
    for (int k = 0; k < 3; ++k) k4[k] *= qop;
#pragma unroll
    for (int k = 0; k < 3; ++k) du[k] = d[k] + (h / 6.0) * (k1[k] + 2.0 * (k2[k] + k3[k]) + k4[k]);
    dun = sqrt(du[0] * du[0] + du[1] * du[1] + du[2] * du[2]);
    // q/p column of the transport (EigenStepper.ipp:50-124): the dk/dlambda
    // vectors
    double c[3];
    crs(d, B1, L1);
    crs(t2, Bm, L2);
    crs(L1, Bm, c);
#pragma unroll
    for (int k = 0; k < 3; ++k) L2[k] += (qop * hh) * c[k];
    crs(t3, Bm, L3);
    crs(L2, Bm, c);
#pragma unroll
    for (int k = 0; k < 3; ++k) L3[k] += (qop * hh) * c[k];
    crs(t4, Bl, L4);
    crs(L3, Bl, c);
#pragma unroll
    for (int k = 0; k < 3; ++k) L4[k] += (qop * h) * c[k];
  }
};

// The direction block of the transport applied to one column g of the
// Jacobian's direction rows: v1 = qop (g x B1), v2 = qop ((g + hh v1) x Bm),
// v3 = qop ((g + hh v2) x Bm), v4 = qop ((g + h v3) x Bl); then
// dFdT g = h g + h^2/6 (v1 + v2 + v3), dGdT g = g + h/6 (v1 + 2 v2 + 2 v3 + v4)
// (the matrices of EigenStepper.ipp:50-124 applied to a vector).  This is the
// backward of that map: aF, aG are the adjoints of dFdT g and dGdT g; it
// accumulates gb, the field adjoints, hb and qb.
KFD_INLINE void dirColumnVJP(const StepFwd &F, double h, double qop, const double g[3],
                             const double aF[3], const double aG[3], double gb[3], double B1b[3],
                             double Bmb[3], double Blb[3], double &hb, double &qb) {
  const double hh = 0.5 * h;
  // x_i = u_i x B (u_1 = g), v_i = qop x_i
  double u2[3], u3[3], u4[3], x1[3], x2[3], x3[3], x4[3], v1[3], v2[3], v3[3], v4[3];
  crs(g, F.B1, x1);
#pragma unroll
  for (int k = 0; k < 3; ++k) v1[k] = qop * x1[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) u2[k] = g[k] + hh * v1[k];
  crs(u2, F.Bm, x2);
#pragma unroll
  for (int k = 0; k < 3; ++k) v2[k] = qop * x2[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) u3[k] = g[k] + hh * v2[k];
  crs(u3, F.Bm, x3);
#pragma unroll
  for (int k = 0; k < 3; ++k) v3[k] = qop * x3[k];
#pragma unroll
  for (int k = 0; k < 3; ++k) u4[k] = g[k] + h * v3[k];
  crs(u4, F.Bl, x4);
#pragma unroll
  for (int k = 0; k < 3; ++k) v4[k] = qop * x4[k];
  double v1b[3], v2b[3], v3b[3], v4b[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    gb[k] += h * aF[k] + aG[k];
    hb += aF[k] * (g[k] + (h / 3.0) * (v1[k] + v2[k] + v3[k])) +
          aG[k] * (v1[k] + 2.0 * (v2[k] + v3[k]) + v4[k]) / 6.0;
    const double a = (h * h / 6.0) * aF[k], b = (h / 6.0) * aG[k];
    v1b[k] = a + b;
    v2b[k] = a + 2.0 * b;
    v3b[k] = a + 2.0 * b;
    v4b[k] = b;
  }
  double w[3], ub[3];
  // v4 = qop (u4 x Bl), u4 = g + h v3
  qb += v4b[0] * x4[0] + v4b[1] * x4[1] + v4b[2] * x4[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    w[k] = qop * v4b[k];
    ub[k] = 0.0;
  }
  crsVJP(u4, F.Bl, w, ub, Blb);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    gb[k] += ub[k];
    v3b[k] += h * ub[k];
    hb += ub[k] * v3[k];
  }
  // v3 = qop (u3 x Bm), u3 = g + hh v2
  qb += v3b[0] * x3[0] + v3b[1] * x3[1] + v3b[2] * x3[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    w[k] = qop * v3b[k];
    ub[k] = 0.0;
  }
  crsVJP(u3, F.Bm, w, ub, Bmb);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    gb[k] += ub[k];
    v2b[k] += hh * ub[k];
    hb += 0.5 * ub[k] * v2[k];
  }
  // v2 = qop (u2 x Bm), u2 = g + hh v1
  qb += v2b[0] * x2[0] + v2b[1] * x2[1] + v2b[2] * x2[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    w[k] = qop * v2b[k];
    ub[k] = 0.0;
  }
  crsVJP(u2, F.Bm, w, ub, Bmb);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    gb[k] += ub[k];
    v1b[k] += hh * ub[k];
    hb += 0.5 * ub[k] * v1[k];
  }
  // v1 = qop (g x B1)
  qb += v1b[0] * x1[0] + v1b[1] * x1[1] + v1b[2] * x1[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) w[k] = qop * v1b[k];
  crsVJP(g, F.B1, w, gb, B1b);
}

// Backward of one RK step.  In: the adjoints of the step's outputs (position
// pb, direction db, the Jacobian's A and G rows in the LDS slots kLDS_AB and
// kLDS_GB of `lj`, and k4b for the propagation derivative of the last step).
// Out (in place): the adjoints of its inputs (position, direction, G rows;
// the A adjoint passes through a step unchanged), and accumulated hb, qb
// (q/p) and sb (field scale).  Gin: the J direction rows before the step,
// element (r, c) at Gin[(3 r + c) * gstride].
KFD_INLINE void stepVJP(const DevField &fld, const StepFwd &F, const double pos[3],
                        const double d[3], double h, double qop, const double *Gin,
                        uint32_t gstride, double pb[3], double db[3], double *lj,
                        const double k4bIn[3], double &hb, double &qb, double &sb,
                        const float *dBs = nullptr) {
  const double hh = h * 0.5;
  double k1b[3] = {0, 0, 0}, k2b[3] = {0, 0, 0}, k3b[3] = {0, 0, 0};
  double k4b[3] = {k4bIn[0], k4bIn[1], k4bIn[2]};
  double t2b[3] = {0, 0, 0}, t3b[3] = {0, 0, 0}, t4b[3] = {0, 0, 0};
  double B1b[3] = {0, 0, 0}, Bmb[3] = {0, 0, 0}, Blb[3] = {0, 0, 0};
  double dInb[3] = {0, 0, 0}, posInb[3] = {pb[0], pb[1], pb[2]};
  // ---- J <- D J:  A[:,2+c] += dFdT G[:,c] (+ dFdL for c = 2), G[:,c] = dGdT G[:,c] (+ dGdL),
  // column by column (one column's temporaries at a time: not unrolled);
  // the next column of G is loaded while the current one is processed
  double gn[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) gn[r] = Gin[(uint32_t)(3 * r) * gstride];
#pragma unroll 1
  for (int c = 0; c < 3; ++c) {
    double g[3], aF[3], aG[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      g[r] = gn[r];
      aF[r] = lj[(kLDS_AB + 5 * r + 2 + c) * KF_BLOCK];
      aG[r] = lj[(kLDS_GB + 3 * r + c) * KF_BLOCK];
    }
    if (c < 2) {
#pragma unroll
      for (int r = 0; r < 3; ++r) gn[r] = Gin[(uint32_t)(3 * r + c + 1) * gstride];
    }
    double gb[3] = {0.0, 0.0, 0.0};
    dirColumnVJP(F, h, qop, g, aF, aG, gb, B1b, Bmb, Blb, hb, qb);
#pragma unroll
    for (int q = 0; q < 3; ++q) lj[(kLDS_GT + 3 * q + c) * KF_BLOCK] = gb[q];
  }
  // ---- dFdL = h^2/6 (L1 + L2 + L3), dGdL = h/6 (L1 + 2 L2 + 2 L3 + L4)
  double dFdLb[3], dGdLb[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    dFdLb[r] = lj[(kLDS_AB + 5 * r + 4) * KF_BLOCK];
    dGdLb[r] = lj[(kLDS_GB + 3 * r + 2) * KF_BLOCK];
  }
  double L1b[3], L2b[3], L3b[3], L4b[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    hb += dFdLb[r] * (h / 3.0) * (F.L1[r] + F.L2[r] + F.L3[r]) +
          dGdLb[r] * (F.L1[r] + 2.0 * (F.L2[r] + F.L3[r]) + F.L4[r]) / 6.0;
    const double a = (h * h / 6.0) * dFdLb[r], g = (h / 6.0) * dGdLb[r];
    L1b[r] = a + g;
    L2b[r] = a + 2.0 * g;
    L3b[r] = a + 2.0 * g;
    L4b[r] = g;
  }
  // ---- L chain: L4 = t4 x Bl + (qop h)(L3 x Bl); L3, L2 with Bm and hh; L1 = d x B1
  {
    double c[3];
    crsVJP(F.t4, F.Bl, L4b, t4b, Blb);
    crs(F.L3, F.Bl, c);
    qb += h * (L4b[0] * c[0] + L4b[1] * c[1] + L4b[2] * c[2]);
    hb += qop * (L4b[0] * c[0] + L4b[1] * c[1] + L4b[2] * c[2]);
    double w[3] = {qop * h * L4b[0], qop * h * L4b[1], qop * h * L4b[2]};
    crsVJP(F.L3, F.Bl, w, L3b, Blb);

    crsVJP(F.t3, F.Bm, L3b, t3b, Bmb);
    crs(F.L2, F.Bm, c);
    qb += hh * (L3b[0] * c[0] + L3b[1] * c[1] + L3b[2] * c[2]);
    hb += 0.5 * qop * (L3b[0] * c[0] + L3b[1] * c[1] + L3b[2] * c[2]);
    double w3[3] = {qop * hh * L3b[0], qop * hh * L3b[1], qop * hh * L3b[2]};
    crsVJP(F.L2, F.Bm, w3, L2b, Bmb);

    crsVJP(F.t2, F.Bm, L2b, t2b, Bmb);
    crs(F.L1, F.Bm, c);
    qb += hh * (L2b[0] * c[0] + L2b[1] * c[1] + L2b[2] * c[2]);
    hb += 0.5 * qop * (L2b[0] * c[0] + L2b[1] * c[1] + L2b[2] * c[2]);
    double w2[3] = {qop * hh * L2b[0], qop * hh * L2b[1], qop * hh * L2b[2]};
    crsVJP(F.L1, F.Bm, w2, L1b, Bmb);

    crsVJP(d, F.B1, L1b, dInb, B1b);
  }
  // ---- outputs: pos' = pos + h d + h^2/6 (k1+k2+k3); d' = du / |du|
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dInb[k] += h * pb[k];
    hb += pb[k] * (d[k] + (h / 3.0) * (F.k1[k] + F.k2[k] + F.k3[k]));
    k1b[k] += (h * h / 6.0) * pb[k];
    k2b[k] += (h * h / 6.0) * pb[k];
    k3b[k] += (h * h / 6.0) * pb[k];
  }
  {
    double dn[3] = {F.du[0] / F.dun, F.du[1] / F.dun, F.du[2] / F.dun};
    const double pr = dn[0] * db[0] + dn[1] * db[1] + dn[2] * db[2];
    double dub[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) dub[k] = (db[k] - dn[k] * pr) / F.dun;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dInb[k] += dub[k];
      hb += dub[k] * (F.k1[k] + 2.0 * (F.k2[k] + F.k3[k]) + F.k4[k]) / 6.0;
      k1b[k] += (h / 6.0) * dub[k];
      k2b[k] += (h / 3.0) * dub[k];
      k3b[k] += (h / 3.0) * dub[k];
      k4b[k] += (h / 6.0) * dub[k];
    }
  }
  // ---- stages, last to first
  double c[3], w[3];
  // k4 = qop (t4 x Bl)
  crs(F.t4, F.Bl, c);
  qb += k4b[0] * c[0] + k4b[1] * c[1] + k4b[2] * c[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) w[k] = qop * k4b[k];
  crsVJP(F.t4, F.Bl, w, t4b, Blb);
  // t4 = d + h k3
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dInb[k] += t4b[k];
    hb += t4b[k] * F.k3[k];
    k3b[k] += h * t4b[k];
  }
  // Bl = s F(p2)
  double p2b[3] = {0, 0, 0};
  if (dBs && fld.grid)
    fieldVJPStored(fld, dBs, 2, F.Bl, Blb, p2b, sb);
  else
    fieldVJP(fld, F.p2, F.Bl, Blb, p2b, sb);
  // p2 = pos + h d + h^2/2 k3
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    posInb[k] += p2b[k];
    dInb[k] += h * p2b[k];
    hb += p2b[k] * (d[k] + h * F.k3[k]);
    k3b[k] += (h * h * 0.5) * p2b[k];
  }
  // k3 = qop (t3 x Bm)
  crs(F.t3, F.Bm, c);
  qb += k3b[0] * c[0] + k3b[1] * c[1] + k3b[2] * c[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) w[k] = qop * k3b[k];
  crsVJP(F.t3, F.Bm, w, t3b, Bmb);
  // t3 = d + hh k2
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dInb[k] += t3b[k];
    hb += 0.5 * t3b[k] * F.k2[k];
    k2b[k] += hh * t3b[k];
  }
  // k2 = qop (t2 x Bm)
  crs(F.t2, F.Bm, c);
  qb += k2b[0] * c[0] + k2b[1] * c[1] + k2b[2] * c[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) w[k] = qop * k2b[k];
  crsVJP(F.t2, F.Bm, w, t2b, Bmb);
  // t2 = d + hh k1
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    dInb[k] += t2b[k];
    hb += 0.5 * t2b[k] * F.k1[k];
    k1b[k] += hh * t2b[k];
  }
  // Bm = s F(p1)
  double p1b[3] = {0, 0, 0};
  if (dBs && fld.grid)
    fieldVJPStored(fld, dBs, 1, F.Bm, Bmb, p1b, sb);
  else
    fieldVJP(fld, F.p1, F.Bm, Bmb, p1b, sb);
  // p1 = pos + hh d + h^2/8 k1
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    posInb[k] += p1b[k];
    dInb[k] += hh * p1b[k];
    hb += p1b[k] * (0.5 * d[k] + 0.25 * h * F.k1[k]);
    k1b[k] += (h * h * 0.125) * p1b[k];
  }
  // k1 = qop (d x B1)
  crs(d, F.B1, c);
  qb += k1b[0] * c[0] + k1b[1] * c[1] + k1b[2] * c[2];
#pragma unroll
  for (int k = 0; k < 3; ++k) w[k] = qop * k1b[k];
  crsVJP(d, F.B1, w, dInb, B1b);
  // B1 = s F(pos)
  if (dBs && fld.grid)
    fieldVJPStored(fld, dBs, 0, F.B1, B1b, posInb, sb);
  else
    fieldVJP(fld, pos, F.B1, B1b, posInb, sb);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pb[k] = posInb[k];
    db[k] = dInb[k];
  }
#pragma unroll
  for (int m = 0; m < 9; ++m) lj[(kLDS_GB + m) * KF_BLOCK] = lj[(kLDS_GT + m) * KF_BLOCK];
}

// ---------------------------------------------------------------------------
// Bound state on a plane and gain-matrix update, spatial 5x5 part, with the
// intermediates their backward pass needs (EigenStepper.ipp:516-540,
// CovarianceEngine.hpp:269-290, GainMatrixUpdater.hpp:34-106).
struct PlaneFwd {
  double u0[3], u1[3], nr[3], cc[3];
  double rho, l20, l21, l30, l31, l32;
  double dp, nv[3], sv[5];
  double jpP[3][5], jpD[3][5];
  double jb[5][5];
  double Cm[5][5], Cp[5][5];
  double xp[5];
  double V00, V01, V11, Si00, Si01, Si11, r[2], u[2], xf[5], chi2;

  // d, pos: direction and position at the plane; A, G: J at the plane; k4:
  // propagation derivative of the last step; Cm: covariance at the segment
  // start; m/V: the hit.
  KFD_INLINE void run(const DevPlane &pl, const double d[3], const double pos[3],
                      const double A[3][5], const double G[3][3], const double k4[3],
                      double qop, const double cm[21], double m0, double m1, double v00,
                      double v01, double v11) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      u0[k] = pl.r[3 * k];
      u1[k] = pl.r[3 * k + 1];
      nr[k] = pl.r[3 * k + 2];
      cc[k] = pl.c[k];
    }
    rho = sqrt(d[0] * d[0] + d[1] * d[1]);
    const double cosT = d[2], sinT = rho, invSinT = 1.0 / rho;
    const double cosP = d[0] * invSinT, sinP = d[1] * invSinT;
    l20 = -sinP * invSinT;
    l21 = cosP * invSinT;
    l30 = cosP * cosT;
    l31 = sinP * cosT;
    l32 = -sinT;
    dp = nr[0] * d[0] + nr[1] * d[1] + nr[2] * d[2];
    const double idp = 1.0 / dp;
#pragma unroll
    for (int k = 0; k < 3; ++k) nv[k] = nr[k] * idp;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      sv[c] = nv[0] * A[0][c] + nv[1] * A[1][c] + nv[2] * A[2][c];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        jpP[q][c] = A[q][c] - d[q] * sv[c];
        jpD[q][c] = ((c >= 2) ? G[q][c - 2] : 0.0) - k4[q] * sv[c];
      }
      jb[0][c] = u0[0] * jpP[0][c] + u0[1] * jpP[1][c] + u0[2] * jpP[2][c];
      jb[1][c] = u1[0] * jpP[0][c] + u1[1] * jpP[1][c] + u1[2] * jpP[2][c];
      jb[2][c] = l20 * jpD[0][c] + l21 * jpD[1][c];
      jb[3][c] = l30 * jpD[0][c] + l31 * jpD[1][c] + l32 * jpD[2][c];
      jb[4][c] = (c == 4) ? 1.0 : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
      for (int c = 0; c < 5; ++c) Cm[q][c] = cm[sidx(q, c)];
    double T[5][5];  // jb Cm
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 5; ++m) s += jb[q][m] * Cm[m][c];
        T[q][c] = s;
      }
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 5; ++m) s += T[q][m] * jb[c][m];
        Cp[q][c] = s;
      }
    xp[0] = u0[0] * (pos[0] - cc[0]) + u0[1] * (pos[1] - cc[1]) + u0[2] * (pos[2] - cc[2]);
    xp[1] = u1[0] * (pos[0] - cc[0]) + u1[1] * (pos[1] - cc[1]) + u1[2] * (pos[2] - cc[2]);
    xp[2] = oolAtan2(d[1], d[0]);
    xp[3] = oolAtan2(rho, d[2]);
    xp[4] = qop;
    // gain-matrix update
    V00 = v00;
    V01 = v01;
    V11 = v11;
    const double S00 = Cp[0][0] + V00, S01 = Cp[0][1] + V01, S11 = Cp[1][1] + V11;
    const double det = S00 * S11 - S01 * S01;
    const double idet = 1.0 / det;
    Si00 = S11 * idet;
    Si01 = -S01 * idet;
    Si11 = S00 * idet;
    r[0] = m0 - xp[0];
    r[1] = m1 - xp[1];
    u[0] = Si00 * r[0] + Si01 * r[1];
    u[1] = Si01 * r[0] + Si11 * r[1];
    chi2 = r[0] * u[0] + r[1] * u[1];
#pragma unroll
    for (int q = 0; q < 5; ++q) xf[q] = xp[q] + Cp[q][0] * u[0] + Cp[q][1] * u[1];
  }

  // Backward of chi2 + update + bound state.  In: xfb (adjoint of the
  // filtered parameters), Cfb (symmetric adjoint of the filtered covariance),
  // chi2 weight 1.  Out: the adjoints of pos, d, A, G, k4, q/p and of the
  // segment-start covariance (Cmb, symmetric), and of the plane (cb, Rb).
  KFD_INLINE void vjp(const double d[3], const double pos[3], const double A[3][5],
                      const double k4[3], const double xfb[5], const double Cfb[5][5],
                      double posb[3], double db[3], double Ab[3][5], double Gb[3][3],
                      double k4b[3], double &qb, double Cmb[5][5], double cb[3],
                      double Rb[3][3]) const {
    // ---- update: xf = xp + W u, Cf = Cp - W Si W^T, chi2 = r^T u, u = Si r, r = m - H xp
    double xpb[5], Cpb[5][5];
    const double W[5][2] = {{Cp[0][0], Cp[0][1]}, {Cp[1][0], Cp[1][1]}, {Cp[2][0], Cp[2][1]},
                            {Cp[3][0], Cp[3][1]}, {Cp[4][0], Cp[4][1]}};
    const double Si[2][2] = {{Si00, Si01}, {Si01, Si11}};
    double Wb[5][2], ub[2] = {0, 0}, rb[2], Sib[2][2];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      xpb[q] = xfb[q];
      Wb[q][0] = xfb[q] * u[0];
      Wb[q][1] = xfb[q] * u[1];
      ub[0] += W[q][0] * xfb[q];
      ub[1] += W[q][1] * xfb[q];
#pragma unroll
      for (int c = 0; c < 5; ++c) Cpb[q][c] = Cfb[q][c];
    }
    // chi2 = r^T Si r (weight 1)
    rb[0] = 2.0 * u[0];
    rb[1] = 2.0 * u[1];
#pragma unroll
    for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
      for (int b2 = 0; b2 < 2; ++b2) Sib[a2][b2] = r[a2] * r[b2] + ub[a2] * r[b2];
    rb[0] += Si[0][0] * ub[0] + Si[0][1] * ub[1];
    rb[1] += Si[1][0] * ub[0] + Si[1][1] * ub[1];
    // Cf term: Wb -= 2 Cfb W Si, Sib -= W^T Cfb W
    {
      double CW[5][2];
#pragma unroll
      for (int q = 0; q < 5; ++q)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          double s = 0.0;
#pragma unroll
          for (int m = 0; m < 5; ++m) s += Cfb[q][m] * W[m][j];
          CW[q][j] = s;
        }
#pragma unroll
      for (int q = 0; q < 5; ++q)
#pragma unroll
        for (int j = 0; j < 2; ++j) Wb[q][j] -= 2.0 * (CW[q][0] * Si[0][j] + CW[q][1] * Si[1][j]);
#pragma unroll
      for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2) {
          double s = 0.0;
#pragma unroll
          for (int m = 0; m < 5; ++m) s += W[m][a2] * CW[m][b2];
          Sib[a2][b2] -= s;
        }
    }
    // Si = S^-1: Sb = -Si Sib Si, symmetrised; S = Cp[0:2,0:2] + V
    {
      double t[2][2], Sb[2][2];
#pragma unroll
      for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2) t[a2][b2] = Sib[a2][0] * Si[0][b2] + Sib[a2][1] * Si[1][b2];
#pragma unroll
      for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
        for (int b2 = 0; b2 < 2; ++b2) Sb[a2][b2] = -(Si[a2][0] * t[0][b2] + Si[a2][1] * t[1][b2]);
      const double off = 0.5 * (Sb[0][1] + Sb[1][0]);
      Cpb[0][0] += Sb[0][0];
      Cpb[1][1] += Sb[1][1];
      Cpb[0][1] += off;
      Cpb[1][0] += off;
    }
    // W = Cp[:, 0:2] of the symmetric Cp
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        Cpb[q][j] += 0.5 * Wb[q][j];
        Cpb[j][q] += 0.5 * Wb[q][j];
      }
    xpb[0] -= rb[0];
    xpb[1] -= rb[1];

    // ---- covariance transport Cp = jb Cm jb^T
    double jbb[5][5];
    {
      double T[5][5];  // Cpb jb
#pragma unroll
      for (int q = 0; q < 5; ++q)
#pragma unroll
        for (int c = 0; c < 5; ++c) {
          double s = 0.0;
#pragma unroll
          for (int m = 0; m < 5; ++m) s += Cpb[q][m] * jb[m][c];
          T[q][c] = s;
        }
#pragma unroll
      for (int q = 0; q < 5; ++q)
#pragma unroll
        for (int c = 0; c < 5; ++c) {
          double s = 0.0, w2 = 0.0;
#pragma unroll
          for (int m = 0; m < 5; ++m) {
            s += jb[m][q] * T[m][c];  // jb^T Cpb jb
            w2 += T[q][m] * Cm[m][c];  // Cpb jb Cm
          }
          Cmb[q][c] = s;
          jbb[q][c] = 2.0 * w2;
        }
    }
    // ---- bound Jacobian rows
    double l20b = 0, l21b = 0, l30b = 0, l31b = 0, l32b = 0, svb[5], nvb[3] = {0, 0, 0};
    double u0b[3] = {0, 0, 0}, u1b[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      posb[k] = 0.0;
      db[k] = 0.0;
      k4b[k] = 0.0;
      cb[k] = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) Gb[k][c] = 0.0;
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      double jpPb[3], jpDb[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        jpPb[q] = jbb[0][c] * u0[q] + jbb[1][c] * u1[q];
        u0b[q] += jbb[0][c] * jpP[q][c];
        u1b[q] += jbb[1][c] * jpP[q][c];
      }
      jpDb[0] = jbb[2][c] * l20 + jbb[3][c] * l30;
      jpDb[1] = jbb[2][c] * l21 + jbb[3][c] * l31;
      jpDb[2] = jbb[3][c] * l32;
      l20b += jbb[2][c] * jpD[0][c];
      l21b += jbb[2][c] * jpD[1][c];
      l30b += jbb[3][c] * jpD[0][c];
      l31b += jbb[3][c] * jpD[1][c];
      l32b += jbb[3][c] * jpD[2][c];
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        Ab[q][c] = jpPb[q];
        db[q] -= sv[c] * jpPb[q];
        s -= d[q] * jpPb[q];
        if (c >= 2) Gb[q][c - 2] = jpDb[q];
        k4b[q] -= sv[c] * jpDb[q];
        s -= k4[q] * jpDb[q];
      }
      svb[c] = s;
    }
    // sv[c] = nv . A[:,c], nv = nr / dp, dp = nr . d
#pragma unroll
    for (int c = 0; c < 5; ++c)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        nvb[q] += svb[c] * A[q][c];
        Ab[q][c] += svb[c] * nv[q];
      }
    double nrb[3];
    const double dpb = -(nvb[0] * nr[0] + nvb[1] * nr[1] + nvb[2] * nr[2]) / (dp * dp);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      nrb[q] = nvb[q] / dp + dpb * d[q];
      db[q] += dpb * nr[q];
    }
    // l's as functions of d (rho^2 = d0^2 + d1^2)
    {
      const double r2 = rho * rho, r3 = r2 * rho, r4 = r2 * r2;
      const double d0 = d[0], d1 = d[1], d2 = d[2];
      db[0] += l20b * (2.0 * d1 * d0 / r4) + l21b * (1.0 / r2 - 2.0 * d0 * d0 / r4) +
               l30b * (d2 / rho - d0 * d0 * d2 / r3) + l31b * (-d0 * d1 * d2 / r3) +
               l32b * (-d0 / rho);
      db[1] += l20b * (-1.0 / r2 + 2.0 * d1 * d1 / r4) + l21b * (-2.0 * d0 * d1 / r4) +
               l30b * (-d0 * d1 * d2 / r3) + l31b * (d2 / rho - d1 * d1 * d2 / r3) +
               l32b * (-d1 / rho);
      db[2] += l30b * (d0 / rho) + l31b * (d1 / rho);
      // parameters: phi = atan2(d1, d0), theta = atan2(rho, d2)
      const double rr = r2 + d2 * d2;
      db[0] += xpb[2] * (-d1 / r2) + xpb[3] * (d2 / rr) * (d0 / rho);
      db[1] += xpb[2] * (d0 / r2) + xpb[3] * (d2 / rr) * (d1 / rho);
      db[2] += xpb[3] * (-rho / rr);
    }
    // loc = (u0, u1) . (pos - c)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      u0b[q] += xpb[0] * (pos[q] - cc[q]);
      u1b[q] += xpb[1] * (pos[q] - cc[q]);
      posb[q] += xpb[0] * u0[q] + xpb[1] * u1[q];
      cb[q] -= xpb[0] * u0[q] + xpb[1] * u1[q];
    }
    qb += xpb[4];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      Rb[q][0] = u0b[q];
      Rb[q][1] = u1b[q];
      Rb[q][2] = nrb[q];
    }
  }
};

// Backward of the free state made from the filtered parameters on the plane
// (updateFromBound): pos = c + u0 xf0 + u1 xf1, dir = dir(phi, theta),
// p = 1 / |xf4|.  pBar is the adjoint of that momentum.  Accumulates into
// xfb, cb, Rb.
KFD_INLINE void freeFromBoundVJP(const PlaneFwd &P, const double posb[3], const double db[3],
                                 double pBar, double xfb[5], double cb[3], double Rb[3][3]) {
  const double ph = P.xf[2], th = P.xf[3];
  const SinCos sph = oolSinCos(ph), sth = oolSinCos(th);
  const double cp = sph.c, sp = sph.s, ct = sth.c, st = sth.s;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    xfb[0] += P.u0[k] * posb[k];
    xfb[1] += P.u1[k] * posb[k];
    Rb[k][0] += P.xf[0] * posb[k];
    Rb[k][1] += P.xf[1] * posb[k];
    cb[k] += posb[k];
  }
  xfb[2] += db[0] * (-sp * st) + db[1] * (cp * st);
  xfb[3] += db[0] * (cp * ct) + db[1] * (sp * ct) + db[2] * (-st);
  const double x4 = P.xf[4];
  xfb[4] += pBar * (-((x4 < 0.0) ? -1.0 : 1.0) / (x4 * x4));
}

// Backward of the Jacobian re-initialisation on a plane (PlaneSurface.ipp:
// 122-156): A[:,0] = u0, A[:,1] = u1, G[0][0] = -d1, G[0][1] = d2 d0 / rho,
// G[1][0] = d0, G[1][1] = d2 d1 / rho, G[2][1] = -rho (d: direction at the
// plane before the update).
KFD_INLINE void jacInitVJP(const double d[3], const double Ab[3][5], const double Gb[3][3],
                           double db[3], double Rb[3][3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    Rb[k][0] += Ab[k][0];
    Rb[k][1] += Ab[k][1];
  }
  const double rho = sqrt(d[0] * d[0] + d[1] * d[1]), r3 = rho * rho * rho;
  const double d0 = d[0], d1 = d[1], d2 = d[2];
  db[0] += Gb[0][1] * (d2 / rho - d2 * d0 * d0 / r3) + Gb[1][0] +
           Gb[1][1] * (-d2 * d1 * d0 / r3) + Gb[2][1] * (-d0 / rho);
  db[1] += -Gb[0][0] + Gb[0][1] * (-d2 * d0 * d1 / r3) +
           Gb[1][1] * (d2 / rho - d2 * d1 * d1 / r3) + Gb[2][1] * (-d1 / rho);
  db[2] += Gb[0][1] * (d0 / rho) + Gb[1][1] * (d1 / rho);
}

// ---------------------------------------------------------------------------
// Material update on a plane (TrackFit::material in double: Bethe mean loss,
// Highland scattering, Landau straggling in q/p), with its backward pass.
// Maps (p, d, plane normal) to the new momentum and the three variances
// added to cov(2,2), cov(3,3), cov(4,4).
struct MaterialFwd {
  bool on;
  double p, m, sgn, c, pcor, q2b2, dq2b2, eps, depsP, depsC, running, drunP, Eloss;
  double E0, nextE, pOut;
  double tt, theta0, rho, dn, sigPhi, vPhi, vTheta, sq, vQop;
  bool ms, el, dens;

  KFD_INLINE void run(const FitArgs &a, const DevPlane &pl, double pIn, const double d[3]) {
    on = pl.has_material && (a.ms || a.eloss);
    ms = a.ms;
    el = a.eloss;
    p = pIn;
    m = a.mass;
    pOut = p;
    vPhi = vTheta = vQop = 0.0;
    if (!on) return;
    const double nr[3] = {pl.r[2], pl.r[5], pl.r[8]};
    const double nd = nr[0] * d[0] + nr[1] * d[1] + nr[2] * d[2];
    sgn = (nd < 0.0) ? -1.0 : 1.0;
    c = fabs(nd);
    pcor = 1.0 / c;
    const double thick = pl.thickness * pcor, xOverX0 = pl.thicknessInX0 * pcor;
    q2b2 = 1.0 + (m / p) * (m / p);
    dq2b2 = -2.0 * m * m / (p * p * p);
    const double K = 0.5 * a.kK * pl.molarElectronDensity;
    eps = K * thick * q2b2;
    depsP = K * thick * dq2b2;
    depsC = K * pl.thickness * q2b2;  // d eps / d pcor
    running = 0.0;
    drunP = 0.0;
    Eloss = 0.0;
    if (el) {
      const double I = pl.meanExcitation, me = a.kMe, bg = p / m, gamma = sqrt(1.0 + bg * bg);
      const double mfrac = me / m, den = 1.0 + 2.0 * gamma * mfrac + mfrac * mfrac;
      const double u = 2.0 * me * bg * bg, wmax = u / den;
      const double beta2 = 1.0 / (1.0 + (m / p) * (m / p));
      dens = !(bg < 10.0);
      const double dhalf =
          dens ? oolLog(bg) + oolLog(a.kPlasma * sqrt(pl.molarElectronDensity) / I) - 0.5 : 0.0;
      running = 0.5 * oolLog(u / I) + 0.5 * oolLog(wmax / I) - beta2 - dhalf;
      const double dgamma = bg / (gamma * m);
      const double dbeta2 = 2.0 * p * m * m / ((p * p + m * m) * (p * p + m * m));
      drunP = 1.0 / p + 0.5 * (2.0 / p - 2.0 * mfrac * dgamma / den) - dbeta2 - (dens ? 1.0 / p : 0.0);
      Eloss = eps * running;
    }
    if (ms) {
      tt = sqrt(xOverX0 * q2b2);
      theta0 = a.kHighland * (1.0 / p) * tt * (1.0 + a.kHighlandLog * oolLog(tt));
      rho = sqrt(d[0] * d[0] + d[1] * d[1]);
      dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      sigPhi = theta0 * (dn / rho);
      vPhi = sigPhi * sigPhi;
      vTheta = theta0 * theta0;
    }
    if (el) {
      sq = sqrt(q2b2) * (1.0 / (p * p)) * (4.0 * eps / a.kLandauConv);
      vQop = sq * sq;
    }
    E0 = sqrt(m * m + p * p);
    nextE = E0 - Eloss;
    pOut = (m < nextE) ? sqrt(nextE * nextE - m * m) : 0.0;
  }

  // In: pOutb, the adjoints of the three variances.  Out: the adjoint of the
  // incoming momentum; accumulates into db and nb (plane normal).
  KFD_INLINE double vjp(const FitArgs &a, const DevPlane &pl, const double d[3], double pOutb,
                        double vPhib, double vThetab, double vQopb, double db[3],
                        double nb[3]) const {
    if (!on) return pOutb;
    double pb = 0.0, pcorb = 0.0;
    // p' = sqrt(nextE^2 - m^2), nextE = E0 - Eloss
    double Elossb = 0.0;
    if (m < nextE && pOut > 0.0) {
      const double nEb = pOutb * nextE / pOut;
      pb += nEb * p / E0;
      Elossb = -nEb;
    }
    double epsb = 0.0;
    if (el) {
      // vQop = sq^2, sq = sqrt(q2b2) p^-2 (4 eps / conv)
      const double sqb = 2.0 * sq * vQopb;
      const double s2 = sqrt(q2b2), f = 4.0 / a.kLandauConv;
      pb += sqb * (s2 * (-2.0 / (p * p * p)) * f * eps + (dq2b2 / (2.0 * s2)) * (1.0 / (p * p)) * f * eps);
      epsb += sqb * s2 * (1.0 / (p * p)) * f;
      // Eloss = eps running(p)
      epsb += Elossb * running;
      pb += Elossb * eps * drunP;
    }
    if (ms) {
      // vPhi = (theta0 dn / rho)^2, vTheta = theta0^2
      const double sigb = 2.0 * sigPhi * vPhib;
      double theta0b = 2.0 * theta0 * vThetab + sigb * (dn / rho);
      const double rb = sigb * theta0;  // adjoint of dn / rho
      const double r3 = rho * rho * rho;
      db[0] += rb * (d[0] / (dn * rho) - dn * d[0] / r3);
      db[1] += rb * (d[1] / (dn * rho) - dn * d[1] / r3);
      db[2] += rb * (d[2] / (dn * rho));
      // theta0 = kH tt (1 + kHL log tt) / p, tt = sqrt(xOverX0 q2b2)
      const double dth_dtt = a.kHighland / p * (1.0 + a.kHighlandLog * oolLog(tt) + a.kHighlandLog);
      const double ttb = theta0b * dth_dtt;
      pb += theta0b * (-theta0 / p);
      const double xOverX0 = pl.thicknessInX0 * pcor;
      pb += ttb * (xOverX0 / (2.0 * tt)) * dq2b2;
      pcorb += ttb * (q2b2 / (2.0 * tt)) * pl.thicknessInX0;
    }
    // eps = K thick(pcor) q2b2(p)
    pb += epsb * depsP;
    pcorb += epsb * depsC;
    // pcor = 1 / |n . d|
    const double cb = -pcorb / (c * c) * sgn;
    const double nr[3] = {pl.r[2], pl.r[5], pl.r[8]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      db[k] += cb * nr[k];
      nb[k] += cb * d[k];
    }
    return pb;
  }
};

// ---------------------------------------------------------------------------
// The reverse sweep of one track.
struct TrackAdjoint {
  const FitArgs &a;
  const AdjArgs &g;
  const int64_t i;    // track
  const int64_t loc;  // position within the launch
  double *stk;  // this wave's step stack
  uint32_t lane;
  double *lds;  // this lane's LDS slots: lds[f * KF_BLOCK]
  float *ldf;   // this lane's float LDS slots: the grid-field gradient of a reversed step
  TrackFit<double, true, 1, false, true> tf;  // recomputes segments from the checkpoints

  KFD_INLINE TrackAdjoint(const FitArgs &args, const AdjArgs &adj, int64_t idx, int64_t local,
                          double *ckpt, double *stack, uint32_t ln, double *ldsLane,
                          float *ldfLane)
      : a(args), g(adj), i(idx), loc(local), stk(stack), lane(ln), lds(ldsLane), ldf(ldfLane),
        tf(args, idx, nullptr, nullptr, nullptr, ckpt, ln) {}

  KFD_INLINE double *sr(int j, int f) const {
    return stk + (uint32_t)(((uint32_t)j * (uint32_t)kSR_N + (uint32_t)f) * 64u + lane);
  }
  KFD_INLINE double *pg(int q, int f) const {
    return g.plane_adj + ((size_t)q * kPG_N + f) * (size_t)a.count + loc;
  }
  KFD_INLINE void addPlane(int plane, const double cb[3], const double Rb[3][3]) const {
    const int q = plane - tf.seqFirst;
    if (q < 0 || q >= g.max_seq) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) *pg(q, kPG_C + k) += cb[k];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) *pg(q, kPG_R + 3 * r + c) += Rb[r][c];
  }
  KFD_INLINE int64_t hitOn(int plane) const {
    const int32_t id = a.planes[plane].geo_id;
    for (int64_t h = tf.h0; h < tf.h1; ++h)
      if (a.hit_geo_id[h] == id) return h;
    return -1;
  }

  KFD_INLINE void run() {
    tf.loadTrack();
    const int nmeas = a.out_nmeas[i];
    const int slots = (tf.size < g.max_seq) ? tf.size : g.max_seq;
    for (int q = 0; q < slots; ++q)
      for (int f = 0; f < kPG_N; ++f) *pg(q, f) = 0.0;
    uint32_t st = 0;
    double sb = 0.0;
    if (nmeas <= 0) {
      g.scale_adj[loc] = 0.0;
      if (g.status) g.status[i] = KF_ADJ_SKIPPED;
      return;
    }
    // adjoint of the state after the update at the surface being processed:
    // position, direction, q/p in registers; the Jacobian (A and G rows) and
    // covariance adjoints in this lane's LDS slots, read where consumed
    double pb[3] = {0, 0, 0}, db[3] = {0, 0, 0}, qb = 0.0;
#pragma unroll
    for (int m = 0; m < kLDS_N; ++m) lds[m * KF_BLOCK] = 0.0;

    // the stepper charge is set once from the start parameters and never
    // updated (the checkpoints do not carry it)
    const double charge = (a.start_params[(size_t)4 * a.n + i] > 0.0) ? 1.0 : -1.0;
    for (int k = nmeas - 1; k >= 0 && st == 0; --k) {
      // ---- recompute segment k from the checkpoint of surface k - 1
      tf.result = true;
      tf.finished = tf.smoothed = false;
      tf.navBreak = false;
      tf.stepFailed = false;
      tf.targetReached = false;
      tf.cur = -1;
      if (k == 0) {
        double pv[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) pv[m] = a.start_params[(size_t)m * a.n + i];
#pragma unroll
        for (int m = 0; m < 21; ++m) tf.cov[m] = a.start_cov[(size_t)m * a.n + i];
        tf.measured = 0;
        tf.outliers = 0;
        tf.next = 0;
        tf.start(pv);
        tf.navStatus();
        if (tf.cur >= 0 && hitOn(tf.cur) >= 0) {  // measured without a step
          st |= KF_ADJ_LOST;
          break;
        }
      } else {
        tf.loadCheckpoint(k - 1);
        tf.q = charge;
        tf.measured = k;
      }
      tf.pathLimit();
      tf.navTarget();
      const double qop = tf.q / tf.p;
      int nst = 0;
      int plane = -1;
      for (;;) {
        if (nst >= g.max_seg_steps) {
          st |= KF_ADJ_OVERFLOW;
          break;
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          *sr(nst, kSR_POS + m) = tf.pos[m];
          *sr(nst, kSR_DIR + m) = tf.dir[m];
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) *sr(nst, kSR_G + 3 * r + c) = tf.G[r][c];
        const int target = (tf.next < tf.size) ? tf.seqFirst + tf.next : -1;
        if (!tf.step()) {
          st |= KF_ADJ_LOST;
          break;
        }
        const double h = tf.ss.get();
        *sr(nst, kSR_H) = h;
        *sr(nst, kSR_TGT) = (target >= 0 && h == tf.ss.v[kActor]) ? (double)(target + 1) : 0.0;
        ++nst;
        tf.navStatus();
        if (tf.cur >= 0 && hitOn(tf.cur) >= 0) {
          plane = tf.cur;
          break;
        }
        if (tf.pathLimit()) {
          st |= KF_ADJ_LOST;
          break;
        }
        tf.navTarget();
        if (tf.navBreak) {
          st |= KF_ADJ_LOST;
          break;
        }
      }
      if (st) break;
      // ---- bound state, chi2, update at the measured plane
      const DevPlane &pl = a.planes[plane];
      const int64_t hit = hitOn(plane);
      double dEnd[3], posEnd[3], k4[3], AEnd[3][5], GEnd[3][3];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        dEnd[m] = tf.dir[m];
        posEnd[m] = tf.pos[m];
        k4[m] = tf.k4d[m];
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 5; ++c) AEnd[r][c] = tf.A[r][c];
#pragma unroll
        for (int c = 0; c < 3; ++c) GEnd[r][c] = tf.G[r][c];
      }
      PlaneFwd P;
      P.run(pl, dEnd, posEnd, AEnd, GEnd, k4, qop, tf.cov, a.hit_meas[hit],
            a.hit_meas[a.n_hits + hit], a.hit_cov[hit], a.hit_cov[a.n_hits + hit],
            a.hit_cov[2 * a.n_hits + hit]);
      // ---- backward at the plane
      double cb[3] = {0, 0, 0}, Rb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      double dPre[3] = {0, 0, 0};  // adjoint of the direction at the plane (before the update)
      {
        double Ab[3][5], Gb[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
#pragma unroll
          for (int c = 0; c < 5; ++c) Ab[r][c] = lds[(kLDS_AB + 5 * r + c) * KF_BLOCK];
#pragma unroll
          for (int c = 0; c < 3; ++c) Gb[r][c] = lds[(kLDS_GB + 3 * r + c) * KF_BLOCK];
        }
        jacInitVJP(dEnd, Ab, Gb, dPre, Rb);
      }
      // material on this plane, recomputed from the filtered state
      const double pF = fabs(1.0 / P.xf[4]);
      double dF[3];
      {
        const SinCos s3 = oolSinCos(P.xf[3]), s2 = oolSinCos(P.xf[2]);
        const double ct = s3.c, stt = s3.s;
        const double v0 = s2.c * stt, v1 = s2.s * stt, v2 = ct;
        const double vn = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
        dF[0] = v0 / vn;
        dF[1] = v1 / vn;
        dF[2] = v2 / vn;
      }
      MaterialFwd Mt;
      Mt.run(a, pl, pF, dF);
      // adjoint of the momentum after the plane from that of q/p
      const double pAfterb = -qb * tf.q / (Mt.pOut * Mt.pOut);
      double dbm[3] = {db[0], db[1], db[2]}, nbm[3] = {0, 0, 0};
      const double pFb = Mt.vjp(a, pl, dF, pAfterb, lds[(kLDS_CB + 12) * KF_BLOCK],
                                lds[(kLDS_CB + 18) * KF_BLOCK], lds[(kLDS_CB + 24) * KF_BLOCK],
                                dbm, nbm);
#pragma unroll
      for (int m = 0; m < 3; ++m) Rb[m][2] += nbm[m];
      double xfb[5] = {0, 0, 0, 0, 0};
      freeFromBoundVJP(P, pb, dbm, pFb, xfb, cb, Rb);
      double posb[3], dbb[3], Abb[3][5], Gbb[3][3], k4b[3], qbb = 0.0, Cmb[5][5], cb2[3],
          Rb2[3][3];
      {
        double Cb[5][5];
#pragma unroll
        for (int r = 0; r < 5; ++r)
#pragma unroll
          for (int c = 0; c < 5; ++c) Cb[r][c] = lds[(kLDS_CB + 5 * r + c) * KF_BLOCK];
        P.vjp(dEnd, posEnd, AEnd, k4, xfb, Cb, posb, dbb, Abb, Gbb, k4b, qbb, Cmb, cb2, Rb2);
      }
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        cb[m] += cb2[m];
        dbb[m] += dPre[m];
#pragma unroll
        for (int c = 0; c < 3; ++c) Rb[m][c] += Rb2[m][c];
      }
      addPlane(plane, cb, Rb);
      // the covariance adjoint waits in LDS while the steps are reversed; the
      // Jacobian adjoints are kept there during the loop
#pragma unroll
      for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int c = 0; c < 5; ++c) lds[(kLDS_CB + 5 * r + c) * KF_BLOCK] = Cmb[r][c];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 5; ++c) lds[(kLDS_AB + 5 * r + c) * KF_BLOCK] = Abb[r][c];
#pragma unroll
        for (int c = 0; c < 3; ++c) lds[(kLDS_GB + 3 * r + c) * KF_BLOCK] = Gbb[r][c];
      }
      // ---- RK steps in reverse
      for (int j = nst - 1; j >= 0; --j) {
        double pin[3], din[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          pin[m] = *sr(j, kSR_POS + m);
          din[m] = *sr(j, kSR_DIR + m);
        }
        const double h = *sr(j, kSR_H);
        const int tgt = (int)*sr(j, kSR_TGT) - 1;
        StepFwd F;
        F.run(a.field, pin, din, h, qop, ldf);
        double hb = 0.0;
        stepVJP(a.field, F, pin, din, h, qop, sr(j, kSR_G), 64u, posb, dbb, lds, k4b, hb, qbb,
                sb, ldf);
#pragma unroll
        for (int m = 0; m < 3; ++m) k4b[m] = 0.0;
        if (tgt >= 0) {
          // h = n . (c - pos) / (n . d) towards plane tgt
          const DevPlane &tp = a.planes[tgt];
          const double n[3] = {tp.r[2], tp.r[5], tp.r[8]};
          const double den = n[0] * din[0] + n[1] * din[1] + n[2] * din[2];
          double tcb[3], tRb[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            const double g1 = hb * n[m] / den;
            posb[m] -= g1;
            dbb[m] -= g1 * h;
            tcb[m] = g1;
            tRb[m][2] = hb * ((tp.c[m] - pin[m]) - h * din[m]) / den;
          }
          addPlane(tgt, tcb, tRb);
        }
      }
      // ---- adjoint of the state after the update at the previous surface
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        pb[m] = posb[m];
        db[m] = dbb[m];
      }
      qb = qbb;
    }
    if (st) {
      for (int q = 0; q < slots; ++q)
        for (int f = 0; f < kPG_N; ++f) *pg(q, f) = 0.0;
      sb = 0.0;
    }
    g.scale_adj[loc] = sb;
    if (g.status) g.status[i] = st;
  }
};

}  // namespace kfd
