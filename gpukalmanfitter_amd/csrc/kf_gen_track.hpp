// kf_gen_track.hpp -- the per-track simulation of the synthetic workload
// generator, shared by the host library (kf_generate.cpp, OpenMP over tracks)
// and the device kernels (kf_generate_gpu.hip, one thread per track).
//
// The including translation unit defines KFG_HD: empty for the host build,
// `__host__ __device__` for the HIP build.
//
// Particle generation, truth propagation with an adaptive RK4 (Nystroem)
// integrator, optional material effects (sampled Highland scattering and Bethe
// energy loss with Gaussian straggling), hit smearing and start-parameter
// smearing.  It restates the shape of the reference harness
// (Run/Common/Processor.hpp:40-253,
// include/ActsExamples/ParametricParticleGenerator.hpp:86-140) with a
// counter-based RNG so that a sample does not depend on thread count or
// platform: every random number is Philox4x32-10 (Salmon et al., SC'11,
// "Parallel random numbers: as easy as 1, 2, 3") of (seed, track, stream,
// draw).  Normal deviates use the Box-Muller transform.  The integer random
// streams are identical on host and device; the floating-point results agree
// to the rounding of the two math libraries.
#pragma once

#include <math.h>
#include <stdint.h>

#include "kf_generate.h"

namespace kfg {

KFG_HD inline double dmin(double a, double b) { return (b < a) ? b : a; }
KFG_HD inline double dmax(double a, double b) { return (a < b) ? b : a; }

// ------------------------------------------------------------ Philox4x32 ----
struct Philox {
  uint32_t k0, k1;
  KFG_HD static inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
    const uint64_t p = (uint64_t)a * b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
  }
  // 10 rounds, multipliers 0xD2511F53 / 0xCD9E8D57, Weyl keys 0x9E3779B9 / 0xBB67AE85
  KFG_HD void gen(uint32_t c[4]) const {
    uint32_t a0 = k0, a1 = k1;
    for (int r = 0; r < 10; ++r) {
      uint32_t h0, l0, h1, l1;
      mulhilo(0xD2511F53u, c[0], h0, l0);
      mulhilo(0xCD9E8D57u, c[2], h1, l1);
      const uint32_t n0 = h1 ^ c[1] ^ a0;
      const uint32_t n1 = l1;
      const uint32_t n2 = h0 ^ c[3] ^ a1;
      const uint32_t n3 = l0;
      c[0] = n0;
      c[1] = n1;
      c[2] = n2;
      c[3] = n3;
      a0 += 0x9E3779B9u;
      a1 += 0xBB67AE85u;
    }
  }
};

// Per-track random stream: uniform doubles in (0, 1) and N(0, 1).
struct TrackRng {
  Philox ph;
  uint32_t track, stream;
  uint32_t draw;
  uint32_t buf[4];
  int left;
  KFG_HD TrackRng(const Philox &p, uint32_t t, uint32_t s)
      : ph(p), track(t), stream(s), draw(0), left(0) {}
  KFG_HD uint32_t next32() {
    if (left == 0) {
      buf[0] = draw++;
      buf[1] = stream;
      buf[2] = track;
      buf[3] = 0x6B66u;  // domain tag
      ph.gen(buf);
      left = 4;
    }
    return buf[4 - (left--)];
  }
  KFG_HD double uniform() {  // 53-bit, strictly inside (0, 1)
    const uint64_t a = next32() >> 5, b = next32() >> 6;
    return ((double)(a * 67108864ull + b) + 0.5) / 9007199254740992.0;
  }
  KFG_HD double normal() {
    const double u1 = uniform(), u2 = uniform();
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  // Standard Landau variate (density (1/pi) int_0^inf exp(-t ln t - x t)
  // sin(pi t) dt, the convention of CERNLIB's RANLAN): the totally skewed
  // stable law of index 1 by the method of Chambers, Mallows and Stuck
  // (J. Amer. Statist. Assoc. 71 (1976) 340): V uniform in (-pi/2, pi/2),
  // W exponential, X = (2/pi) [(pi/2 + V) tan V - ln((pi/2) W cos V / (pi/2 + V))].
  KFG_HD double landau() {
    const double hp = 1.5707963267948966;
    const double V = 3.141592653589793 * (uniform() - 0.5);
    const double W = -log(uniform());
    return ((hp + V) * tan(V) - log(hp * W * cos(V) / (hp + V))) / hp;
  }
  // Gamma(k, 1) variate, k > 0: the squeeze-rejection method of Marsaglia
  // and Tsang (ACM TOMS 26 (2000) 363) for shape k + 1 when k < 1, times
  // U^(1/k).  Each trial accepts with probability > 0.95; after 64 trials
  // the last one is kept, so the loop always ends.
  KFG_HD double gamma(double k) {
    const double a = (k < 1.0) ? k + 1.0 : k;
    const double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    double g = d;
    for (int trial = 0; trial < 64; ++trial) {
      const double x = normal();
      double v = 1.0 + c * x;
      if (v <= 0.0) continue;
      v = v * v * v;
      g = d * v;
      const double u = uniform();
      const double x2 = x * x;
      if (u < 1.0 - 0.0331 * (x2 * x2) || log(u) < 0.5 * x2 + d * (1.0 - v + log(v))) break;
    }
    if (k < 1.0) g *= exp(log(uniform()) / k);
    return g;
  }
};

// ------------------------------------------------------------------ field ----
struct GenField {
  int grid;
  double b[3];
  const float *v;
  int n[3];
  double lo[3], hi[3];
  double scale;
  KFG_HD void get(const double p[3], double out[3]) const {
    if (!grid) {
      for (int k = 0; k < 3; ++k) out[k] = scale * b[k];
      return;
    }
    int idx[3];
    double f[3];
    for (int d = 0; d < 3; ++d) {
      const double x = dmin(dmax(p[d], lo[d]), hi[d]);
      const double w = (hi[d] - lo[d]) / (n[d] - 1);
      int i = (int)floor((x - lo[d]) / w);
      i = (i < 0) ? 0 : ((i > n[d] - 2) ? n[d] - 2 : i);
      idx[d] = i;
      f[d] = (x - (lo[d] + i * w)) / w;
    }
    double acc[3] = {0, 0, 0};
    for (int c = 0; c < 8; ++c) {
      const int dx = (c >> 2) & 1, dy = (c >> 1) & 1, dz = c & 1;
      const double wgt = (dx ? f[0] : 1 - f[0]) * (dy ? f[1] : 1 - f[1]) * (dz ? f[2] : 1 - f[2]);
      const size_t o = (((size_t)(idx[0] + dx) * n[1] + idx[1] + dy) * n[2] + idx[2] + dz) * 3;
      for (int k = 0; k < 3; ++k) acc[k] += wgt * v[o + k];
    }
    for (int k = 0; k < 3; ++k) out[k] = scale * acc[k];
  }
};

KFG_HD inline void cross3(const double a[3], const double b[3], double r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}
KFG_HD inline double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// One RK4-Nystroem step of length h for d(dir)/ds = qop * dir x B.
// Returns the embedded error estimate h^2 |k1 - k2 - k3 + k4|_1.
KFG_HD inline double rkStep(const GenField &fld, double qop, double pos[3], double dir[3],
                            double h) {
  double B[3], k1[3], k2[3], k3[3], k4[3], t[3], p1[3];
  fld.get(pos, B);
  cross3(dir, B, k1);
  for (int i = 0; i < 3; ++i) k1[i] *= qop;
  for (int i = 0; i < 3; ++i) p1[i] = pos[i] + 0.5 * h * dir[i] + 0.125 * h * h * k1[i];
  fld.get(p1, B);
  for (int i = 0; i < 3; ++i) t[i] = dir[i] + 0.5 * h * k1[i];
  cross3(t, B, k2);
  for (int i = 0; i < 3; ++i) k2[i] *= qop;
  for (int i = 0; i < 3; ++i) t[i] = dir[i] + 0.5 * h * k2[i];
  cross3(t, B, k3);
  for (int i = 0; i < 3; ++i) k3[i] *= qop;
  for (int i = 0; i < 3; ++i) p1[i] = pos[i] + h * dir[i] + 0.5 * h * h * k3[i];
  fld.get(p1, B);
  for (int i = 0; i < 3; ++i) t[i] = dir[i] + h * k3[i];
  cross3(t, B, k4);
  for (int i = 0; i < 3; ++i) k4[i] *= qop;
  double err = 0;
  for (int i = 0; i < 3; ++i) err += fabs(k1[i] - k2[i] - k3[i] + k4[i]);
  for (int i = 0; i < 3; ++i) {
    pos[i] += h * dir[i] + h * h / 6.0 * (k1[i] + k2[i] + k3[i]);
    dir[i] += h / 6.0 * (k1[i] + 2.0 * (k2[i] + k3[i]) + k4[i]);
  }
  const double dn = sqrt(dot3(dir, dir));
  for (int i = 0; i < 3; ++i) dir[i] /= dn;
  return h * h * err;
}

// Propagates to the plane (centre c, normal n).  Returns false if the plane is
// not reached (the track turns away first).  Path length is added to *path.
KFG_HD inline bool propagateToPlane(const GenField &fld, double qop, double pos[3],
                                    double dir[3], const double c[3], const double n[3],
                                    double *path) {
  const double hmax = 20.0;
  for (int it = 0; it < 100000; ++it) {
    const double denom = dot3(dir, n);
    const double d[3] = {c[0] - pos[0], c[1] - pos[1], c[2] - pos[2]};
    const double dist = dot3(n, d);  // signed distance along the normal
    if (fabs(dist) < 1e-7) return true;
    if (denom <= 1e-3 && dist > 0) return false;  // moving away or grazing
    const double s = dist / denom;
    if (s < -1.0) return false;
    double h = dmax(-hmax, dmin(s, hmax));
    // step-doubling error control
    for (int tries = 0; tries < 60; ++tries) {
      double p2[3] = {pos[0], pos[1], pos[2]}, d2[3] = {dir[0], dir[1], dir[2]};
      const double err = rkStep(fld, qop, p2, d2, h);
      if (err <= 1e-7 || fabs(h) < 1e-3) {
        for (int i = 0; i < 3; ++i) {
          pos[i] = p2[i];
          dir[i] = d2[i];
        }
        *path += h;
        break;
      }
      h *= 0.5;
    }
  }
  return false;
}

// RPP formulas for the material simulation, as the reference's simulation
// (ActsFatras BetheBloch and Highland): the energy loss is drawn from a
// Landau distribution whose location is the most probable loss
// (computeEnergyLossLandau, Interactions.ipp:220-240, RPP eq. 33.11) and
// whose scale is the Gaussian-equivalent width of the Landau FWHM 4 eps
// (computeEnergyLossLandauSigma, :298-314); Highland scattering width.  Also
// the mean Bethe loss the fit subtracts (for the fit-model simulation).
struct GenMaterial {
  double thickness, x0, ar, z, molarRho;
};
KFG_HD inline void materialEffects(const GenMaterial &m, double mass, double p, double pathCorr,
                                   double *mpvLoss, double *meanLoss, double *sigmaE,
                                   double *theta0) {
  const double kMe = 0.5109989461e-3;
  const double kK = 0.307075e-3 * 100.0;
  const double x = m.thickness * pathCorr;
  const double q2OverBeta2 = 1.0 + (mass / p) * (mass / p);
  const double beta2 = 1.0 / q2OverBeta2;
  const double bg = p / mass;
  const double gamma = sqrt(1.0 + bg * bg);
  const double Ne = m.z * m.molarRho;
  const double I = 16e-9 * pow(m.z, 0.9);
  const double eps = 0.5 * kK * Ne * x * q2OverBeta2;
  const double u = 2 * kMe * bg * bg;
  const double mfrac = kMe / mass;
  const double wmax = 2 * kMe * bg * bg / (1.0 + 2 * gamma * mfrac + mfrac * mfrac);
  double dhalf = 0;
  if (bg >= 10.0) dhalf = log(bg) + log(28.816e-9 * sqrt(Ne) / I) - 0.5;
  // the fit's mean loss: RPP eq. 34.5 over the slab, 2 eps [ln(u Wmax / I^2)/2
  // - beta^2 - delta/2], with the reference's prefactor eps instead of 2 eps
  // (Interactions.ipp:150-176)
  *meanLoss = eps * (0.5 * log(u * wmax / (I * I)) - beta2 - dhalf);
  // the reference's most probable value takes the particle's mass in its
  // mass term (Interactions.ipp:235; RPP eq. 33.11 has the electron's)
  const double tMass = 2 * mass * bg * bg;
  *mpvLoss = eps * (log(tMass / I) + log(eps / I) + 0.2 - beta2 - 2 * dhalf);
  *sigmaE = 4 * eps / (2 * sqrt(2 * log(2.0)));
  const double t = sqrt(x / m.x0 * q2OverBeta2);
  *theta0 = 13.6e-3 / p * t * (1.0 + 0.038 * 2 * log(t));
}

// Outputs of the simulation: per-track SoA arrays ([component][n]) and the
// dense hit scratch [n][max_planes] (u, v, geo id); truth may be NULL.
struct SimOut {
  double *start_params, *start_cov, *perigee;
  int32_t *seq_first, *seq_count, *n_hits;
  double *scratch_u, *scratch_v;
  int32_t *scratch_geo;
  double *truth;
};

// Simulates track `it` through the (true) planes.
KFG_HD inline void simulateTrack(const kfg_config &cfg, const kf_plane *true_planes,
                                 const GenField &fld, const Philox &ph, int64_t it,
                                 const SimOut &o) {
  const int64_t n = cfg.n_tracks;
  const int maxp = cfg.max_planes;
  TrackRng r(ph, (uint32_t)it, 1u);
  const double q = (r.uniform() < 0.5) ? -1.0 : 1.0;
  const double p = cfg.p_min + (cfg.p_max - cfg.p_min) * r.uniform();
  const double phi = cfg.phi_min + (cfg.phi_max - cfg.phi_min) * r.uniform();
  const double eta = cfg.eta_min + (cfg.eta_max - cfg.eta_min) * r.uniform();
  const double theta = 2.0 * atan(exp(-eta));
  double vtx[3];
  for (int k = 0; k < 3; ++k) vtx[k] = cfg.vertex_sigma * r.normal();
  int cnt = cfg.min_planes;
  if (cfg.max_planes > cfg.min_planes)
    cnt += (int)(r.uniform() * (cfg.max_planes - cfg.min_planes + 1));
  const int cap = (cfg.max_planes < cfg.n_planes) ? cfg.max_planes : cfg.n_planes;
  cnt = (cnt < cap) ? cnt : cap;
  o.seq_first[it] = 0;
  o.seq_count[it] = cnt;
  for (int k = 0; k < 3; ++k) o.perigee[k * n + it] = vtx[k];
  if (o.truth) {
    const double tv[6] = {0.0, 0.0, phi, theta, q / p, 0.0};
    for (int k = 0; k < 6; ++k) o.truth[k * n + it] = tv[k];
  }

  // truth propagation through the (true) planes
  TrackRng rs(ph, (uint32_t)it, 2u);
  double pos[3] = {vtx[0], vtx[1], vtx[2]};
  double dir[3] = {cos(phi) * sin(theta), sin(phi) * sin(theta), cos(theta)};
  double pp = p, path = 0;
  int nh = 0;
  for (int k = 0; k < cnt; ++k) {
    const kf_plane &pl = true_planes[k];
    const double nrm[3] = {pl.rot[2], pl.rot[5], pl.rot[8]};
    if (!propagateToPlane(fld, q / pp, pos, dir, pl.center, nrm, &path)) break;
    // local coordinates of the crossing
    const double d[3] = {pos[0] - pl.center[0], pos[1] - pl.center[1], pos[2] - pl.center[2]};
    const double lu = pl.rot[0] * d[0] + pl.rot[3] * d[1] + pl.rot[6] * d[2];
    const double lv = pl.rot[1] * d[0] + pl.rot[4] * d[1] + pl.rot[7] * d[2];
    const double su = cfg.hit_sigma_u * rs.normal();
    const double sv = cfg.hit_sigma_v * rs.normal();
    const bool drop = (cfg.hole_prob > 0.0) && (rs.uniform() < cfg.hole_prob);
    if (!drop) {
      o.scratch_u[it * maxp + nh] = lu + su;
      o.scratch_v[it * maxp + nh] = lv + sv;
      o.scratch_geo[it * maxp + nh] = pl.geo_id;
      ++nh;
    }
    if ((cfg.material == KFG_MATERIAL_FIT_MODEL || cfg.material == KFG_MATERIAL_FIT_MODEL_ELOSS) &&
        pl.thickness > 0.0) {
      // the fit's own material model (a pull test of the material update):
      // Gaussian projected scattering angles, theta by theta0 and phi by
      // theta0 / sin(theta), and a Gaussian loss of the mean Bethe value and
      // the Landau-FWHM width, both for the slab along the path
      const GenMaterial m{pl.thickness, pl.x0, pl.ar, pl.z, pl.molar_rho};
      const double pc = 1.0 / fabs(dot3(nrm, dir));
      double mpvLoss, meanLoss, sigmaE, theta0;
      materialEffects(m, cfg.mass, pp, pc, &mpvLoss, &meanLoss, &sigmaE, &theta0);
      if (cfg.material == KFG_MATERIAL_FIT_MODEL) {
        const double st = sqrt(dir[0] * dir[0] + dir[1] * dir[1]);
        const double th = atan2(st, dir[2]) + theta0 * rs.normal();
        const double ph = atan2(dir[1], dir[0]) + theta0 / st * rs.normal();
        dir[0] = sin(th) * cos(ph);
        dir[1] = sin(th) * sin(ph);
        dir[2] = cos(th);
      }
      const double E = sqrt(cfg.mass * cfg.mass + pp * pp) - (meanLoss + sigmaE * rs.normal());
      if (E <= cfg.mass * 1.0001) break;  // stopped
      pp = sqrt(E * E - cfg.mass * cfg.mass);
    } else if (cfg.material && pl.thickness > 0.0) {
      // the reference's simulation (ActsFatras MinimalSimulator): Highland
      // scattering, then Bethe-Bloch (Landau) loss, then Bethe-Heitler loss,
      // with the slab thickness scaled by the incidence
      const GenMaterial m{pl.thickness, pl.x0, pl.ar, pl.z, pl.molar_rho};
      const double pc = 1.0 / fabs(dot3(nrm, dir));
      double mpvLoss, meanLoss, sigmaE, theta0;
      materialEffects(m, cfg.mass, pp, pc, &mpvLoss, &meanLoss, &sigmaE, &theta0);
      // scattering (Scattering.hpp, Highland.hpp): a deflection by theta ~
      // N(0, sqrt(2) theta0) towards a uniformly oriented direction
      // perpendicular to the track
      {
        const double psi = 6.283185307179586 * (rs.uniform() - 0.5);
        const double th = 1.4142135623730951 * theta0 * rs.normal();
        double a[3] = {0, 0, 1};
        if (fabs(dir[2]) > 0.9) a[0] = 1, a[2] = 0;
        double e1[3], e2[3];
        cross3(dir, a, e1);
        const double l1 = sqrt(dot3(e1, e1));
        for (int i = 0; i < 3; ++i) e1[i] /= l1;
        cross3(dir, e1, e2);
        const double ct = cos(th), st = sin(th), cp = cos(psi), sn = sin(psi);
        for (int i = 0; i < 3; ++i) dir[i] = ct * dir[i] + st * (cp * e1[i] + sn * e2[i]);
        const double dn = sqrt(dot3(dir, dir));
        for (int i = 0; i < 3; ++i) dir[i] /= dn;
      }
      // Bethe-Bloch (BetheBloch.hpp): Landau(most probable loss, width)
      double loss = mpvLoss + sigmaE * rs.landau();
      if (loss < 0) loss = 0;  // the Landau's left tail, rare at these widths
      double E = sqrt(cfg.mass * cfg.mass + pp * pp) - loss;
      if (E <= cfg.mass * 1.0001) break;  // stopped
      // Bethe-Heitler (BetheHeitler.hpp): E z with z = exp(-u),
      // u ~ Gamma(t / ln 2) and t the path in radiation lengths
      const double u = rs.gamma(pl.thickness * pc / pl.x0 / 0.6931471805599453);
      E = E * exp(-u);
      if (E <= cfg.mass * 1.0001) break;  // stopped
      pp = sqrt(E * E - cfg.mass * cfg.mass);
    }
  }
  o.n_hits[it] = nh;

  // start parameters: truth at the perigee, smeared (Processor.hpp:177-253)
  TrackRng rp(ph, (uint32_t)it, 3u);
  const double kPi = 3.14159265358979323846;
  const double sp_ = cfg.sigma_p_rel * p;
  double sp[6];
  sp[0] = cfg.sigma_d0 * rp.normal();
  sp[1] = cfg.sigma_z0 * rp.normal();
  sp[5] = cfg.sigma_t * rp.normal();
  double nphi = phi + cfg.sigma_phi * rp.normal();
  double nth = theta + cfg.sigma_theta * rp.normal();
  if (nth < 0) nth = -nth, nphi += kPi;  // normalizePhiTheta
  if (nth > kPi) nth = 2 * kPi - nth, nphi += kPi;
  nphi = nphi - 2 * kPi * floor((nphi + kPi) / (2 * kPi));
  sp[2] = nphi;
  sp[3] = nth;
  const double np = dmax(0.0, p + sp_ * rp.normal());
  sp[4] = q / np;
  for (int k = 0; k < 6; ++k) o.start_params[k * n + it] = sp[k];
  const double sqop = sp_ / (p * p);
  const double diag[6] = {cfg.sigma_d0 * cfg.sigma_d0, cfg.sigma_z0 * cfg.sigma_z0,
                          cfg.sigma_phi * cfg.sigma_phi, cfg.sigma_theta * cfg.sigma_theta,
                          sqop * sqop, cfg.sigma_t * cfg.sigma_t};
  int kk = 0;
  for (int a = 0; a < 6; ++a)
    for (int b = a; b < 6; ++b) o.start_cov[(int64_t)(kk++) * n + it] = (a == b) ? diag[a] : 0.0;
}

// Copies track it's hits from the dense scratch into the CSR arrays.
KFG_HD inline void packTrack(const kfg_config &cfg, const int32_t *n_hits,
                             const int64_t *hit_offset, int64_t m, const double *scratch_u,
                             const double *scratch_v, const int32_t *scratch_geo, double sigma_u,
                             double sigma_v, int32_t *hit_geo, double *hit_meas, double *hit_cov,
                             int64_t it) {
  const int maxp = cfg.max_planes;
  for (int k = 0; k < n_hits[it]; ++k) {
    const int64_t h = hit_offset[it] + k;
    hit_geo[h] = scratch_geo[it * maxp + k];
    hit_meas[h] = scratch_u[it * maxp + k];
    hit_meas[m + h] = scratch_v[it * maxp + k];
    hit_cov[h] = sigma_u * sigma_u;
    hit_cov[m + h] = 0.0;
    hit_cov[2 * m + h] = sigma_v * sigma_v;
  }
}

}  // namespace kfg
