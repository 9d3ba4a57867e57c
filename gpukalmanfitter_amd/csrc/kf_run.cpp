// kf_run.cpp -- command-line fit driver (the counterpart of the reference's
// Run/GPU/KalmanFitterGPUTest.cu, same option letters, :36-54).
//
// Generates a seeded workload with libkfgen.so, copies it to the GPU, fits it
// with the C++ facade (include/kfmi.hpp) over the C ABI, and prints the fit
// time; -o 1 writes the fitted parameters of the successful tracks in the
// reference's CSV format (Run/Common/Writer.hpp:80-110).
//
//   -t,--tracks N          number of tracks (default 10000)
//   -e,--streams N         split the batch over N HIP streams (one fitter each)
//   -r,--threads N         host threads of the generator (0 = all)
//   -u,--multiple-devices  1 is refused: this build runs on one GPU
//   -o,--output 0|1        write params-fitted_<config>_<N>.csv
//   -d,--device gpu        the fit runs on the GPU ("cpu" is refused: the CPU
//                          restatement of the fit is test infrastructure)
//   -g,--grid-size, -b,--block-size
//                          accepted for compatibility; the launch shape is
//                          one track per lane, 128-lane blocks
//   -s,--shared-memory 1   the counterpart of the reference's block-per-track
//                          kernel: each track on a group of 4 lanes
//                          (kf_schedule.lanes_per_track = 4)
//   -m,--smoothing 0|1     run the smoother (default 1)
//   -a,--machine NAME      label of the timing line
// Options of this build:
//   -c,--config 1|2|3|5    workload (detectors/fields of the benchmark configs)
//   -p,--precision fp64|mixed
//   -n,--repeat N          timed fits (default 10, the median is reported)
//   --reference-overstep   keep the reference's 0.01 mm overstep limit
//   --with-copies          time each fit with its inputs copied in from pinned
//                          host memory and its results copied out, every
//                          stream's chunk in its own stream (the reference
//                          driver's timing also covers the transfers)
//   --per-track K          also fit the first K tracks one at a time through
//                          kfmi::KalmanFitter::fit (the reference's per-track
//                          call) and report their largest difference from
//                          the batch fit
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/kfmi.hpp"
#include "kf_generate.h"

namespace {

constexpr double kTesla = 0.000299792458;
constexpr double kNs = 1e-9 * 299792458000.0;

// Generator settings of the benchmark configs (the Python workloads,
// gpukalmanfitter_amd/configs.py; tests/test_cpp_facade.py checks they agree).
bool preset(int id, kfg_config &c, bool &grid) {
  std::memset(&c, 0, sizeof c);
  c.n_tracks = 10000;
  c.seed = 1234567890;
  c.n_planes = 10;
  c.min_planes = 10;
  c.max_planes = 10;
  c.material = 0;
  c.first_x = 30.0;
  c.pitch = 30.0;
  c.half_u = 500.0;
  c.half_v = 500.0;
  c.thickness = 0.0;
  c.hole_prob = 0.0;
  c.p_min = 1.0;
  c.p_max = 10.0;
  c.phi_min = -0.3;
  c.phi_max = 0.3;
  c.eta_min = -0.3;
  c.eta_max = 0.3;
  c.vertex_sigma = 0.010;
  c.hit_sigma_u = 0.050;
  c.hit_sigma_v = 0.050;
  c.sigma_d0 = 0.100;
  c.sigma_z0 = 0.100;
  c.sigma_phi = 1e-3;
  c.sigma_theta = 1e-3;
  c.sigma_p_rel = 0.05;
  c.sigma_t = 5.0 * kNs;
  c.mass = 105.6583755 * 1e-3;
  c.true_field_scale = 1.0;
  grid = false;
  switch (id) {
    case 1:
      c.seed = 1234567891;
      return true;
    case 2:
      c.n_tracks = 1048576;
      c.seed = 1234567892;
      return true;
    case 3:
      c.n_tracks = 2000000;
      c.seed = 1234567893;
      c.n_planes = 30;
      c.min_planes = 5;
      c.max_planes = 30;
      c.first_x = 25.0;
      c.pitch = 25.0;
      c.hole_prob = 0.05;
      grid = true;
      return true;
    case 5:
      c.n_tracks = 4194304;
      c.seed = 1234567895;
      c.n_planes = 40;
      c.min_planes = 8;
      c.max_planes = 40;
      c.first_x = 25.0;
      c.pitch = 25.0;
      c.material = 1;
      c.thickness = 0.300;
      c.p_min = 0.5;
      c.p_max = 20.0;
      grid = true;
      return true;
    default:
      return false;
  }
}

void usage() {
  std::fprintf(stderr,
               "Usage: kf_run [options]\n"
               "\t-t,--tracks N\n\t-e,--streams N\n\t-r,--threads N\n"
               "\t-u,--multiple-devices 0\n\t-o,--output 0|1\n\t-d,--device gpu\n"
               "\t-g,--grid-size X*Y\n\t-b,--block-size X*Y*Z\n\t-s,--shared-memory 0|1\n"
               "\t-m,--smoothing 0|1\n\t-a,--machine NAME\n\t-c,--config 1|2|3|5\n"
               "\t-p,--precision fp64|mixed\n\t-n,--repeat N\n\t--reference-overstep\n"
               "\t--with-copies\n\t--per-track K\n"
               "\t--print-config (prints the generator settings and exits)\n");
}

template <typename T>
T *devAlloc(size_t n) {
  void *p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) {
    std::fprintf(stderr, "ERROR: hipMalloc of %zu bytes failed\n", n * sizeof(T));
    std::exit(1);
  }
  return (T *)p;
}
template <typename T>
void toDev(T *d, const std::vector<T> &h) {
  if (!h.empty() && hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
    std::fprintf(stderr, "ERROR: hipMemcpy to the device failed\n");
    std::exit(1);
  }
}
template <typename T>
void toHost(std::vector<T> &h, const T *d) {
  if (!h.empty() && hipMemcpy(h.data(), d, h.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) {
    std::fprintf(stderr, "ERROR: hipMemcpy from the device failed\n");
    std::exit(1);
  }
}

}  // namespace

int main(int argc, char **argv) {
  int64_t nTracks = -1;
  int nStreams = 1, nThreads = 0, config = 1, repeat = 10, lanesPerTrack = 1;
  int64_t perTrack = 0;
  bool output = false, smoothing = true, refOverstep = false, printConfig = false;
  bool withCopies = false;
  std::string device = "gpu", machine, precision = "fp64";
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](const char *what) -> std::string {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "ERROR: %s needs a value\n", what);
        std::exit(1);
      }
      return argv[++i];
    };
    if (a == "-h" || a == "--help") {
      usage();
      return 0;
    } else if (a == "-t" || a == "--tracks") {
      nTracks = std::atoll(next("--tracks").c_str());
    } else if (a == "-e" || a == "--streams") {
      nStreams = std::atoi(next("--streams").c_str());
    } else if (a == "-r" || a == "--threads") {
      nThreads = std::atoi(next("--threads").c_str());
    } else if (a == "-u" || a == "--multiple-devices") {
      if (std::atoi(next("--multiple-devices").c_str()) == 1) {
        std::fprintf(stderr, "ERROR: this build runs on one GPU (--multiple-devices 1)\n");
        return 1;
      }
    } else if (a == "-o" || a == "--output") {
      output = std::atoi(next("--output").c_str()) == 1;
    } else if (a == "-d" || a == "--device") {
      device = next("--device");
    } else if (a == "-g" || a == "--grid-size" || a == "-b" || a == "--block-size") {
      const std::string v = next("--grid-size/--block-size");
      std::printf("INFO: %s %s ignored: one track per lane, 128-lane blocks\n", a.c_str(), v.c_str());
    } else if (a == "-s" || a == "--shared-memory") {
      lanesPerTrack = std::atoi(next("--shared-memory").c_str()) == 1 ? 4 : 1;
    } else if (a == "-m" || a == "--smoothing") {
      smoothing = std::atoi(next("--smoothing").c_str()) == 1;
    } else if (a == "-a" || a == "--machine") {
      machine = next("--machine");
    } else if (a == "-c" || a == "--config") {
      config = std::atoi(next("--config").c_str());
    } else if (a == "-p" || a == "--precision") {
      precision = next("--precision");
    } else if (a == "-n" || a == "--repeat") {
      repeat = std::max(1, std::atoi(next("--repeat").c_str()));
    } else if (a == "--reference-overstep") {
      refOverstep = true;
    } else if (a == "--with-copies") {
      withCopies = true;
    } else if (a == "--per-track") {
      perTrack = std::atoll(next("--per-track").c_str());
    } else if (a == "--print-config") {
      printConfig = true;
    } else {
      std::fprintf(stderr, "ERROR: unknown argument %s\n", a.c_str());
      usage();
      return 1;
    }
  }
  if (device != "gpu") {
    std::fprintf(stderr, "ERROR: --device %s: this build fits on the GPU only\n", device.c_str());
    return 1;
  }
  if (precision != "fp64" && precision != "mixed") {
    std::fprintf(stderr, "ERROR: --precision must be fp64 or mixed\n");
    return 1;
  }
  kfg_config cfg;
  bool grid = false;
  if (!preset(config, cfg, grid)) {
    std::fprintf(stderr, "ERROR: --config %d: choose 1, 2, 3 or 5\n", config);
    return 1;
  }
  if (nTracks > 0) cfg.n_tracks = nTracks;
  if (printConfig) {
    std::printf("{\"n_tracks\": %lld, \"seed\": %llu, \"n_planes\": %d, \"min_planes\": %d, "
                "\"max_planes\": %d, \"material\": %d, \"first_x\": %.17g, \"pitch\": %.17g, "
                "\"half_u\": %.17g, \"half_v\": %.17g, \"thickness\": %.17g, \"hole_prob\": %.17g, "
                "\"p_min\": %.17g, \"p_max\": %.17g, \"phi_min\": %.17g, \"phi_max\": %.17g, "
                "\"eta_min\": %.17g, \"eta_max\": %.17g, \"vertex_sigma\": %.17g, "
                "\"hit_sigma_u\": %.17g, \"hit_sigma_v\": %.17g, \"sigma_d0\": %.17g, "
                "\"sigma_z0\": %.17g, \"sigma_phi\": %.17g, \"sigma_theta\": %.17g, "
                "\"sigma_p_rel\": %.17g, \"sigma_t\": %.17g, \"mass\": %.17g, "
                "\"align_shift_sigma\": %.17g, \"align_rot_sigma\": %.17g, "
                "\"true_field_scale\": %.17g, \"grid\": %d}\n",
                (long long)cfg.n_tracks, (unsigned long long)cfg.seed, cfg.n_planes,
                cfg.min_planes, cfg.max_planes, cfg.material, cfg.first_x, cfg.pitch, cfg.half_u,
                cfg.half_v, cfg.thickness, cfg.hole_prob, cfg.p_min, cfg.p_max, cfg.phi_min,
                cfg.phi_max, cfg.eta_min, cfg.eta_max, cfg.vertex_sigma, cfg.hit_sigma_u,
                cfg.hit_sigma_v, cfg.sigma_d0, cfg.sigma_z0, cfg.sigma_phi, cfg.sigma_theta,
                cfg.sigma_p_rel, cfg.sigma_t, cfg.mass, cfg.align_shift_sigma,
                cfg.align_rot_sigma, cfg.true_field_scale, grid ? 1 : 0);
    return 0;
  }
  const int64_t n = cfg.n_tracks;
  if (nStreams < 1 || nStreams > n) nStreams = 1;

  // ---- workload (host) ----
  std::vector<kf_plane> planes(cfg.n_planes);
  kfg_planes(&cfg, planes.data());
  std::vector<float> gridValues;
  kfg_field fld{};
  kfmi::MagneticField field = kfmi::MagneticField::constant(0.0, 0.0, 2.0 * kTesla);
  if (grid) {
    const int32_t nx = 101, ny = 101, nz = 201;
    const double lo[3] = {-1000.0, -1000.0, -2000.0}, hi[3] = {1000.0, 1000.0, 2000.0};
    gridValues.resize((size_t)nx * ny * nz * 3);
    kfg_fill_grid(gridValues.data(), nx, ny, nz, lo, hi, 2.0 * kTesla, 0.05 * kTesla);
    fld.grid = 1;
    fld.n[0] = nx;
    fld.n[1] = ny;
    fld.n[2] = nz;
    for (int k = 0; k < 3; ++k) {
      fld.lo[k] = lo[k];
      fld.hi[k] = hi[k];
    }
    fld.values = gridValues.data();
    field = kfmi::MagneticField::interpolated(gridValues.data(), nx, ny, nz, {lo[0], lo[1], lo[2]},
                                              {hi[0], hi[1], hi[2]});
  } else {
    fld.b[2] = 2.0 * kTesla;
  }
  const int maxp = cfg.max_planes;
  std::vector<double> sp(6 * n), sc(21 * n), pg(3 * n), truth(6 * n);
  std::vector<int32_t> sf(n), scount(n), nh(n);
  std::vector<double> su((size_t)n * maxp), sv((size_t)n * maxp);
  std::vector<int32_t> sg((size_t)n * maxp);
  auto t0 = std::chrono::steady_clock::now();
  if (kfg_simulate(&cfg, planes.data(), &fld, sp.data(), sc.data(), pg.data(), sf.data(),
                   scount.data(), nh.data(), su.data(), sv.data(), sg.data(), truth.data(),
                   nThreads) != 0) {
    std::fprintf(stderr, "ERROR: kfg_simulate failed\n");
    return 1;
  }
  std::vector<int64_t> off(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + nh[i];
  const int64_t m = off[n];
  std::vector<int32_t> hg(m);
  std::vector<double> hm(2 * m), hc(3 * m);
  kfg_pack_hits(&cfg, nh.data(), off.data(), su.data(), sv.data(), sg.data(), cfg.hit_sigma_u,
                cfg.hit_sigma_v, hg.data(), hm.data(), hc.data(), nThreads);
  std::vector<double>().swap(su);
  std::vector<double>().swap(sv);
  std::vector<int32_t>().swap(sg);
  auto t1 = std::chrono::steady_clock::now();
  std::printf("INFO: %s precision, config %d, %lld tracks, %lld hits, generated in %.1f ms\n",
              precision.c_str(), config, (long long)n, (long long)m,
              std::chrono::duration<double, std::milli>(t1 - t0).count());

  // ---- device buffers (one batch; streams take contiguous track ranges) ----
  int32_t *dsf = devAlloc<int32_t>(n), *dsn = devAlloc<int32_t>(n);
  int32_t *dhg = devAlloc<int32_t>(m);
  double *dhm = devAlloc<double>(2 * m), *dhc = devAlloc<double>(3 * m);
  double *dpar = devAlloc<double>(6 * n), *dcov = devAlloc<double>(21 * n);
  double *dchi = devAlloc<double>(n);
  int32_t *dndf = devAlloc<int32_t>(n), *dnm = devAlloc<int32_t>(n);
  uint32_t *dst = devAlloc<uint32_t>(n);
  toDev(dhg, hg);
  toDev(dhm, hm);
  toDev(dhc, hc);
  toDev(dsf, sf);
  toDev(dsn, scount);

  // Per stream: its own fitter (contexts are not shared across streams), SoA
  // slices of the inputs (per-track arrays are [component][n], so each
  // stream's chunk is re-laid out as [component][chunk]) and its own outputs.
  struct Chunk {
    int64_t first, count;
    double *sp, *sc, *pg, *par, *cov, *chi;
    int64_t *off;
    int32_t *ndf, *nm;
    uint32_t *st;
    hipStream_t stream;
  };
  std::vector<Chunk> chunks(nStreams);
  std::vector<std::unique_ptr<kfmi::KalmanFitter>> fitters;
  // n / nStreams tracks per stream, one more for the first n % nStreams
  // streams: every chunk holds at least one track (nStreams <= n)
  const int64_t per = n / nStreams, extra = n % nStreams;
  for (int s = 0; s < nStreams; ++s) {
    Chunk &c = chunks[s];
    c.first = s * per + std::min<int64_t>(s, extra);
    c.count = per + (s < extra ? 1 : 0);
    auto slice = [&](const std::vector<double> &v, int comps) {
      std::vector<double> out((size_t)comps * c.count);
      for (int k = 0; k < comps; ++k)
        std::copy(v.begin() + (size_t)k * n + c.first, v.begin() + (size_t)k * n + c.first + c.count,
                  out.begin() + (size_t)k * c.count);
      return out;
    };
    c.sp = devAlloc<double>(6 * c.count);
    c.sc = devAlloc<double>(21 * c.count);
    c.pg = devAlloc<double>(3 * c.count);
    c.off = devAlloc<int64_t>(c.count + 1);
    toDev(c.sp, slice(sp, 6));
    toDev(c.sc, slice(sc, 21));
    toDev(c.pg, slice(pg, 3));
    toDev(c.off, std::vector<int64_t>(off.begin() + c.first, off.begin() + c.first + c.count + 1));
    c.par = dpar + 6 * c.first;  // chunk-local [6][count] blocks, gathered after the fit
    c.cov = dcov + 21 * c.first;
    c.chi = dchi + c.first;
    c.ndf = dndf + c.first;
    c.nm = dnm + c.first;
    c.st = dst + c.first;
    if (hipStreamCreate(&c.stream) != hipSuccess) {
      std::fprintf(stderr, "ERROR: hipStreamCreate failed\n");
      return 1;
    }
    fitters.emplace_back(new kfmi::KalmanFitter(planes, field, 0));
    if (lanesPerTrack > 1) {
      kf_schedule sched;
      kf_schedule_default(&sched);
      sched.lanes_per_track = lanesPerTrack;
      if (kf_set_schedule(fitters.back()->context(), &sched) != KF_OK) {
        std::fprintf(stderr, "ERROR: %s\n", kf_last_error(fitters.back()->context()));
        return 1;
      }
    }
  }
  kfmi::KalmanFitterOptions opt;
  opt.smoothing = smoothing;
  opt.propagator.precision = (precision == "mixed") ? KF_PRECISION_MIXED : KF_PRECISION_FP64;
  if (!refOverstep) opt.propagator.overstep_limit = 1.0;
  opt.propagator.max_states = maxp;
  // --with-copies: pinned host images of every input and output; each
  // stream's chunk copies its own ranges (per-track arrays in the chunk's
  // [component][count] blocks, its hit range of every hit component)
  struct HostChunk {
    double *sp = nullptr, *sc = nullptr, *pg = nullptr, *par = nullptr, *cov = nullptr,
           *chi = nullptr;
    int64_t *off = nullptr;
    int32_t *ndf = nullptr, *nm = nullptr;
    uint32_t *st = nullptr;
  };
  std::vector<HostChunk> hchunks(nStreams);
  int32_t *hsf = nullptr, *hsn = nullptr, *hhg = nullptr;
  double *hhm = nullptr, *hhc = nullptr;
  bool copyFailed = false;
  auto pinned = [&](auto *&ptr, size_t count, const auto *src) {
    using T = std::remove_reference_t<decltype(*ptr)>;
    void *q = nullptr;
    if (hipHostMalloc(&q, (count ? count : 1) * sizeof(T)) != hipSuccess) {
      copyFailed = true;
      return;
    }
    ptr = (T *)q;
    if (src) std::memcpy(ptr, src, count * sizeof(T));
  };
  if (withCopies) {
    pinned(hsf, (size_t)n, sf.data());
    pinned(hsn, (size_t)n, scount.data());
    pinned(hhg, (size_t)m, hg.data());
    pinned(hhm, (size_t)2 * m, hm.data());
    pinned(hhc, (size_t)3 * m, hc.data());
    for (int s = 0; s < nStreams && !copyFailed; ++s) {
      const Chunk &c = chunks[s];
      HostChunk &h = hchunks[s];
      const size_t k = (size_t)c.count;
      pinned(h.sp, 6 * k, (const double *)nullptr);
      pinned(h.sc, 21 * k, (const double *)nullptr);
      pinned(h.pg, 3 * k, (const double *)nullptr);
      pinned(h.off, k + 1, off.data() + c.first);
      pinned(h.par, 6 * k, (const double *)nullptr);
      pinned(h.cov, 21 * k, (const double *)nullptr);
      pinned(h.chi, k, (const double *)nullptr);
      pinned(h.ndf, k, (const int32_t *)nullptr);
      pinned(h.nm, k, (const int32_t *)nullptr);
      pinned(h.st, k, (const uint32_t *)nullptr);
      if (copyFailed) break;
      for (int comp = 0; comp < 6; ++comp)
        std::memcpy(h.sp + comp * k, sp.data() + (size_t)comp * n + c.first, k * sizeof(double));
      for (int comp = 0; comp < 21; ++comp)
        std::memcpy(h.sc + comp * k, sc.data() + (size_t)comp * n + c.first, k * sizeof(double));
      for (int comp = 0; comp < 3; ++comp)
        std::memcpy(h.pg + comp * k, pg.data() + (size_t)comp * n + c.first, k * sizeof(double));
    }
    if (copyFailed) {
      std::fprintf(stderr, "ERROR: hipHostMalloc failed (--with-copies)\n");
      return 1;
    }
  }
  auto copyIn = [&](int s) {
    const Chunk &c = chunks[s];
    const HostChunk &h = hchunks[s];
    const size_t k = (size_t)c.count;
    const int64_t h0 = off[c.first], nh = off[c.first + c.count] - h0;
    auto cp = [&](void *d, const void *src, size_t bytes) {
      if (bytes && hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, c.stream) != hipSuccess)
        copyFailed = true;
    };
    cp(c.sp, h.sp, 6 * k * sizeof(double));
    cp(c.sc, h.sc, 21 * k * sizeof(double));
    cp(c.pg, h.pg, 3 * k * sizeof(double));
    cp(c.off, h.off, (k + 1) * sizeof(int64_t));
    cp(dsf + c.first, hsf + c.first, k * sizeof(int32_t));
    cp(dsn + c.first, hsn + c.first, k * sizeof(int32_t));
    cp(dhg + h0, hhg + h0, nh * sizeof(int32_t));
    for (int comp = 0; comp < 2; ++comp)
      cp(dhm + comp * m + h0, hhm + comp * m + h0, nh * sizeof(double));
    for (int comp = 0; comp < 3; ++comp)
      cp(dhc + comp * m + h0, hhc + comp * m + h0, nh * sizeof(double));
  };
  auto copyOut = [&](int s) {
    const Chunk &c = chunks[s];
    const HostChunk &h = hchunks[s];
    const size_t k = (size_t)c.count;
    auto cp = [&](void *dst, const void *d, size_t bytes) {
      if (bytes && hipMemcpyAsync(dst, d, bytes, hipMemcpyDeviceToHost, c.stream) != hipSuccess)
        copyFailed = true;
    };
    cp(h.par, c.par, 6 * k * sizeof(double));
    cp(h.cov, c.cov, 21 * k * sizeof(double));
    cp(h.chi, c.chi, k * sizeof(double));
    cp(h.ndf, c.ndf, k * sizeof(int32_t));
    cp(h.nm, c.nm, k * sizeof(int32_t));
    cp(h.st, c.st, k * sizeof(uint32_t));
  };
  auto fitAll = [&]() {
    for (int s = 0; s < nStreams; ++s) {
      Chunk &c = chunks[s];
      if (withCopies) copyIn(s);
      kf_tracks t{};
      t.n_tracks = c.count;
      t.n_hits = m;  // hit arrays are shared; offsets index the full arrays
      t.start_params = c.sp;
      t.start_cov = c.sc;
      t.perigee = c.pg;
      t.seq_first = dsf + c.first;
      t.seq_count = dsn + c.first;
      t.hit_offset = c.off;
      t.hit_geo_id = dhg;
      t.hit_meas = dhm;
      t.hit_cov = dhc;
      kf_results r{};
      r.params = c.par;
      r.cov = c.cov;
      r.chi2 = c.chi;
      r.ndf = c.ndf;
      r.n_measured = c.nm;
      r.status = c.st;
      fitters[s]->fitBatch(t, opt, r, c.stream);
      if (withCopies) copyOut(s);
    }
    for (int s = 0; s < nStreams; ++s) (void)hipStreamSynchronize(chunks[s].stream);
  };
  fitAll();  // warm-up (allocates the fitters' stores)
  std::vector<double> ms;
  for (int k = 0; k < repeat; ++k) {
    auto a = std::chrono::steady_clock::now();
    fitAll();
    auto b = std::chrono::steady_clock::now();
    ms.push_back(std::chrono::duration<double, std::milli>(b - a).count());
  }
  std::sort(ms.begin(), ms.end());
  const double med = ms[ms.size() / 2];
  std::printf("INFO: Time (ms) for KF track fitting on device 0 : %f\n", med);
  std::printf("INFO: %s %.4g tracks/s (median of %d fits, %d stream(s), %d lane(s) per track%s)\n",
              machine.empty() ? "MI355X" : machine.c_str(), n / (med * 1e-3), repeat, nStreams,
              lanesPerTrack, withCopies ? ", inputs copied in and results out in the timed region" : "");
  if (copyFailed) {
    std::fprintf(stderr, "ERROR: hipMemcpyAsync failed (--with-copies)\n");
    return 1;
  }
  if (withCopies) {
    // every copied-out result array must equal the device's, element for
    // element (per chunk: [6][count] params, [21][count] cov, then chi2, ndf,
    // n_measured and status of its tracks)
    int64_t diff = 0;
    std::vector<double> parDev(6 * n), covDev(21 * n), chiDev(n);
    std::vector<int32_t> ndfDev(n), nmDev(n);
    std::vector<uint32_t> stDev(n);
    toHost(parDev, dpar);
    toHost(covDev, dcov);
    toHost(chiDev, dchi);
    toHost(ndfDev, dndf);
    toHost(nmDev, dnm);
    toHost(stDev, dst);
    auto differ = [](const void *a, const void *b, size_t bytes) {
      return std::memcmp(a, b, bytes) != 0;
    };
    for (int s = 0; s < nStreams; ++s) {
      const Chunk &c = chunks[s];
      const HostChunk &h = hchunks[s];
      for (int64_t j = 0; j < 6 * c.count; ++j)
        diff += differ(&h.par[j], &parDev[6 * c.first + j], sizeof(double));
      for (int64_t j = 0; j < 21 * c.count; ++j)
        diff += differ(&h.cov[j], &covDev[21 * c.first + j], sizeof(double));
      for (int64_t j = 0; j < c.count; ++j) {
        diff += differ(&h.chi[j], &chiDev[c.first + j], sizeof(double));
        diff += (h.ndf[j] != ndfDev[c.first + j]) + (h.nm[j] != nmDev[c.first + j]) +
                (h.st[j] != stDev[c.first + j]);
      }
    }
    std::printf("INFO: copied-out result elements differing from the device's: %lld\n",
                (long long)diff);
  }

  // ---- results ----
  std::vector<double> par(6 * n);
  std::vector<uint32_t> st(n);
  toHost(par, dpar);
  toHost(st, dst);
  int64_t ok = 0;
  for (int64_t i = 0; i < n; ++i) ok += (st[i] & KF_STATUS_OK) ? 1 : 0;
  std::printf("INFO: %lld of %lld tracks fitted\n", (long long)ok, (long long)n);
  if (perTrack > 0) {
    // the reference's per-track call on the same inputs; chunk 0 holds the
    // first tracks in [6][count] / [21][count] blocks
    std::vector<double> cov(21 * n);
    toHost(cov, dcov);
    const Chunk &c0 = chunks[0];
    double worst = 0.0;
    int64_t agree = 0;
    kfmi::KalmanFitter single(planes, field, 0);
    for (int64_t i = 0; i < std::min(perTrack, c0.count); ++i) {
      std::vector<kfmi::PixelSourceLink> sls;
      for (int64_t h = off[i]; h < off[i + 1]; ++h) {
        kfmi::PixelSourceLink sl;
        sl.values = {hm[h], hm[m + h]};
        sl.cov = {{{hc[h], hc[m + h]}, {hc[m + h], hc[2 * m + h]}}};
        sl.geometryId = hg[h];
        sls.push_back(sl);
      }
      kfmi::BoundParameters start;
      std::vector<double> c21(21);
      for (int k = 0; k < 6; ++k) start.parameters[k] = sp[(size_t)k * n + i];
      for (int k = 0; k < 21; ++k) c21[k] = sc[(size_t)k * n + i];
      start.covariance = kfmi::unpackSym(c21.data());
      kfmi::LineSurface perigee;
      for (int k = 0; k < 3; ++k) perigee.center[k] = pg[(size_t)k * n + i];
      kfmi::KalmanFitterOptions o = opt;
      o.referenceSurface = &perigee;
      kfmi::KalmanFitterResult res;
      const bool ok1 = single.fit(sls, start, o, res, single.planes().data() + sf[i], scount[i]);
      const bool okb = (st[i] & KF_STATUS_OK) != 0;
      if (ok1 == okb && res.status == st[i]) ++agree;
      for (int k = 0; k < 6; ++k)
        worst = std::max(worst, std::fabs(res.fittedParameters.parameters[k] -
                                          par[6 * c0.first + (size_t)k * c0.count + i]));
      const kfmi::SymMatrix6 bc = kfmi::unpackSym(cov.data() + 21 * c0.first + i, c0.count);
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b)
          worst = std::max(worst, std::fabs(res.fittedParameters.covariance[a][b] - bc[a][b]));
    }
    std::printf("INFO: per-track fit of %lld tracks: %lld with the batch's status, largest "
                "difference %.3g\n", (long long)std::min(perTrack, c0.count), (long long)agree,
                worst);
  }
  if (output) {
    const std::string name = "params-fitted_config" + std::to_string(config) + "_" +
                             std::to_string(n) + ".csv";
    std::ofstream f(name);
    f << "trackIdx, fit_BoundLoc0, fit_BoundLoc1, fit_BoundPhi, fit_BoundTheta, "
         "fit_BoundQOverP, fit_BoundTime\n";
    f.precision(17);
    for (int s = 0; s < nStreams; ++s) {
      const Chunk &c = chunks[s];
      for (int64_t j = 0; j < c.count; ++j) {
        const int64_t i = c.first + j;
        if (!(st[i] & KF_STATUS_OK)) continue;
        const double *p = par.data() + 6 * c.first;  // chunk block [6][count]
        f << i;
        for (int k = 0; k < 6; ++k) f << "," << p[(size_t)k * c.count + j];
        f << "\n";
      }
    }
    std::printf("INFO: wrote %s\n", name.c_str());
  }
  for (Chunk &c : chunks) {
    (void)hipStreamDestroy(c.stream);
    (void)hipFree(c.sp);
    (void)hipFree(c.sc);
    (void)hipFree(c.pg);
    (void)hipFree(c.off);
  }
  fitters.clear();
  for (void *p : {(void *)dsf, (void *)dsn, (void *)dhg,
                  (void *)dhm, (void *)dhc, (void *)dpar, (void *)dcov, (void *)dchi, (void *)dndf,
                  (void *)dnm, (void *)dst})
    (void)hipFree(p);
  return 0;
}
