#!/usr/bin/env python3
"""Throughput benchmark of the MI355X Kalman fitter (one JSON line on stdout).

Workload (config 2): 1,048,576 synthetic tracks through 10 x-normal planes at
x = 30..300 mm in a uniform 2 T field, no material, mu+- with p in [1, 10] GeV,
|phi| <= 0.3, |eta| <= 0.3; one 2-D hit per plane; start parameters smeared
(d0/z0 100 um, phi/theta 1 mrad, q/p 5 %).  A step is one kf_fit() over the
whole batch with inputs already resident in HBM; the timed region is exactly
--steps fits bracketed by device synchronisation.  Fits run in every precision
given by --precision (first one is the headline `value`).

A fit is two launches on one stream: the forward filter (start, steps,
gain-matrix update on every surface, per-surface store writes) and the
smoother plus back-propagation to the perigee (store reads).  The roofline
object is for the longer launch of the headline precision.  The path is bound
by the vector ALU: the roofline time of a launch is T_roof = F32 / 157.3 +
F64 / 78.6 TFLOP/s (the MI355X vector peaks) for its algorithmic flops, the
flops of the reference formulation with structural zeros and identities
removed, counted per track by the counting build of the CPU oracle and split
into the float and double classes of the mixed mode; achieved = F / T, peak
= F / T_roof, frac = T_roof / T.  Beside it: the launch's external I/O bytes
(inputs it reads, results it writes) against 8 TB/s, the bytes of its
working storage (smoother store, hand-over) and the HBM traffic rocprofv3
measured (FETCH_SIZE + WRITE_SIZE).  Launch durations are the mean over the
timed steps of HIP events the library records on the launch stream before,
between and after the two launches.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fitted tracks/sec on 1 GPU: FP32-mixed vs FP64 at fixed error; adjoint trk/s"
# Algorithmic flops per config-2 track (mean over 20,000 tracks of the counting
# oracle, liboracle_count.so; tests/test_bench_constants.py re-measures them):
# whole fit, the smoother/back-propagation launch's share, and the float-class
# parts of the mixed mode.
FLOPS_PER_TRACK = 30610.0
FLOPS_BWD_PER_TRACK = 17876.0
FLOPS32_PER_TRACK = 8109.0
FLOPS32_BWD_PER_TRACK = 784.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# External I/O bytes per config-2 track and launch (10 hits): what the caller
# hands in and gets back, read or written by that launch.
#   forward  = inputs 696 (start parameters 48 + covariance 168 + perigee 24
#              + sequence 8 + hit offset 8 + 10 hits x 44) + chi2/ndf/n_measured 16;
#   backward = perigee, sequence and hit offsets 48 + outputs 220 (params 48,
#              cov 168, status 4).
EXT_BYTES_PER_TRACK = {"forward": 712, "backward": 268}
# The same plus the launch's working storage (tests/test_bench_constants.py
# re-derives them from the layouts): the smoother store, 10 measured states x
# (12 x 8 + 50 x sizeof(S)), written forward and read backward, and the
# 272-byte forward-to-backward hand-over (8 int32 + 30 double).
DESIGN_BYTES_PER_TRACK = {("fp64", "forward"): 5944, ("fp64", "backward"): 5500,
                          ("mixed", "forward"): 3944, ("mixed", "backward"): 3500}
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch (fetched + written) of the two fit kernels on this
# workload, from rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel
# calibration (profiles/kf_fit_config2_1M_tracks.md).
TRAFFIC_TRACKS = 1048576
TRAFFIC_BYTES_PER_LAUNCH = {
    ("fp64", "forward"): 2.318e9 + 5.662e9, ("fp64", "backward"): 5.555e9 + 0.231e9,
    ("mixed", "forward"): 2.265e9 + 3.481e9, ("mixed", "backward"): 3.510e9 + 0.231e9,
}
PARAM_NAMES = ("loc0", "loc1", "phi", "theta", "qop", "t")


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--tracks", type=int, default=None, help="override the workload size")
    ap.add_argument("--precision", default="mixed,fp64",
                    help="comma list of mixed,fp64 (the first is the headline value)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--cpu-sample", type=int, default=1048576,
                    help="tracks of the same workload timed on the CPU oracle")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ragged", action="store_true",
                    help="skip the ragged-batch leg (config 3, 2,000,000 tracks)")
    ap.add_argument("--no-material", action="store_true",
                    help="skip the 4,194,304-track material workload leg")
    ap.add_argument("--no-intersect", action="store_true",
                    help="skip the batched ray x triangle intersection leg")
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the 10,000-track config-1 leg")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the chi2-adjoint leg (config 4, 262,144 tracks)")
    ap.add_argument("--field", choices=("grid", "flat", "const"), default="grid",
                    help="main loop on a grid-field workload (--config 3 or 5): its grid "
                         "(default), a grid of the same box holding the grid's central "
                         "value everywhere (same trajectories as a constant field: the "
                         "difference to 'const' is the cost of the lookups), or that "
                         "constant field")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer (calibrates rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE against known bytes) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB: reads and writes mib MiB of HBM."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4}), flush=True)


def cpu_baseline(batch, opts_c, threads, sample, gpu_results, others=None):
    """The CPU oracle (FP64 restatement of the reference fit, OpenMP over
    tracks as Run/CPU/KalmanFitterCPUTest.cpp) timed on a bounded sample.
    Its results on that sample also give the per-parameter error of each GPU
    precision mode (|x_gpu - x_cpu| / max(|x_cpu|, sigma_cpu), max and 99.9 %
    quantile over the fitted tracks); `others` maps a name to (batch sample,
    GPU results on it) of another workload, checked the same way (the third
    value returned)."""
    import numpy as np
    import oracle
    from gpukalmanfitter_amd import abi
    n = min(sample, batch.n_tracks)
    sub = batch.subset(np.arange(n))
    if threads <= 0:
        threads = os.cpu_count() or 1
    oracle.time_fit(sub.subset(np.arange(min(n, 2000))), opts_c, threads)  # warm
    sec = oracle.time_fit(sub, opts_c, threads)
    n1 = min(n, 50000)
    sec1 = oracle.time_fit(sub.subset(np.arange(n1)), opts_c, 1)
    base = {"value": n / sec, "unit": "tracks/s", "cores": threads, "kind": "port",
            "sample": f"first {n} tracks of the same workload, FP64 oracle "
                      f"(oracle/kfo_fitter.hpp), OpenMP over tracks, {threads} threads, "
                      f"fit loop timed alone ({sec:.2f} s)",
            "single_thread": {"value": n1 / sec1, "unit": "tracks/s",
                              "sample": f"first {n1} tracks, 1 thread ({sec1:.2f} s)"},
            "host": host_cpu()}
    ores, _, _, _ = oracle.fit(sub, opts_c, n_threads=threads)
    other_errors = {}
    for name, (ob, ores_gpu) in (others or {}).items():
        o, _, _, _ = oracle.fit(ob, opts_c, n_threads=threads)
        other_errors[name] = dict(param_errors(ores_gpu, o, ob.n_tracks),
                                  sample=f"first {ob.n_tracks} tracks")
    return base, param_errors(gpu_results, ores, n), other_errors


def host_cpu():
    """CPU model and the number of host cores the process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "cores_usable": usable, "cores_online": os.cpu_count()}


def param_errors(gpu_results, ores, n):
    """Per-parameter deviation of the first n tracks of each GPU result from
    the oracle's: |x_gpu - x_cpu| / max(|x_cpu|, sigma_cpu), max and 99.9 %
    quantile over the tracks the oracle fitted."""
    import numpy as np
    from gpukalmanfitter_amd import abi
    ok = (ores.status & abi.KF_STATUS_OK) != 0
    sig = np.sqrt(np.abs(ores.cov[abi.DIAG_INDEX][:, ok]))
    scale = np.maximum(np.abs(ores.params[:, ok]), sig)
    errors = {}
    for prec, r in gpu_results.items():
        dev = np.abs(r["params"][:, :n][:, ok] - ores.params[:, ok]) / scale
        errors[prec] = {
            "status_equal": bool(np.array_equal(r["status"][:n], ores.status)),
            "max": {nm: float(dev[k].max()) for k, nm in enumerate(PARAM_NAMES)},
            "q999": {nm: float(np.quantile(dev[k], 0.999)) for k, nm in enumerate(PARAM_NAMES)},
        }
    return errors


def launch_figures(prec, phase, n, ms):
    """Flops, roofline time and bytes of one launch of `phase` (module docstring)."""
    mixed = prec == "mixed"
    f = FLOPS_PER_TRACK - FLOPS_BWD_PER_TRACK if phase == "forward" else FLOPS_BWD_PER_TRACK
    g = 0.0
    if mixed:
        g = FLOPS32_PER_TRACK - FLOPS32_BWD_PER_TRACK if phase == "forward" else FLOPS32_BWD_PER_TRACK
    t_roof = n * (g / (FP32_PEAK_TFLOPS * 1e12) + (f - g) / (FP64_PEAK_TFLOPS * 1e12))
    t = ms * 1e-3
    ext = EXT_BYTES_PER_TRACK[phase]
    design = DESIGN_BYTES_PER_TRACK[(prec, phase)]
    traffic = TRAFFIC_BYTES_PER_LAUNCH.get((prec, phase))
    if traffic is not None:
        traffic *= n / TRAFFIC_TRACKS
    # the launch of the default schedule on a constant field without track
    # states (the last two template arguments: kFieldConst = 1, no recording)
    return {"kernel": "kf_fit_kernel<%s, %s, true, false, 1, false, 1, false>" % (
                "float" if mixed else "double", "true" if phase == "forward" else "false"),
            "kernel_ms": ms, "flops_per_track": f, "fp32_flops_per_track": g,
            "t_roof_ms": t_roof * 1e3, "tflop_per_s": f * n / t / 1e12,
            "vector_peak_tflop_per_s": f * n / t_roof / 1e12, "vector_frac": t_roof / t,
            "hbm_ext": {"bytes_per_track": ext, "gb_per_s": ext * n / t / 1e9,
                        "frac": ext * n / t / 1e9 / HBM_PEAK_GBS},
            "hbm_design": {"bytes_per_track": design, "gb_per_s": design * n / t / 1e9,
                           "frac": design * n / t / 1e9 / HBM_PEAK_GBS},
            "traffic": traffic,
            "traffic_over_ext": traffic / (ext * n) if traffic else None,
            "traffic_over_design": traffic / (design * n) if traffic else None}


def roofline(prec, n, phase_ms):
    """Vector-ALU roofline of the dominant launch (see the module docstring)."""
    ms = {"forward": phase_ms[0], "backward": phase_ms[1]}
    k = max(ms, key=ms.get)
    other = "backward" if k == "forward" else "forward"
    dom = launch_figures(prec, k, n, ms[k])
    return {"bound": "vector_alu", "achieved": dom["tflop_per_s"],
            "peak": dom["vector_peak_tflop_per_s"], "unit": "TFLOP/s",
            "frac": dom["vector_frac"], "traffic": dom["traffic"],
            "kernel": dom["kernel"], "kernel_ms": ms[k], "t_roof_ms": dom["t_roof_ms"],
            "peak_note": "F / (F32 / 157.3 + F64 / 78.6 TFLOP/s): the MI355X FP32 and FP64 "
                         "vector peaks weighted by the launch's float and double flops",
            "hbm": dict(dom["hbm_ext"], peak_gb_per_s=HBM_PEAK_GBS,
                        note="external I/O of the launch (inputs read, results written) "
                             "against 8 TB/s"),
            "dominant_launch": dom, "other_launch": launch_figures(prec, other, n, ms[other])}


def adjoint_leg(steps, warmup, sample):
    """Gradient of the summed chi2 with respect to 60 alignment parameters and
    the field scale (kf_chi2_gradient) on the alignment workload: 262,144
    tracks, 10 misaligned planes, field scale 1.003, no material effects; then
    the forward fit of the batch in both precisions.  Returns the figures, the
    first `sample` tracks and both fits' results on them."""
    import numpy as np
    import torch
    from gpukalmanfitter_amd import configs
    from gpukalmanfitter_amd import generate as G
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(4)
    kf = KalmanFitter(b.planes, b.field)
    kf.set_field_scale(1.003)
    dev = DeviceBatch(b)
    opts = KalmanFitterOptions(**dict(configs.WORKLOAD_FIT_OPTIONS, multiple_scattering=False,
                                      energy_loss=False))
    for _ in range(warmup):
        kf.chi2_gradient(dev, opts, synchronize=False)
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        kf.chi2_gradient(dev, opts, synchronize=False)
        kms.append(kf.last_kernel_ms())
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    _, _, _, flags, _ = kf.chi2_gradient(dev, opts)
    # the forward fit of the same batch in both precisions (error vs the oracle)
    fit_rates, res = {}, {}
    for prec in ("mixed", "fp64"):
        fopts = KalmanFitterOptions(precision=prec, **dict(configs.WORKLOAD_FIT_OPTIONS,
                                                            multiple_scattering=False,
                                                            energy_loss=False))
        kf.fit_batch(dev, fopts)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            kf.fit_batch(dev, fopts)
        torch.cuda.synchronize()
        fit_rates[prec] = b.n_tracks * steps / (time.perf_counter() - t0)
        r = dev.results_numpy()
        res[prec] = {k: r[k][..., :sample].copy() for k in ("params", "status")}
    kf.close()
    sub = b.subset(np.arange(min(sample, b.n_tracks)))
    sub.field.scale = 1.003
    return {"workload": "config4-alignment: 262144 tracks, 10 misaligned x-normal planes, Bz=2 T "
                        "x 1.003, d(sum chi2)/d(60 alignment params + field scale)",
            "tracks_per_s": b.n_tracks * steps / el, "ms_per_step": el * 1e3 / steps,
            "kernel_ms": float(np.mean(kms)), "tracks_left_out": int((flags != 0).sum()),
            "dtype": "f64", "forward_fit_tracks_per_s": fit_rates}, sub, res


def ragged_leg(steps, warmup, sample, precision="mixed"):
    """The ragged batch: 2,000,000 tracks over 5-30 planes with 5 % holes in
    the 101x101x201 grid field, fitted with the schedule the library's
    autotuner picks on this batch (length sort, persistent work queue, graph
    replay), beside the default one-thread-per-track schedule, then in FP64
    with the tuned schedule.  Returns the figures, the first `sample` tracks
    and both fits' results on them."""
    import numpy as np
    import torch
    from gpukalmanfitter_amd import configs
    from gpukalmanfitter_amd import generate as G
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    t0 = time.time()
    b = G.generate(3)
    print(f"[bench] generated {b.n_tracks} ragged tracks in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    opts = KalmanFitterOptions(precision=precision, **configs.WORKLOAD_FIT_OPTIONS)

    def timed():
        for _ in range(warmup):
            kf.fit_batch(dev, opts)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            kf.fit_batch(dev, opts)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    default_s = timed()
    # lane groups (the counterpart of the reference's block-per-track kernel)
    lane_rates = {}
    for lanes in (2, 4):
        kf.set_schedule(lanes_per_track=lanes, sort_by_length=1, graph=1)
        lane_rates[lanes] = b.n_tracks / timed()
    best, tune_ms = kf.autotune(dev, opts)
    tuned_s = timed()
    r = dev.results_numpy()
    res = {precision: {k: r[k][..., :sample].copy() for k in ("params", "status")}}
    # the FP64 fit with the same schedule
    opts = KalmanFitterOptions(precision="fp64", **configs.WORKLOAD_FIT_OPTIONS)
    fp64_s = timed()
    r = dev.results_numpy()
    res["fp64"] = {k: r[k][..., :sample].copy() for k in ("params", "status")}
    kf.close()
    return {"workload": "config3-ragged-grid: 2000000 tracks, U{5..30} planes at 25 mm pitch, "
                        "5% holes, 101x101x201 FP32 grid field", "precision": precision,
            "tracks_per_s": b.n_tracks / tuned_s, "ms_per_step": tuned_s * 1e3,
            "schedule": best, "autotune_best_kernel_ms": tune_ms,
            "default_schedule_tracks_per_s": b.n_tracks / default_s,
            "lane_groups_tracks_per_s": {f"{k} lanes/track, sorted, graph": v
                                         for k, v in lane_rates.items()},
            "fp64_same_schedule_tracks_per_s": b.n_tracks / fp64_s,
            "mixed_over_fp64": fp64_s / tuned_s}, b.subset(np.arange(min(sample, b.n_tracks))), res


def material_leg(steps, warmup, sample):
    """The material workload: 4,194,304 tracks over 8-40 planes of 300 um
    silicon (Highland scattering, Bethe energy loss) in the grid field.  The
    FP32-mixed and FP64 fits, each with the schedule the autotuner picks for
    it (and with 2 and 4 lanes per track beside it), and the chi2 adjoint
    with respect to 6 x 40 alignment parameters
    and the field scale (material differentiated).  Returns the figures, the
    first `sample` tracks and both fits' results on them (checked against
    the oracle in the CPU leg).  The working set exceeds the library's
    default memory budget, so every call runs in slices."""
    import numpy as np
    import torch
    from gpukalmanfitter_amd import abi, configs
    from gpukalmanfitter_amd import generate as G
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    t0 = time.perf_counter()
    b = G.generate(5)
    cpu_gen_s = time.perf_counter() - t0
    print(f"[bench] generated {b.n_tracks} material tracks in {cpu_gen_s:.1f}s",
          file=sys.stderr, flush=True)
    # the same workload from the GPU generator (arrays left in HBM); the first
    # call loads the library and warms up
    G.generate_device(5, n_tracks=65536)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    w = G.generate_device(5)
    torch.cuda.synchronize()
    gpu_gen_s = time.perf_counter() - t0
    same_hits = bool(w.n_hits == b.n_hits and np.array_equal(
        w.tensors["hit_offset"].cpu().numpy(), b.hit_offset))
    del w
    torch.cuda.empty_cache()
    print(f"[bench] GPU generator: {b.n_tracks} tracks in {gpu_gen_s * 1e3:.1f} ms "
          f"(host {cpu_gen_s:.2f} s), hit offsets equal {same_hits}", file=sys.stderr, flush=True)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    out = {"workload": "config5-material: 4194304 tracks, U{8..40} planes of 300 um silicon at "
                       "25 mm pitch, p 0.5-20 GeV, Highland + Bethe, 101x101x201 FP32 grid field",
           "memory_budget_bytes": 16 << 30,
           "generator": {"gpu_s": gpu_gen_s, "gpu_tracks_per_s": b.n_tracks / gpu_gen_s,
                         "host_s": cpu_gen_s, "host_tracks_per_s": b.n_tracks / cpu_gen_s,
                         "host_threads": os.cpu_count(), "hit_offsets_equal": same_hits,
                         "note": "seeded simulation + hit packing; host = libkfgen.so (OpenMP) "
                                 "incl. Python array setup, GPU = libkfgen_gpu.so incl. one "
                                 "device sync for the hit count"}}
    res = {}
    rates = {}
    out["schedule"] = {}
    out["lane_groups_tracks_per_s"] = {}
    for prec in ("mixed", "fp64"):
        opts = KalmanFitterOptions(precision=prec, **configs.WORKLOAD_FIT_OPTIONS)
        for lanes in (2, 4):
            kf.set_schedule(lanes_per_track=lanes, sort_by_length=1, graph=1)
            out["lane_groups_tracks_per_s"][f"{prec}, {lanes} lanes/track, sorted, graph"] = \
                b.n_tracks / timed(lambda: kf.fit_batch(dev, opts))
        best, _ = kf.autotune(dev, opts)
        out["schedule"][prec] = best
        sec = timed(lambda: kf.fit_batch(dev, opts))
        rates[prec] = b.n_tracks / sec
        r = dev.results_numpy()
        res[prec] = {k: (r[k][..., :sample].copy()) for k in ("params", "status")}
        out["fitted_fraction_" + prec] = float(((r["status"] & abi.KF_STATUS_OK) != 0).mean())
        print(f"[bench] material {prec}: {rates[prec]:.4g} tracks/s ({sec * 1e3:.2f} ms)",
              file=sys.stderr, flush=True)
    out["tracks_per_s"] = rates
    out["mixed_over_fp64"] = rates["mixed"] / rates["fp64"]
    opts = KalmanFitterOptions(**configs.WORKLOAD_FIT_OPTIONS)
    sec = timed(lambda: kf.chi2_gradient(dev, opts, synchronize=False))
    _, _, _, flags, _ = kf.chi2_gradient(dev, opts)
    out["adjoint"] = {"tracks_per_s": b.n_tracks / sec, "ms_per_step": sec * 1e3,
                      "parameters": 6 * len(b.planes) + 1,
                      "tracks_left_out": int((flags != 0).sum()), "dtype": "f64"}
    print(f"[bench] material adjoint: {b.n_tracks / sec:.4g} tracks/s ({sec * 1e3:.1f} ms)",
          file=sys.stderr, flush=True)
    kf.close()
    return out, b.subset(np.arange(min(sample, b.n_tracks))), res


def config1_leg(steps):
    """The reference's CPU-runnable case: 10,000 tracks through the 10-plane
    detector, fitted in both precisions; returns the figures, the batch and
    the results (checked against the oracle in the CPU leg)."""
    import torch
    from gpukalmanfitter_amd import configs
    from gpukalmanfitter_amd import generate as G
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(1)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    rates, res = {}, {}
    for prec in ("mixed", "fp64"):
        opts = KalmanFitterOptions(precision=prec, **configs.WORKLOAD_FIT_OPTIONS)
        kf.fit_batch(dev, opts)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            kf.fit_batch(dev, opts)
        torch.cuda.synchronize()
        rates[prec] = b.n_tracks * steps / (time.perf_counter() - t0)
        r = dev.results_numpy()
        res[prec] = {k: r[k].copy() for k in ("params", "status")}
    kf.close()
    return {"workload": "config1-fp64-10k: 10000 tracks, 10 x-normal planes x=30..300 mm, "
                        "Bz=2 T, no material", "tracks_per_s": rates}, b, res


def end_to_end(kf, dev, opts, steps):
    """The fit as the reference's driver times it (inputs from host memory,
    results back to host): per step, H2D copies of every input array from
    pinned host buffers into the resident device buffers, the fit, and D2H
    copies of every result array into pinned host buffers, on one stream,
    timed between device synchronisations.  Allocation is outside the timed
    region (the reference's timing also includes its cudaMalloc/free)."""
    import torch
    host_in = {k: v.cpu().pin_memory() for k, v in dev.t.items()}
    host_out = {k: torch.empty(v.shape, dtype=v.dtype).pin_memory() for k, v in dev.out.items()}
    h2d = sum(v.numel() * v.element_size() for v in host_in.values())
    d2h = sum(v.numel() * v.element_size() for v in host_out.values())

    def step():
        for k, v in host_in.items():
            dev.t[k].copy_(v, non_blocking=True)
        kf.fit_batch(dev, opts)
        for k, v in host_out.items():
            v.copy_(dev.out[k], non_blocking=True)

    step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t) / steps
    return {"tracks_per_s": dev.n_tracks / sec, "ms_per_step": sec * 1e3, "h2d_bytes": h2d,
            "d2h_bytes": d2h, "pcie_gb_per_s": (h2d + d2h) / sec / 1e9,
            "note": "pinned host inputs -> H2D -> fit -> D2H results per step; not the "
                    "headline value (inputs resident in HBM)"}


def pull_summary(res, batch):
    """Gaussian fits of the pulls (fitted - truth) / error of the successful
    tracks, per perigee parameter (the perigeeParamResolution.C report):
    a physics check of the timed fit's output."""
    from gpukalmanfitter_amd import writers as W
    tab = W.params_table(res["params"], res["cov"], res["status"], batch.truth, batch.perigee)
    rep = W.resolution_report(tab, residual_range="auto")
    return {k: {"mean": round(v["pull"]["mean"], 4), "sigma": round(v["pull"]["sigma"], 4)}
            for k, v in rep.items()}


def intersect_leg(steps, warmup, n_rays=16384, cpu_rays=16384):
    """Batched ray x triangle intersection (kf_intersect): rays from the
    beam spot against a 10-layer barrel of about 37k triangles (the size of
    the reference's TrackML mesh), boundary check on.  Device time from HIP
    events around the call on its stream.  Returns the figures and the first
    `cpu_rays` rays with their GPU hits, for the CPU baseline."""
    import numpy as np
    import torch
    from gpukalmanfitter_amd import intersect as I
    surfaces = I.polygon_planes(I.barrel_mesh())
    p, d = I.make_rays(n_rays, seed=17)
    ix = I.Intersector(surfaces)
    pos = torch.from_numpy(p).cuda()
    dirs = torch.from_numpy(d).cuda()
    cap = 64 * n_rays
    out = torch.empty(cap * I.HIT_DTYPE.itemsize, dtype=torch.uint8, device="cuda:0")
    lib = I._typed()
    scratch = torch.empty(int(lib.kf_intersect_scratch_bytes(n_rays, len(surfaces))),
                          dtype=torch.uint8,
                          device="cuda:0")
    n_hits = torch.zeros(1, dtype=torch.int64, device="cuda:0")

    def call():
        ix.run_device(pos, dirs, n_rays, True, cap, None, out, scratch, n_hits, False)

    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        call()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    total = int(n_hits.item())
    pairs = n_rays * len(surfaces)
    got = np.frombuffer(out[:min(total, cap) * I.HIT_DTYPE.itemsize].cpu().numpy().tobytes(),
                        dtype=I.HIT_DTYPE)
    res = {"workload": f"{n_rays} rays x {len(surfaces)} triangles (10-layer barrel, "
                       "ConvexPolygonBounds<3>), boundary check",
           "ms_per_call": ms, "pair_tests_per_s": pairs / (ms * 1e-3),
           "rays_per_s": n_rays / (ms * 1e-3), "hits": total, "hits_per_ray": total / n_rays,
           "dtype": "f64"}
    print(f"[bench] intersect: {res['pair_tests_per_s']:.4g} pair tests/s ({ms:.3f} ms, "
          f"{total} hits)", file=sys.stderr, flush=True)
    return res, (surfaces, p[:, :cpu_rays], d[:, :cpu_rays], got[got["ray"] < cpu_rays])


def intersect_cpu_baseline(surfaces, p, d, got, threads):
    """The intersection oracle (oracle/kfo_intersect.cpp, OpenMP over rays)
    timed on the ray sample, and whether its hit list (pairs, paths,
    positions) is byte-identical to the GPU's on that sample."""
    import numpy as np
    import oracle
    threads = threads if threads > 0 else (os.cpu_count() or 1)
    p, d = np.ascontiguousarray(p), np.ascontiguousarray(d)
    oracle.time_intersect(surfaces, p[:, :64], d[:, :64], True, threads)  # warm
    sec = oracle.time_intersect(surfaces, p, d, True, threads)
    ref = oracle.intersect(surfaces, p, d, True, threads)
    same = len(got) == len(ref) and got.tobytes() == ref.tobytes()
    n = p.shape[1]
    return {"pair_tests_per_s": n * len(surfaces) / sec, "cores": threads, "kind": "port",
            "sample": f"first {n} rays, oracle (oracle/kfo_intersect.cpp), OpenMP over rays "
                      f"({sec:.2f} s)"}, same


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); "
              "tracks are independent: give each GPU's process a contiguous range of tracks",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from gpukalmanfitter_amd import abi, configs
    from gpukalmanfitter_amd import generate as G
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions

    if not torch.cuda.is_available():
        print("bench.py: no GPU", file=sys.stderr)
        sys.exit(1)
    cfg = configs.CONFIGS[args.config]
    t0 = time.time()
    batch = G.generate(args.config, n_tracks=args.tracks)
    gen_s = time.time() - t0
    n = batch.n_tracks
    print(f"[bench] generated {n} tracks ({cfg['name']}) in {gen_s:.1f}s", file=sys.stderr, flush=True)
    if args.field != "grid" and batch.field.kind == "grid":
        v = batch.field.values
        c = v[v.shape[0] // 2, v.shape[1] // 2, v.shape[2] // 2]
        if args.field == "flat":
            batch.field = G.FieldSpec(kind="grid", values=np.broadcast_to(c, v.shape).copy(),
                                      lo=batch.field.lo, hi=batch.field.hi)
        else:
            batch.field = G.FieldSpec(kind="const", b=tuple(float(x) for x in c))
        print(f"[bench] field: {args.field} ({tuple(float(x) for x in c)})", file=sys.stderr,
              flush=True)

    kf = KalmanFitter(batch.planes, batch.field)
    t0 = time.time()
    dev = DeviceBatch(batch)
    torch.cuda.synchronize()
    h2d_s = time.time() - t0
    precisions = [p.strip() for p in args.precision.split(",") if p.strip()]
    rates = {}
    kernel_ms = {}
    phase_ms = {}
    ms_step = {}
    results = {}
    for prec in precisions:
        opts = KalmanFitterOptions(precision=prec, **configs.WORKLOAD_FIT_OPTIONS)
        for _ in range(args.warmup):
            kf.fit_batch(dev, opts)
        torch.cuda.synchronize()
        kms = []
        pms = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            kf.fit_batch(dev, opts)
            kms.append(kf.last_kernel_ms())
            try:
                pms.append(kf.last_phase_ms())
            except Exception:  # a batch fitted in memory slices has no phase split
                pms.append((float("nan"), float("nan")))
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        rates[prec] = n * args.steps / el
        ms_step[prec] = el * 1e3 / args.steps
        kernel_ms[prec] = float(np.mean(kms))
        phase_ms[prec] = [float(x) for x in np.mean(np.array(pms), axis=0)]
        results[prec] = dev.results_numpy()
        st = results[prec]["status"]
        ok = float(((st & abi.KF_STATUS_OK) != 0).mean())
        print(f"[bench] {prec}: {rates[prec]:.4g} tracks/s, {ms_step[prec]:.3f} ms/step, "
              f"kernels {kernel_ms[prec]:.3f} ms (forward {phase_ms[prec][0]:.3f}, "
              f"backward {phase_ms[prec][1]:.3f}), fitted {ok:.4f}", file=sys.stderr, flush=True)

    head = precisions[0]
    roof = roofline(head, n, phase_ms[head]) if args.config == 2 else None
    pulls = pull_summary(results[head], batch)
    e2e = end_to_end(kf, dev, KalmanFitterOptions(precision=head, **configs.WORKLOAD_FIT_OPTIONS),
                     max(args.steps // 2, 3))
    print(f"[bench] end to end ({head}, H2D + fit + D2H): {e2e['tracks_per_s']:.4g} tracks/s, "
          f"{e2e['ms_per_step']:.2f} ms/step, {e2e['pcie_gb_per_s']:.1f} GB/s over PCIe",
          file=sys.stderr, flush=True)
    others = {}
    c1 = None
    if not args.no_config1:
        c1, c1b, c1r = config1_leg(max(args.steps // 2, 3))
        others["config1"] = (c1b, c1r)
    ragged = None
    if not args.no_ragged:
        ragged, rsub, rres = ragged_leg(max(args.steps // 2, 3), 1, 10000)
        others["config3"] = (rsub, rres)
        print(f"[bench] ragged: {ragged['tracks_per_s']:.4g} tracks/s autotuned "
              f"({ragged['schedule']}), {ragged['default_schedule_tracks_per_s']:.4g} default",
              file=sys.stderr, flush=True)
    adj = None
    if not args.no_adjoint:
        adj, asub, ares = adjoint_leg(max(args.steps // 2, 3), 1, 10000)
        others["config4"] = (asub, ares)
        print(f"[bench] adjoint: {adj['tracks_per_s']:.4g} tracks/s, {adj['ms_per_step']:.3f} ms/step",
              file=sys.stderr, flush=True)
    material = None
    if not args.no_material:
        material, msub, mres = material_leg(max(args.steps // 4, 2), 1, 10000)
        others["material"] = (msub, mres)
    inter = None
    if not args.no_intersect:
        inter, isample = intersect_leg(max(args.steps // 2, 3), 2)
    cpu = None
    errors = None
    error_by_config = None
    if not args.no_cpu:
        cpu, errors, oerr = cpu_baseline(batch,
                                         abi.default_options(**configs.WORKLOAD_FIT_OPTIONS),
                                         args.cpu_threads, args.cpu_sample, results, others)
        if material is not None:
            material["error_vs_oracle"] = oerr["material"]
        if ragged is not None:
            ragged["error_vs_oracle"] = oerr["config3"]
        if adj is not None:
            adj["forward_fit_error_vs_oracle"] = oerr["config4"]
        if c1 is not None:
            c1["error_vs_oracle"] = oerr["config1"]
        summary = {"config2": errors}
        summary.update({{"material": "config5"}.get(k, k): v for k, v in oerr.items()})
        error_by_config = {
            cfg_name: {prec: {"max_over_params": max(e["max"].values()),
                              "status_equal": e["status_equal"]}
                       for prec, e in errs.items() if isinstance(e, dict) and "max" in e}
            for cfg_name, errs in sorted(summary.items())}
        if inter is not None:
            inter["cpu_baseline"], inter["oracle_bit_identical_on_sample"] = \
                intersect_cpu_baseline(*isample, args.cpu_threads)
            print(f"[bench] intersect cpu baseline {inter['cpu_baseline']['pair_tests_per_s']:.4g}"
                  f" pair tests/s, bit-identical {inter['oracle_bit_identical_on_sample']}",
                  file=sys.stderr, flush=True)
        print(f"[bench] cpu baseline {cpu['value']:.4g} tracks/s on {cpu['cores']} threads",
              file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": rates[head],
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step[head],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": cfg["name"] + f": {n} tracks, " + (
            "10 x-normal planes x=30..300 mm, Bz=2 T, no material, mu+- p 1-10 GeV" if args.config in (1, 2)
            else "see gpukalmanfitter_amd/configs.py"), "n_tracks": n, "precision": head,
            "overstep_limit_mm": configs.WORKLOAD_FIT_OPTIONS["overstep_limit"],
            **({"field": args.field} if args.field != "grid" else {})},
        "tracks_per_s": rates,
        "pulls": pulls,
        "end_to_end": e2e,
        "kernel_ms": kernel_ms,
        "h2d_s": h2d_s,
        "roofline": roof,
        "cpu_baseline": cpu,
        "error_vs_oracle": errors,
        "error_vs_oracle_by_config": error_by_config,
        "config1": c1,
        "ragged": ragged,
        "adjoint": adj,
        "material": material,
        "intersect": inter,
    }
    if "fp64" in rates and "mixed" in rates:
        out["mixed_over_fp64"] = rates["mixed"] / rates["fp64"]
    print(json.dumps(out), flush=True)
    kf.close()


if __name__ == "__main__":
    main()
