#!/usr/bin/env python3
"""Throughput benchmark of the MI355X Kalman fitter (one JSON line on stdout).

Workload (config 2): 1,048,576 synthetic tracks through 10 x-normal planes at
x = 30..300 mm in a uniform 2 T field, no material, mu+- with p in [1, 10] GeV,
|phi| <= 0.3, |eta| <= 0.3; one 2-D hit per plane; start parameters smeared
(d0/z0 100 um, phi/theta 1 mrad, q/p 5 %).  A step is one kf_fit() over the
whole batch with inputs already resident in HBM; the timed region is exactly
--steps fits bracketed by device synchronisation.  Fits run in every precision
given by --precision (first one is the headline `value`).

The roofline object prices the fit kernel against MI355X's fp64 vector peak
(78.6 TFLOP/s, equal to its fp64 matrix peak) with the algorithmic flop count
per track of the reference formulation (structural zeros and identities
removed), measured by the counting build of the CPU oracle on this workload:
FLOPS_PER_TRACK below.  The kernel duration is the mean over the timed steps of
HIP events recorded by the library on the launch stream.
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
# oracle, liboracle_count.so; tests/test_bench_constants.py re-measures it).
FLOPS_PER_TRACK = 30610.0
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE (scaled by the
# copy-kernel calibration), per precision; None until profiled.
TRAFFIC_BYTES_PER_LAUNCH = {}


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--tracks", type=int, default=None, help="override the workload size")
    ap.add_argument("--precision", default="fp64", help="comma list of fp64,mixed")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--cpu-sample", type=int, default=1048576,
                    help="tracks of the same workload timed on the CPU oracle")
    ap.add_argument("--no-cpu", action="store_true")
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


def cpu_baseline(batch, opts_c, threads, sample):
    """The CPU oracle (FP64 restatement of the reference fit, OpenMP over
    tracks as Run/CPU/KalmanFitterCPUTest.cpp) on a bounded sample."""
    import numpy as np
    import oracle
    n = min(sample, batch.n_tracks)
    sub = batch.subset(np.arange(n))
    if threads <= 0:
        threads = os.cpu_count() or 1
    oracle.time_fit(sub.subset(np.arange(min(n, 2000))), opts_c, threads)  # warm
    sec = oracle.time_fit(sub, opts_c, threads)
    return {"value": n / sec, "unit": "tracks/s", "cores": threads, "kind": "port",
            "sample": f"first {n} tracks of the same workload, FP64 oracle "
                      f"(oracle/kfo_fitter.hpp), OpenMP over tracks, {threads} threads, "
                      f"fit loop timed alone ({sec:.2f} s)"}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); "
              "tracks are independent, see DESIGN.md for how the batch would shard", file=sys.stderr)
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

    kf = KalmanFitter(batch.planes, batch.field)
    t0 = time.time()
    dev = DeviceBatch(batch)
    torch.cuda.synchronize()
    h2d_s = time.time() - t0
    precisions = [p.strip() for p in args.precision.split(",") if p.strip()]
    rates = {}
    kernel_ms = {}
    ms_step = {}
    for prec in precisions:
        opts = KalmanFitterOptions(precision=prec, **configs.WORKLOAD_FIT_OPTIONS)
        for _ in range(args.warmup):
            kf.fit_batch(dev, opts)
        torch.cuda.synchronize()
        kms = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            kf.fit_batch(dev, opts)
            kms.append(kf.last_kernel_ms())
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        rates[prec] = n * args.steps / el
        ms_step[prec] = el * 1e3 / args.steps
        kernel_ms[prec] = float(np.mean(kms))
        st = dev.results_numpy()["status"]
        ok = float(((st & abi.KF_STATUS_OK) != 0).mean())
        print(f"[bench] {prec}: {rates[prec]:.4g} tracks/s, {ms_step[prec]:.3f} ms/step, "
              f"kernel {kernel_ms[prec]:.3f} ms, fitted {ok:.4f}", file=sys.stderr, flush=True)

    head = precisions[0]
    roof = None
    if FLOPS_PER_TRACK:
        achieved = FLOPS_PER_TRACK * n / (kernel_ms[head] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "traffic": TRAFFIC_BYTES_PER_LAUNCH.get(head),
                "compute": "fp64 vector ALU (MI355X fp64 vector peak equals its fp64 "
                           "matrix peak, 78.6 TFLOP/s); algorithmic flops/track "
                           f"{FLOPS_PER_TRACK:.0f} from the counting oracle",
                "kernel_ms": kernel_ms[head]}
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(batch, abi.default_options(**configs.WORKLOAD_FIT_OPTIONS),
                           args.cpu_threads, args.cpu_sample)
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
            "overstep_limit_mm": configs.WORKLOAD_FIT_OPTIONS["overstep_limit"]},
        "tracks_per_s": rates,
        "kernel_ms": kernel_ms,
        "h2d_s": h2d_s,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "fp64" in rates and "mixed" in rates:
        out["mixed_over_fp64"] = rates["mixed"] / rates["fp64"]
    print(json.dumps(out), flush=True)
    kf.close()


if __name__ == "__main__":
    main()
