"""The constants bench.py prices the kernel with are re-measured here."""
import os
import re

import numpy as np

from gpukalmanfitter_amd import abi, configs
from gpukalmanfitter_amd import generate as G

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_flops_per_track_constants(built):
    src = open(os.path.join(ROOT, "bench.py")).read()

    def const(name):
        return float(re.search(rf"^{name} = ([0-9.e+]+)", src, re.M).group(1))

    b = G.generate(2, n_tracks=20000)
    o, _, flops, _ = oracle.fit(b, abi.default_options(**configs.WORKLOAD_FIT_OPTIONS), counting=True)
    for name, measured in (("FLOPS_PER_TRACK", flops), ("FLOPS_BWD_PER_TRACK", o.flops_bwd),
                           ("FLOPS32_PER_TRACK", o.flops32),
                           ("FLOPS32_BWD_PER_TRACK", o.flops32_bwd)):
        c = const(name)
        assert abs(measured.mean() - c) / c < 0.01, (name, measured.mean())


def test_bytes_per_track_constants():
    """bench.EXT_BYTES_PER_TRACK and DESIGN_BYTES_PER_TRACK from the ABI arrays
    and the store layout of the kernels
    (include/kf_abi.h, gpukalmanfitter_amd/csrc/kf_kernels.hpp) and the
    config-2 track shape (10 hits, 10 measured states)."""
    src = open(os.path.join(ROOT, "bench.py")).read()
    hdr = open(os.path.join(ROOT, "gpukalmanfitter_amd", "csrc", "kf_kernels.hpp")).read()
    sd = int(re.search(r"#define KF_SD_N (\d+)", hdr).group(1))
    # data elements of the S store (the layout pads the Jacobian to a
    # 16-byte group; padding is not algorithmic traffic)
    ss = int(re.search(r"#define KF_SS_USED (\d+)", hdr).group(1))
    import ast
    table = ast.literal_eval(re.search(r"DESIGN_BYTES_PER_TRACK = (\{.*?\})", src, re.S).group(1))
    ext = ast.literal_eval(re.search(r"EXT_BYTES_PER_TRACK = (\{.*?\})", src, re.S).group(1))
    b = G.generate(2, n_tracks=1000)
    assert (np.diff(b.hit_offset) == 10).all() and (b.seq_count == 10).all()
    hits = 10 * (4 + 2 * 8 + 3 * 8)
    inputs = 6 * 8 + 21 * 8 + 3 * 8 + 2 * 4 + 8 + hits
    carry = 8 * 4 + 30 * 8
    outputs = 6 * 8 + 21 * 8 + 4
    assert ext["forward"] == inputs + 16
    assert ext["backward"] == 48 + outputs
    for prec, ssize in (("fp64", 8), ("mixed", 4)):
        store = 10 * (sd * 8 + ss * ssize)
        assert table[(prec, "forward")] == inputs + 16 + carry + store
        assert table[(prec, "backward")] == store + carry + 48 + outputs
