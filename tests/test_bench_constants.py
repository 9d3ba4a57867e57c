"""The constants bench.py prices the kernel with are re-measured here."""
import os
import re

import numpy as np

from gpukalmanfitter_amd import abi, configs
from gpukalmanfitter_amd import generate as G

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_flops_per_track_constant(built):
    src = open(os.path.join(ROOT, "bench.py")).read()
    const = float(re.search(r"^FLOPS_PER_TRACK = ([0-9.e+]+)", src, re.M).group(1))
    b = G.generate(2, n_tracks=20000)
    _, _, flops, _ = oracle.fit(b, abi.default_options(**configs.WORKLOAD_FIT_OPTIONS), counting=True)
    assert abs(flops.mean() - const) / const < 0.01, flops.mean()
