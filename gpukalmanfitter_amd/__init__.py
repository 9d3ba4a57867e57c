"""MI355X-native Kalman track fitter (drop-in for gpuKalmanFitter's per-track fit).

* :mod:`gpukalmanfitter_amd.fitter`   -- KalmanFitter over the C ABI of libkfmi.so
* :mod:`gpukalmanfitter_amd.generate` -- seeded synthetic workloads (libkfgen.so)
* :mod:`gpukalmanfitter_amd.configs`  -- the five workload definitions
* :mod:`gpukalmanfitter_amd.abi`      -- ctypes mirror of include/kf_abi.h
"""
from . import abi, configs  # noqa: F401

__all__ = ["abi", "configs"]
