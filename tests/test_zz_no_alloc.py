"""kf_fit_noalloc (the fit without device allocation, on the library's fixed
per-device workspace) against kf_fit, through the C ABI.

Two calls in a row on one context, the second with a smaller batch: each
must give kf_fit's results bit for bit (the same kernels; slicing never
changes a result).  The first batch needs several workspace slices (FP64
material tracks of up to 40 states: 40 x 512 B per track, ~12,900 tracks
per slice), the second fits in one, with track states.
"""
import ctypes as C

import numpy as np
import pytest

from gpukalmanfitter_amd import abi, configs
from gpukalmanfitter_amd import generate as G

pytestmark = pytest.mark.gpu

W = configs.WORKLOAD_FIT_OPTIONS
KEYS = ("params", "cov", "chi2", "ndf", "n_measured", "status")


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def fit_both(kf, b, precision, track_states=False):
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitterOptions
    import torch
    opts = KalmanFitterOptions(precision=precision, **W)
    ref_dev = DeviceBatch(b, track_states=track_states)
    kf.fit_batch(ref_dev, opts)
    torch.cuda.synchronize()
    ref = ref_dev.results_numpy()
    dev = DeviceBatch(b, track_states=track_states)
    kf.fit_batch_noalloc(dev, opts)
    torch.cuda.synchronize()
    return ref, dev.results_numpy()


def test_noalloc_equals_fit_on_two_calls(torch_cuda):
    from gpukalmanfitter_amd.fitter import KalmanFitter
    big = G.generate(5, n_tracks=30000)
    small = G.generate(5, n_tracks=2000, seed=99)
    kf = KalmanFitter(big.planes, big.field)
    per_track = big.max_states * 512 + 272
    assert 30000 * per_track > abi.KF_NOALLOC_WORKSPACE_BYTES  # several slices
    print(f"[noalloc] call 1: {big.n_tracks} tracks fp64", flush=True)
    ref, got = fit_both(kf, big, "fp64")
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k
    assert ((got["status"] & abi.KF_STATUS_OK) != 0).mean() > 0.99
    print(f"[noalloc] call 2: {small.n_tracks} tracks mixed, track states", flush=True)
    ref, got = fit_both(kf, small, "mixed", track_states=True)
    for k in KEYS + ("track_states",):
        assert np.array_equal(got[k], ref[k]), k
    kf.close()


def test_noalloc_rejects_a_slice_beyond_the_workspace(torch_cuda):
    """A state capacity whose 64-track slice exceeds the workspace is refused
    with a message, before any launch."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, FitError, KalmanFitter, KalmanFitterOptions
    b = G.generate(1, n_tracks=64)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    with pytest.raises(FitError, match="fixed workspace"):
        kf.fit_batch_noalloc(dev, KalmanFitterOptions(max_states=9000, **W))
    # the context still fits afterwards
    kf.fit_batch_noalloc(dev, KalmanFitterOptions(**W))
    kf.close()


def test_noalloc_c_abi_symbol(torch_cuda):
    from gpukalmanfitter_amd._lib import load_library
    lib = load_library("libkfmi.so")
    assert isinstance(lib.kf_fit_noalloc, C._CFuncPtr)
