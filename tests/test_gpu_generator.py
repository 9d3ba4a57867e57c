"""The workload generator on the GPU (libkfgen_gpu.so) against the host
generator (libkfgen.so): the same tracks from the same random streams.

Integer decisions (charge, sequence length, dropped hits, hit counts, geo
ids) are identical; floating-point values agree to the rounding of the two
math libraries, amplified by the truth propagation's adaptive step control
(a step accepted on one side and halved on the other moves a crossing by at
most the integration tolerance): start parameters within 1e-12 relative,
hit positions within 1e-6 mm (the hit resolution is 50 um).  A fit of the
GPU-generated tracks matches the fit of the host-generated ones.
"""
import numpy as np
import pytest

from gpukalmanfitter_amd import abi, configs
from gpukalmanfitter_amd import generate as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.parametrize("cfg,n", [(1, 20000), (3, 20000), (4, 20000), (5, 20000)])
def test_device_generator_matches_host(torch_cuda, cfg, n):
    h = G.generate(cfg, n_tracks=n)
    w = G.generate_device(cfg, n_tracks=n)
    d = w.to_host()
    assert w.n_hits == h.n_hits and w.max_states == h.max_states
    for k in ("seq_first", "seq_count", "hit_offset", "hit_geo_id"):
        assert np.array_equal(getattr(d, k), getattr(h, k)), k
    for k in ("start_params", "start_cov", "perigee", "truth"):
        a, b = getattr(d, k), getattr(h, k)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-300), (k, np.abs(a - b).max())
    assert np.abs(d.hit_meas - h.hit_meas).max() <= 1e-6
    assert np.array_equal(d.hit_cov, h.hit_cov)


def test_fit_of_device_generated_batch(torch_cuda):
    """Config 5 sample fitted from the generator's device arrays (no host copy)
    equals the fit of the host-generated sample to the generators' agreement."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    n = 20000
    h = G.generate(5, n_tracks=n)
    w = G.generate_device(5, n_tracks=n)
    kf = KalmanFitter(w.planes, w.field)
    opts = KalmanFitterOptions(**configs.WORKLOAD_FIT_OPTIONS)
    dw = w.device_batch()
    kf.fit_batch(dw, opts)
    rw = dw.results_numpy()
    dh = DeviceBatch(h)
    kf.fit_batch(dh, opts)
    rh = dh.results_numpy()
    kf.close()
    assert np.array_equal(rw["status"], rh["status"])
    ok = (rh["status"] & abi.KF_STATUS_OK) != 0
    assert ok.mean() > 0.99
    sig = np.sqrt(rh["cov"][abi.DIAG_INDEX][:, ok])
    dev = np.abs(rw["params"][:, ok] - rh["params"][:, ok]) / sig
    assert dev.max() < 1e-3, dev.max()
