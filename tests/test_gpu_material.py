"""The material update on the GPU: pulls on the fit-model simulation, in both
precisions, and FP64 parity with the oracle on the same tracks.

The tracks and bars are those of tests/material_pulls.py (60,000 tracks of
the material detector simulated with the fit's own material model, |eta| <
1.5): perigee pulls N(0, 1) within 0.03, and the charge-signed q/p pull of the
loss-only simulation fitted without multiple scattering.
"""
import numpy as np
import pytest

import material_pulls as MP
import oracle
from parity import compare_fit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def gpu_fit(batch, precision, **opt):
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    import torch
    kf = KalmanFitter(batch.planes, batch.field)
    dev = DeviceBatch(batch)
    kf.fit_batch(dev, KalmanFitterOptions(precision=precision, **opt))
    torch.cuda.synchronize()
    out = dev.results_numpy()
    kf.close()
    return out


def opts(o):
    return {k: getattr(o, k) for k in ("overstep_limit", "multiple_scattering", "energy_loss")}


@pytest.fixture(scope="module")
def fit_model(torch_cuda):
    return MP.fit_model_batch()


@pytest.fixture(scope="module")
def eloss(torch_cuda):
    return MP.eloss_batch()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_material_pulls_fit_model(fit_model, precision):
    b = fit_model
    g = gpu_fit(b, precision, **opts(MP.fit_model_options()))
    MP.check_fit_model(g["params"], g["cov"], g["status"], b.truth)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_energy_loss_pulls_fit_model(eloss, precision):
    b = eloss
    g = gpu_fit(b, precision, **opts(MP.eloss_options()))
    MP.check_eloss(g["params"], g["cov"], g["status"], b.truth)


@pytest.mark.parametrize("which", ["fit_model", "eloss"])
def test_fp64_parity_on_fit_model_simulation(fit_model, eloss, which):
    """The first 5,000 tracks of each sample against the oracle, 1e-9; the
    covariance of the loss-only sample, fitted without multiple scattering,
    within 1e-8: there a one-ulp change of the inputs moves the oracle's own
    perigee covariance by 3e-9 (tests/test_tolerance_pins.py)."""
    b = (fit_model if which == "fit_model" else eloss).subset(np.arange(5000))
    o_opt = MP.fit_model_options() if which == "fit_model" else MP.eloss_options()
    g = gpu_fit(b, "fp64", **opts(o_opt))
    o, _, _, _ = oracle.fit(b, o_opt)
    assert np.array_equal(g["status"], o.status)
    pd, cd, c2 = compare_fit(g, o)
    assert pd <= 1e-9 and c2 <= 1e-9, (pd, c2)
    assert cd <= (1e-9 if which == "fit_model" else 1e-8), cd
