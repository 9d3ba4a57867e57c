"""The chi2 adjoint (kf_chi2_gradient) on the GPU.

Objective X = summed predicted-residual chi2 of the FP64 forward filter over
the batch; gradient with respect to 6 alignment parameters per plane (local
shifts and rotations, gpukalmanfitter_amd/alignment.py) and the field scale,
with and without material effects.
Checks: the objective equals the FP64 fit's chi2 sum; every gradient
component matches central finite differences of the FP64 fit to 1e-4, with
steps of 1e-6 mm for shifts, 1e-6 (relative) for the field scale and 1e-8 rad
for rotations: the navigator calls a surface reached within 1e-4 mm, and a
1e-6 rad tilt moves the intersection of a track 100 mm from the plane centre
by that much, flipping the step count of enough of the 262,144 tracks to
perturb a 1e-6 central difference at the 1e-4 level (the adjoint holds such
discrete decisions fixed; the oracle shows it on 20,000 tracks:
tests/test_tolerance_pins.py).  The raw (centre, rotation) adjoints
drive a torch.autograd.Function whose gradient through the alignment
parameterisation matches finite differences away from zero as well.
"""
import numpy as np
import pytest

from gpukalmanfitter_amd import abi, alignment, configs
from gpukalmanfitter_amd import generate as G

pytestmark = pytest.mark.gpu

W = dict(configs.WORKLOAD_FIT_OPTIONS, multiple_scattering=False, energy_loss=False)
TOL = 1e-4
H = 1e-6  # shifts [mm] and relative field scale
H_ROT = 1e-8  # rotations [rad]


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def chi2_of(kf, dev, opts, planes=None, scale=None):
    if planes is not None:
        kf.set_geometry(planes)
    if scale is not None:
        kf.set_field_scale(scale)
    kf.fit_batch(dev, opts)
    r = dev.results_numpy()
    return float(np.sum(r["chi2"]))


def fd_check(kf, dev, opts, planes, scale, grad, gscale, components):
    """Central differences of the FP64 fit's chi2 sum; returns the worst
    relative deviation of the listed (plane, parameter) components and of the
    field scale.  The deviation of a component is taken against
    max(|fd|, 1e-3 * the largest |fd| of its parameter kind)."""
    fds = {}
    for (p, j) in components:
        h = H if j < 3 else H_ROT
        d = np.zeros((len(planes), 6))
        d[p, j] = h
        fp = chi2_of(kf, dev, opts, alignment.apply(planes, d), scale)
        fm = chi2_of(kf, dev, opts, alignment.apply(planes, -d), scale)
        fds[(p, j)] = (fp - fm) / (2 * h)
    fp = chi2_of(kf, dev, opts, planes, scale * (1 + H))
    fm = chi2_of(kf, dev, opts, planes, scale * (1 - H))
    fd_s = (fp - fm) / (2 * H * scale)
    kf.set_geometry(planes)
    kf.set_field_scale(scale)
    worst = 0.0
    for (p, j), fd in fds.items():
        floor = 1e-3 * max(abs(v) for (q, k), v in fds.items() if k == j)
        dev_ = abs(grad[p, j] - fd) / max(abs(fd), floor)
        worst = max(worst, dev_)
        print(f"plane {p} param {j}: adjoint {grad[p, j]: .9e} fd {fd: .9e} dev {dev_:.1e}")
    ds = abs(gscale - fd_s) / abs(fd_s)
    print(f"field scale: adjoint {gscale: .9e} fd {fd_s: .9e} dev {ds:.1e}")
    return worst, ds


def test_adjoint_config4_full_matches_fd_and_forward(torch_cuda):
    """262,144 tracks on the misaligned-detector workload, field scale 1.003:
    all 61 components against central differences of the FP64 fit."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(4)
    kf = KalmanFitter(b.planes, b.field)
    kf.set_field_scale(1.003)
    dev = DeviceBatch(b)
    opts = KalmanFitterOptions(**W)
    grad, gscale, chi2, flags, raw = kf.chi2_gradient(dev, opts)
    print(f"chi2 sum {chi2:.6f}, adjoint time {kf.last_kernel_ms():.2f} ms")
    assert (flags == 0).all(), np.unique(flags, return_counts=True)
    x = chi2_of(kf, dev, opts)
    assert abs(chi2 - x) <= 1e-12 * abs(x), (chi2, x)
    comps = [(p, j) for p in range(len(b.planes)) for j in range(6)]
    worst, ds = fd_check(kf, dev, opts, b.planes, 1.003, grad, gscale, comps)
    kf.close()
    assert worst <= TOL and ds <= TOL, (worst, ds)


def test_adjoint_grid_field_holes_ragged(torch_cuda):
    """Config 3 sample (grid field, holes, 5-30 planes): a spread of planes."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(3, n_tracks=20000)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    opts = KalmanFitterOptions(**W)
    grad, gscale, chi2, flags, raw = kf.chi2_gradient(dev, opts)
    assert (flags == 0).all()
    comps = [(p, j) for p in (0, 4, 11, 19, 29) for j in range(6)]
    worst, ds = fd_check(kf, dev, opts, b.planes, 1.0, grad, gscale, comps)
    kf.close()
    print("config 3 adjoint vs central differences: worst", worst, "field scale", ds)
    assert worst <= TOL and ds <= TOL, (worst, ds)


def test_adjoint_behind_torch_autograd(torch_cuda):
    """torch.autograd.Function over (centres, rotations, field scale) with the
    raw adjoints in backward; autograd chains them through the alignment
    parameterisation at a non-zero point and matches central differences."""
    torch = torch_cuda
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(4, n_tracks=20000)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    opts = KalmanFitterOptions(**W)
    c0, r0 = alignment.centres_and_rotations(b.planes)

    class Chi2(torch.autograd.Function):
        @staticmethod
        def forward(ctx, centres, rots, scale):
            kf.set_geometry(alignment.with_geometry(b.planes, centres.detach().numpy(),
                                                    rots.detach().numpy()))
            kf.set_field_scale(float(scale))
            _, gs, chi2, flags, raw = kf.chi2_gradient(dev, opts)
            assert (flags == 0).all()
            ctx.save_for_backward(torch.from_numpy(raw[:, :3].copy()),
                                  torch.from_numpy(raw[:, 3:].reshape(-1, 3, 3).copy()),
                                  torch.tensor(gs, dtype=torch.float64))
            return torch.tensor(chi2, dtype=torch.float64)

        @staticmethod
        def backward(ctx, g):
            gc, gr, gs = ctx.saved_tensors
            return g * gc, g * gr, g * gs

    def rot(a, b_, c):
        one, zero = torch.ones((), dtype=torch.float64), torch.zeros((), dtype=torch.float64)
        rx = torch.stack([torch.stack([one, zero, zero]),
                          torch.stack([zero, torch.cos(a), -torch.sin(a)]),
                          torch.stack([zero, torch.sin(a), torch.cos(a)])])
        ry = torch.stack([torch.stack([torch.cos(b_), zero, torch.sin(b_)]),
                          torch.stack([zero, one, zero]),
                          torch.stack([-torch.sin(b_), zero, torch.cos(b_)])])
        rz = torch.stack([torch.stack([torch.cos(c), -torch.sin(c), zero]),
                          torch.stack([torch.sin(c), torch.cos(c), zero]),
                          torch.stack([zero, zero, one])])
        return rx @ ry @ rz

    C0 = torch.from_numpy(c0)
    R0 = torch.from_numpy(r0)

    def objective(theta, scale):
        cs = torch.stack([C0[i] + R0[i] @ theta[i, :3] for i in range(len(b.planes))])
        rs = torch.stack([R0[i] @ rot(theta[i, 3], theta[i, 4], theta[i, 5])
                          for i in range(len(b.planes))])
        return Chi2.apply(cs, rs, scale)

    rng = np.random.default_rng(5)
    t0 = np.zeros((len(b.planes), 6))
    t0[:, :3] = rng.normal(0, 0.02, (len(b.planes), 3))
    t0[:, 3:] = rng.normal(0, 1e-4, (len(b.planes), 3))
    theta = torch.tensor(t0, requires_grad=True)
    scale = torch.tensor(0.995, dtype=torch.float64, requires_grad=True)
    x = objective(theta, scale)
    x.backward()
    g = theta.grad.numpy()
    worst = 0.0
    for (p, j) in [(0, 0), (0, 4), (3, 1), (3, 5), (6, 2), (6, 3), (9, 0), (9, 5)]:
        h = H if j < 3 else H_ROT
        tp = t0.copy()
        tp[p, j] += h
        tm = t0.copy()
        tm[p, j] -= h
        with torch.no_grad():
            fd = (float(objective(torch.tensor(tp), scale)) -
                  float(objective(torch.tensor(tm), scale))) / (2 * h)
        worst = max(worst, abs(g[p, j] - fd) / max(abs(fd), 1.0))
        print(f"theta[{p},{j}]: autograd {g[p, j]: .9e} fd {fd: .9e}")
    with torch.no_grad():
        fds = (float(objective(theta.detach(), scale + H)) -
               float(objective(theta.detach(), scale - H))) / (2 * H)
    print(f"scale: autograd {float(scale.grad): .9e} fd {fds: .9e}")
    kf.close()
    assert worst <= TOL, worst
    assert abs(float(scale.grad) - fds) <= TOL * abs(fds)


def test_adjoint_memory_slices(torch_cuda):
    """A memory budget of about 1,350 tracks splits the forward filter and
    the reverse sweep into 15 slices whose sums add in slice order: the same
    gradient as one slice up to summation order, the same objective and
    flags."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(5, n_tracks=20000)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    opts = KalmanFitterOptions(**configs.WORKLOAD_FIT_OPTIONS)
    grad, gscale, chi2, flags, raw = kf.chi2_gradient(dev, opts)
    kf.set_memory_budget(64 << 20)
    grad2, gscale2, chi22, flags2, raw2 = kf.chi2_gradient(dev, opts)
    kf.close()
    assert np.array_equal(flags, flags2)
    assert abs(chi22 - chi2) <= 1e-12 * abs(chi2)
    assert abs(gscale2 - gscale) <= 1e-10 * abs(gscale)
    for j in range(6):
        scale = np.abs(grad[:, j]).max()
        assert np.abs(grad2[:, j] - grad[:, j]).max() <= 1e-10 * scale, j
    assert np.abs(raw2 - raw).max() <= 1e-10 * np.abs(raw).max()


def test_adjoint_material_config5(torch_cuda):
    """Config 5 sample (300 um silicon on every plane, multiple scattering and
    energy loss on, grid field): the material update is differentiated too."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(5, n_tracks=20000)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    opts = KalmanFitterOptions(**configs.WORKLOAD_FIT_OPTIONS)
    grad, gscale, chi2, flags, raw = kf.chi2_gradient(dev, opts)
    assert (flags == 0).all(), np.unique(flags, return_counts=True)
    x = chi2_of(kf, dev, opts)
    assert abs(chi2 - x) <= 1e-12 * abs(x), (chi2, x)
    comps = [(p, j) for p in (0, 9, 25, 39) for j in range(6)]
    worst, ds = fd_check(kf, dev, opts, b.planes, 1.0, grad, gscale, comps)
    kf.close()
    print("config 5 adjoint vs central differences: worst", worst, "field scale", ds)
    assert worst <= TOL and ds <= TOL, (worst, ds)
