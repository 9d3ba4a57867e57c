"""Parity of the HIP fitter (through the C ABI) with the CPU oracle.

Bar: the plain FP64 mode matches the oracle within 1e-9 relative (metric in
tests/parity.py) on every output, per-track status bits are identical, and at
the full 1,048,576-track size the fit is deterministic and its pulls are
N(0, 1).  Inputs are the seeded workloads of gpukalmanfitter_amd.configs.
"""
import ctypes
import math

import numpy as np
import pytest

from gpukalmanfitter_amd import abi, configs
from gpukalmanfitter_amd import generate as G

import oracle
from parity import compare_fit, compare_states  # noqa: E402
from test_host_fit import state_bar  # noqa: E402

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-9


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def run_gpu(batch, track_states=False, planes=None, field_scale=None, **opt):
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    import torch
    kf = KalmanFitter(planes if planes is not None else batch.planes, batch.field)
    if field_scale is not None:
        kf.set_field_scale(field_scale)
    dev = DeviceBatch(batch, track_states=track_states)
    kf.fit_batch(dev, KalmanFitterOptions(**opt))
    torch.cuda.synchronize()
    out = dev.results_numpy()
    kf.close()
    return out


def c_options(**opt):
    o = abi.default_options()
    for k, v in opt.items():
        if k == "precision":
            v = abi.KF_PRECISION_FP64 if v == "fp64" else abi.KF_PRECISION_MIXED
        setattr(o, k, int(v) if isinstance(v, bool) else v)
    return o


def check_parity(g, o, tol=TOL_FP64):
    assert np.array_equal(g["status"], o.status), np.unique(g["status"] ^ o.status)
    assert np.array_equal(g["n_measured"], o.n_measured)
    assert np.array_equal(g["ndf"], o.ndf)
    pd, cd, c2 = compare_fit(g, o)
    assert pd <= tol and cd <= tol and c2 <= tol, (pd, cd, c2)
    return pd, cd, c2


W = configs.WORKLOAD_FIT_OPTIONS


def test_fp64_config1_full(torch_cuda):
    b = G.generate(1)
    g = run_gpu(b, **W)
    o, _, _, _ = oracle.fit(b, c_options(**W))
    assert (o.status == (abi.KF_STATUS_OK | abi.KF_STATUS_SMOOTHED | abi.KF_STATUS_FINISHED)).all()
    print("config1 fp64 max dev (params, cov, chi2):", check_parity(g, o))


def test_fp64_reference_default_options(torch_cuda):
    """Reference defaults (0.01 mm overstep): the lost tracks are lost on both."""
    b = G.generate(1, n_tracks=3000)
    g = run_gpu(b)
    o, _, _, _ = oracle.fit(b, c_options())
    lost = (o.status & abi.KF_STATUS_OK) == 0
    assert lost.sum() > 100
    check_parity(g, o)


@pytest.mark.parametrize("cfg,n", [(3, 3000), (5, 3000)])
def test_fp64_ragged_grid_holes_material(torch_cuda, cfg, n):
    b = G.generate(cfg, n_tracks=n)
    g = run_gpu(b, **W)
    o, _, _, _ = oracle.fit(b, c_options(**W))
    assert ((o.status & abi.KF_STATUS_OK) != 0).mean() > 0.99
    check_parity(g, o)


@pytest.mark.parametrize("sched", [dict(), dict(sort_by_length=1), dict(store_by_plane=1),
                                   dict(sort_by_length=1, store_by_plane=1)])
def test_fp64_fewer_state_slots_than_planes(torch_cuda, sched):
    """Config 3 (5-30 planes, 5 % holes) with room for 22 states: with store
    slots by plane, a sequence of up to 22 planes keeps its states in the
    slots of its planes' positions and a longer one shifts them back into
    the 22 slots; a track that measures more than 22 states fails, as the
    oracle's fixed state capacity.  FP64 within 1e-9 of the oracle, track
    states included, with either slot rule, with and without the length
    sort."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    import torch
    b = G.generate(3, n_tracks=3000)
    assert (b.seq_count > 22).any() and (b.seq_count <= 22).any()
    kf = KalmanFitter(b.planes, b.field)
    if sched:
        kf.set_schedule(**sched)
    dev = DeviceBatch(b, track_states=True, max_states=22)
    kf.fit_batch(dev, KalmanFitterOptions(max_states=22, **W))
    torch.cuda.synchronize()
    g = dev.results_numpy()
    kf.close()
    o, _, _, _ = oracle.fit(b, c_options(max_states=22, **W), track_states=True)
    long_ok = (b.seq_count > 22) & (np.diff(b.hit_offset) <= 22)
    assert long_ok.any() and ((o.status[long_ok] & abi.KF_STATUS_OK) != 0).all()
    check_parity(g, o)
    smoothed = ((o.status & abi.KF_STATUS_SMOOTHED) != 0) & ((o.status & abi.KF_STATUS_SMOOTHER_NAN) == 0)
    worst = compare_states(g["track_states"], o.track_states, o.n_measured, smoothed)
    for k, v in worst.items():
        assert v <= (0 if k == "geo_id" else TOL_FP64), (k, v)


def test_fp64_misaligned_detector_and_field_scale(torch_cuda):
    b = G.generate(4, n_tracks=2000)
    # fit with the nominal planes and a non-unit field scale
    b.field.scale = 1.003
    g = run_gpu(b, field_scale=1.003, **W)
    o, _, _, _ = oracle.fit(b, c_options(**W))
    check_parity(g, o)
    # and with the true (misaligned) planes
    g2 = run_gpu(b, planes=b.true_planes, field_scale=1.003, **W)
    o2, _, _, _ = oracle.fit(b, c_options(**W), planes=b.true_planes)
    check_parity(g2, o2)


def test_fp64_track_states(torch_cuda):
    b = G.generate(3, n_tracks=800)
    g = run_gpu(b, track_states=True, **W)
    o, _, _, _ = oracle.fit(b, c_options(**W), track_states=True)
    check_parity(g, o)
    smoothed = ((o.status & abi.KF_STATUS_SMOOTHED) != 0) & ((o.status & abi.KF_STATUS_SMOOTHER_NAN) == 0)
    worst = compare_states(g["track_states"], o.track_states, o.n_measured, smoothed)
    for k, v in worst.items():
        assert v <= (0 if k == "geo_id" else TOL_FP64), (k, v)


@pytest.mark.parametrize("opt", [dict(smoothing=False), dict(multiple_scattering=False),
                                 dict(energy_loss=False), dict(max_steps=7),
                                 dict(tolerance=1e-6)])
def test_fp64_option_variants(torch_cuda, opt):
    b = G.generate(5, n_tracks=1000)
    full = dict(W)
    full.update(opt)
    g = run_gpu(b, **full)
    o, _, _, _ = oracle.fit(b, c_options(**full))
    assert np.array_equal(g["status"], o.status)
    pd, cd, c2 = compare_fit(g, o)
    assert pd <= TOL_FP64 and c2 <= TOL_FP64, (pd, c2)
    # Without multiple-scattering noise the RTS recursion
    # C_s = C_f - G (C_p - C_s') G^T cancels almost completely in var(q/p) on
    # long tracks and amplifies the ~1e-15 covariance differences of the
    # forward pass (symmetric storage, J accumulated per step) up to ~1e6 at the
    # perigee; the oracle's own rounding is amplified alike: one ulp of the
    # start covariance moves its perigee covariance by 1e-9 to 1e-8
    # (tests/test_tolerance_pins.py).
    assert cd <= (1e-8 if opt == dict(multiple_scattering=False) else TOL_FP64), cd


def test_edge_cases(torch_cuda):
    b = G.generate(1, n_tracks=6)
    # 0: q/p = 0 (bad input); 1: sequence outside the plane table; 2: no hits;
    # 3: foreign hit (plane not in sequence); 4, 5 regular
    b.start_params[4, 0] = 0.0
    b.seq_first[1] = 5  # 5 + 10 > 10 planes
    counts = np.diff(b.hit_offset).copy()
    counts[2] = 0
    keep = np.ones(b.n_hits, dtype=bool)
    keep[b.hit_offset[2]:b.hit_offset[3]] = False
    b.hit_geo_id[b.hit_offset[3] + 2] = 77
    b.hit_offset = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    b.hit_geo_id = b.hit_geo_id[keep]
    b.hit_meas = np.ascontiguousarray(b.hit_meas[:, keep])
    b.hit_cov = np.ascontiguousarray(b.hit_cov[:, keep])
    g = run_gpu(b, **W)
    st = g["status"]
    assert st[0] & abi.KF_STATUS_BAD_INPUT and st[1] & abi.KF_STATUS_BAD_INPUT
    assert st[2] & abi.KF_STATUS_BAD_INPUT
    assert not st[3] & abi.KF_STATUS_OK and g["n_measured"][3] == 9
    assert (st[4:] & abi.KF_STATUS_OK).all()
    # the oracle agrees on every track it can take (it does not range-check
    # the sequence, so track 1 is left out)
    sub = b.subset([0, 2, 3, 4, 5])
    o, _, _, _ = oracle.fit(sub, c_options(**W))
    assert np.array_equal(o.status, st[[0, 2, 3, 4, 5]])


def _with_hits(b, order):
    """b with the hits of every track replaced by order[t] (hit indices)."""
    import copy
    nb = copy.copy(b)
    idx = np.concatenate([np.asarray(o, dtype=np.int64) for o in order])
    nb.hit_offset = np.concatenate([[0], np.cumsum([len(o) for o in order])]).astype(np.int64)
    nb.hit_geo_id = np.ascontiguousarray(b.hit_geo_id[idx])
    nb.hit_meas = np.ascontiguousarray(b.hit_meas[:, idx])
    nb.hit_cov = np.ascontiguousarray(b.hit_cov[:, idx])
    return nb


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hit_order_and_duplicate_hits(torch_cuda, precision):
    """The hit of a plane is the track's first hit with that geo id, in hit
    order.  Shuffled hit lists fit bitwise as the ordered ones (the search
    leaves its in-order fast path); a copy of a track's fourth hit with a
    moved measurement, placed before its second hit, is the one used; the
    oracle agrees on every track."""
    b = G.generate(3, n_tracks=3000)
    rng = np.random.default_rng(3)
    ranges = [np.arange(b.hit_offset[t], b.hit_offset[t + 1]) for t in range(b.n_tracks)]
    shuffled = [rng.permutation(r) if t % 2 == 0 else r for t, r in enumerate(ranges)]
    dup = []
    for t, r in enumerate(ranges):
        if t % 3 == 0 and len(r) >= 4:
            dup.append(np.concatenate([r[:1], [r[3]], r[1:]]))
        else:
            dup.append(r)
    bd = _with_hits(b, dup)
    moved = np.zeros(bd.n_hits, dtype=bool)
    for t in range(b.n_tracks):
        if t % 3 == 0 and len(ranges[t]) >= 4:
            moved[bd.hit_offset[t] + 1] = True
    bd.hit_meas = bd.hit_meas.copy()
    bd.hit_meas[0, moved] += 0.1  # 2 sigma: changes the fit of those tracks
    opts = dict(W, precision=precision)
    ref = run_gpu(b, **opts)
    got = run_gpu(_with_hits(b, shuffled), **opts)
    for k in ("params", "cov", "chi2", "status", "n_measured"):
        assert np.array_equal(got[k], ref[k]), k
    g = run_gpu(bd, **opts)
    o, _, _, _ = oracle.fit(bd, c_options(**W))
    sel = [t for t in range(b.n_tracks) if t % 3 == 0 and len(ranges[t]) >= 4]
    assert np.abs(g["chi2"][sel] - ref["chi2"][sel]).max() > 1.0  # the copy was used
    if precision == "fp64":
        check_parity(g, o)
    else:
        pd, cd, _ = compare_fit(g, o)
        assert pd <= 2 * TOL_MIXED_PARAMS and cd <= 2 * TOL_MIXED_COV, (pd, cd)


def test_empty_batch_and_single_track(torch_cuda):
    b = G.generate(1, n_tracks=1)
    g = run_gpu(b, **W)
    o, _, _, _ = oracle.fit(b, c_options(**W))
    check_parity(g, o)
    e = b.subset(np.zeros(0, dtype=np.int64))
    g0 = run_gpu(e, **W)
    assert g0["status"].shape == (1,) and g0["status"][0] == 0  # untouched buffer


def test_per_track_fit_api(torch_cuda):
    """KalmanFitter.fit(): the reference's per-track call shape."""
    from gpukalmanfitter_amd.fitter import (BoundParameters, KalmanFitter, KalmanFitterOptions,
                                            PixelSourceLink, unpack_sym)
    b = G.generate(1, n_tracks=3)
    kf = KalmanFitter(b.planes, b.field)
    i = 2
    h0, h1 = b.hit_offset[i], b.hit_offset[i + 1]
    sls = [PixelSourceLink((b.hit_meas[0, h], b.hit_meas[1, h]),
                           ((b.hit_cov[0, h], b.hit_cov[1, h]), (b.hit_cov[1, h], b.hit_cov[2, h])),
                           int(b.hit_geo_id[h])) for h in range(h0, h1)]
    start = BoundParameters(b.start_params[:, i], unpack_sym(b.start_cov[:, i]))
    ok, res = kf.fit(sls, start, KalmanFitterOptions(**W), b.perigee[:, i], 0, 10)
    o, _, _, _ = oracle.fit(b.subset([i]), c_options(**W), track_states=True)
    assert ok and res.smoothed and res.finished and res.measurement_states == 10
    assert np.allclose(res.fitted_parameters.parameters, o.params[:, 0], rtol=1e-12, atol=1e-15)
    assert res.fitted_states.shape == (10, abi.KF_TS_STRIDE)
    kf.close()


def test_non_default_stream_and_kernel_timer(torch_cuda):
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    torch = torch_cuda
    b = G.generate(1, n_tracks=2048)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        kf.fit_batch(dev, KalmanFitterOptions(**W), stream=s)
    s.synchronize()
    assert kf.last_kernel_ms() > 0
    o, _, _, _ = oracle.fit(b, c_options(**W))
    check_parity(dev.results_numpy(), o)
    kf.close()


def test_full_size_config2_properties(torch_cuda):
    """1,048,576 tracks: every track fitted, bitwise deterministic across runs,
    pulls N(0, 1), and a 5000-track sample equal to the oracle."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    torch = torch_cuda
    b = G.generate(2)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    opts = KalmanFitterOptions(**W)
    kf.fit_batch(dev, opts)
    r1 = dev.results_numpy()
    kf.fit_batch(dev, opts)
    r2 = dev.results_numpy()
    kf.close()
    assert np.array_equal(r1["params"], r2["params"]) and np.array_equal(r1["cov"], r2["cov"])
    ok = (r1["status"] & abi.KF_STATUS_OK) != 0
    assert ok.all()
    pull = (r1["params"] - b.truth) / np.sqrt(r1["cov"][abi.DIAG_INDEX])
    n = b.n_tracks
    for k in range(6):
        assert abs(pull[k].mean()) < 5.0 / math.sqrt(n) + 0.01, k
        assert abs(pull[k].std() - 1.0) < 0.02, k
    idx = np.sort(np.random.default_rng(7).choice(n, 5000, replace=False))
    sub = b.subset(idx)
    o, _, _, _ = oracle.fit(sub, c_options(**W))
    g = {k: (v[..., idx] if v.ndim > 1 else v[idx]) for k, v in r1.items()}
    check_parity(g, o)


# ---------------------------------------------------------------- mixed mode --
# FP32 transport Jacobian, covariance transport and material formulas; FP64
# positions, directions, RK stages, 2x2 inversion, Joseph update and smoother.
# Bar: every fitted parameter within 1e-5 of the FP64 oracle (metric of
# tests/parity.py), identical status bits.  The smoothed covariance goes
# through the RTS cancellation C_f - G (C_p - C_s') G^T, which amplifies the
# float rounding of the stored filtered covariances; its bar is 1e-3 (float-
# rounded RTS operands move the oracle's own smoothed covariances by ~1e-4 on
# 300 tracks: tests/test_tolerance_pins.py).
TOL_MIXED_PARAMS = 1e-5
TOL_MIXED_COV = 1e-3


@pytest.mark.parametrize("cfg,n", [(1, 10000), (3, 4000), (5, 4000), (4, 4000)])
def test_mixed_accuracy_vs_oracle(torch_cuda, cfg, n):
    b = G.generate(cfg, n_tracks=n)
    g = run_gpu(b, precision="mixed", **W)
    o, _, _, _ = oracle.fit(b, c_options(**W))
    assert np.array_equal(g["status"], o.status), np.unique(g["status"] ^ o.status)
    assert np.array_equal(g["n_measured"], o.n_measured)
    pd, cd, c2 = compare_fit(g, o)
    print(f"config {cfg} mixed max dev (params, cov, chi2): {pd:.2e} {cd:.2e} {c2:.2e}")
    assert pd <= TOL_MIXED_PARAMS and cd <= TOL_MIXED_COV and c2 <= TOL_MIXED_PARAMS, (pd, cd, c2)


def test_mixed_track_states_and_options(torch_cuda):
    b = G.generate(5, n_tracks=1000)
    g = run_gpu(b, track_states=True, precision="mixed", **W)
    o, _, _, _ = oracle.fit(b, c_options(**W), track_states=True)
    assert np.array_equal(g["status"], o.status)
    smoothed = ((o.status & abi.KF_STATUS_SMOOTHED) != 0) & ((o.status & abi.KF_STATUS_SMOOTHER_NAN) == 0)
    # a state's chi2 is a squared residual in units of its sigma: its float
    # rounding is bounded against max(chi2, 1), not against chi2 itself when
    # the residual is small
    worst = compare_states(g["track_states"], o.track_states, o.n_measured, smoothed,
                           chi2_floor=1.0)
    print("mixed track-state deviations:", worst)
    assert worst["geo_id"] == 0
    for k in ("predicted_par", "filtered_par", "smoothed_par", "path", "chi2"):
        assert worst[k] <= TOL_MIXED_PARAMS, (k, worst[k])
    for k in ("predicted_cov", "filtered_cov", "jacobian"):
        assert worst[k] <= 1e-5, (k, worst[k])
    assert worst["smoothed_cov"] <= TOL_MIXED_COV, worst["smoothed_cov"]
    for opt in (dict(smoothing=False), dict(energy_loss=False), dict(max_steps=7)):
        full = dict(W)
        full.update(opt)
        g = run_gpu(b, precision="mixed", **full)
        o, _, _, _ = oracle.fit(b, c_options(**full))
        assert np.array_equal(g["status"], o.status), opt
        pd, cd, c2 = compare_fit(g, o)
        assert pd <= TOL_MIXED_PARAMS and cd <= TOL_MIXED_COV, (opt, pd, cd)


def test_mixed_full_size_config2_pulls(torch_cuda):
    """1,048,576 tracks in mixed mode: all fitted, pulls N(0, 1), a 10,000-track
    subset within the error bar of the oracle."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(2)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    kf.fit_batch(dev, KalmanFitterOptions(precision="mixed", **W))
    r = dev.results_numpy()
    kf.close()
    assert ((r["status"] & abi.KF_STATUS_OK) != 0).all()
    pull = (r["params"] - b.truth) / np.sqrt(r["cov"][abi.DIAG_INDEX])
    n = b.n_tracks
    for k in range(6):
        assert abs(pull[k].mean()) < 5.0 / math.sqrt(n) + 0.01, k
        assert abs(pull[k].std() - 1.0) < 0.02, k
    idx = np.sort(np.random.default_rng(11).choice(n, 10000, replace=False))
    o, _, _, _ = oracle.fit(b.subset(idx), c_options(**W))
    g = {k: (v[..., idx] if v.ndim > 1 else v[idx]) for k, v in r.items()}
    assert np.array_equal(g["status"], o.status)
    pd, cd, c2 = compare_fit(g, o)
    assert pd <= TOL_MIXED_PARAMS and cd <= TOL_MIXED_COV, (pd, cd)


# ------------------------------------------------------------- scheduling --
SCHEDULES = [dict(sort_by_length=1), dict(persistent=1, waves_per_cu=4, chunks_per_grab=1),
             dict(sort_by_length=1, persistent=1, waves_per_cu=8, chunks_per_grab=4, graph=1),
             dict(graph=1), dict(store_by_plane=1),
             dict(sort_by_length=1, store_by_plane=1, persistent=1, waves_per_cu=4, graph=1)]


@pytest.mark.parametrize("records", [True, False])
@pytest.mark.parametrize("sched", SCHEDULES)
def test_schedules_give_identical_results(torch_cuda, sched, records):
    """Length sort, persistent work queue, graph replay and store slots by
    plane change where a track runs or keeps its states, never its result:
    bitwise equal to the default schedule, and the oracle on a ragged batch
    with holes (grid field).  With track states (the recording build) and
    without (the grid build without record writes)."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(3, n_tracks=5000)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b, track_states=records)
    opts = KalmanFitterOptions(**W)
    kf.fit_batch(dev, opts)
    ref = dev.results_numpy()
    kf.set_schedule(**sched)
    for _ in range(3):  # graph: capture, then two replays
        dev.clear_results()
        kf.fit_batch(dev, opts)
        got = dev.results_numpy()
        for k in ("params", "cov", "chi2", "status", "n_measured", "track_states"):
            if k in ref:
                assert np.array_equal(got[k], ref[k]), k
    # identical calls replay the one capture
    assert kf.graph_captures() == (1 if sched.get("graph") else 0)
    kf.close()
    o, _, _, _ = oracle.fit(b, c_options(**W))
    check_parity(ref, o)


@pytest.mark.parametrize("records", [True, False])
@pytest.mark.parametrize("cfg", [3, 5])
def test_store_by_plane_mixed_bitwise(torch_cuda, cfg, records):
    """Mixed precision: store slots by plane (default and length-sorted
    persistent schedules) give bitwise the results of the default schedule,
    with and without track states."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(cfg, n_tracks=4000)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b, track_states=records)
    opts = KalmanFitterOptions(precision="mixed", **W)
    kf.fit_batch(dev, opts)
    ref = dev.results_numpy()
    for sched in (dict(store_by_plane=1), dict(store_by_plane=1, sort_by_length=1, persistent=1)):
        kf.set_schedule(**sched)
        dev.clear_results()
        kf.fit_batch(dev, opts)
        got = dev.results_numpy()
        for k in ("params", "cov", "chi2", "status", "n_measured", "track_states"):
            if k in ref:
                assert np.array_equal(got[k], ref[k]), (sched, k)
    kf.close()


def test_store_by_plane_refused_with_lane_groups(torch_cuda):
    from gpukalmanfitter_amd.fitter import FitError, KalmanFitter
    b = G.generate(1, n_tracks=64)
    kf = KalmanFitter(b.planes, b.field)
    with pytest.raises(FitError, match="one lane per track"):
        kf.set_schedule(store_by_plane=1, lanes_per_track=2)
    kf.close()


@pytest.mark.parametrize("lanes", [2, 4])
@pytest.mark.parametrize("cfg", [1, 3, 5])
def test_lane_groups_parity(torch_cuda, lanes, cfg):
    """2 or 4 lanes per track (the counterpart of the reference's
    block-per-track kernel): FP64 within 1e-9 of the oracle with identical
    status bits, the track states too, and within 1e-12 of one lane per
    track; the mixed mode within its bars."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(cfg, n_tracks=3000)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b, track_states=True)
    one = {}
    for prec in ("fp64", "mixed"):
        kf.set_schedule()
        kf.fit_batch(dev, KalmanFitterOptions(precision=prec, **W))
        one[prec] = dev.results_numpy()
    kf.set_schedule(lanes_per_track=lanes, sort_by_length=1, graph=1)
    for prec in ("fp64", "mixed"):
        for _ in range(2):  # graph: capture, then a replay
            dev.clear_results()
            kf.fit_batch(dev, KalmanFitterOptions(precision=prec, **W))
            got = dev.results_numpy()
        assert np.array_equal(got["status"], one[prec]["status"])
        pd, cd, c2 = compare_fit(got, _as_result(one[prec]))
        # the group builds the float transport column by column: the same
        # operations, but the compiler may contract them differently
        bar = (1e-12, 1e-12) if prec == "fp64" else (1e-6, TOL_MIXED_COV)
        assert pd <= bar[0] and c2 <= bar[0] and cd <= bar[1], (prec, pd, cd, c2)
        print(f"lanes {lanes} config {cfg} {prec}: vs one lane per track {max(pd, cd, c2):.2g}, "
              f"bitwise {np.array_equal(got['params'], one[prec]['params'])}")
    o, _, _, _ = oracle.fit(b, c_options(**W), track_states=True)
    dev.clear_results()
    kf.fit_batch(dev, KalmanFitterOptions(precision="fp64", **W))
    got = dev.results_numpy()
    check_parity(got, o)
    smoothed = ((o.status & abi.KF_STATUS_SMOOTHED) != 0) & ((o.status & abi.KF_STATUS_SMOOTHER_NAN) == 0)
    worst = compare_states(got["track_states"], o.track_states, o.n_measured, smoothed)
    for k, v in worst.items():
        assert v <= state_bar(cfg, k), (k, v)
    kf.close()


def _as_result(r):
    """GPU results in the shape compare_fit takes as its reference."""
    class R:
        pass
    o = R()
    o.status, o.params, o.cov, o.chi2 = r["status"], r["params"], r["cov"], r["chi2"]
    return o


def test_graph_recaptures_only_when_arguments_change(torch_cuda):
    """The replay key holds every argument of the captured launches: equal
    calls replay, a changed option or output array captures again."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(3, n_tracks=3000)
    kf = KalmanFitter(b.planes, b.field)
    kf.set_schedule(graph=1)
    dev = DeviceBatch(b)
    opts = KalmanFitterOptions(**W)
    for _ in range(4):
        kf.fit_batch(dev, opts)
    assert kf.graph_captures() == 1
    ref = dev.results_numpy()
    kf.fit_batch(dev, KalmanFitterOptions(precision="mixed", **W))
    assert kf.graph_captures() == 2
    kf.fit_batch(dev, opts)
    kf.fit_batch(dev, opts)
    assert kf.graph_captures() == 3
    dev2 = DeviceBatch(b)
    kf.fit_batch(dev2, opts)
    assert kf.graph_captures() == 4
    for k in ("params", "cov", "status"):
        assert np.array_equal(dev2.results_numpy()[k], ref[k]), k
    kf.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("sched", [dict(), SCHEDULES[2]])
def test_memory_slices_give_identical_results(torch_cuda, precision, sched):
    """A memory budget below the batch's working set splits the fit into
    slices of positions (here 7-8 of at most 832 tracks, and the 64-track
    minimum); every track's result is bitwise that of the one-slice fit."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(3, n_tracks=5000)
    kf = KalmanFitter(b.planes, b.field)
    kf.set_schedule(**sched)
    dev = DeviceBatch(b, track_states=True)
    opts = KalmanFitterOptions(precision=precision, **W)
    kf.fit_batch(dev, opts)
    ref = dev.results_numpy()
    for budget in (32 << 20, 1):
        kf.set_memory_budget(budget)
        for _ in range(2):  # graph: capture, then a replay
            dev.clear_results()
            kf.fit_batch(dev, opts)
            got = dev.results_numpy()
            for k in ("params", "cov", "chi2", "status", "n_measured", "track_states"):
                assert np.array_equal(got[k], ref[k]), (budget, k)
    kf.close()


def test_autotuner_on_ragged_batch(torch_cuda):
    """The autotuner times its candidate schedules on the batch and keeps the
    fastest; the kept schedule reproduces the default results."""
    from gpukalmanfitter_amd.fitter import DeviceBatch, KalmanFitter, KalmanFitterOptions
    b = G.generate(3, n_tracks=200000)
    kf = KalmanFitter(b.planes, b.field)
    dev = DeviceBatch(b)
    opts = KalmanFitterOptions(precision="mixed", **W)
    kf.fit_batch(dev, opts)
    ref = dev.results_numpy()
    best, ms = kf.autotune(dev, opts)
    print("autotuned schedule", best, f"{ms:.3f} ms")
    assert best["lanes_per_track"] in (1, 2, 4) and ms > 0
    assert kf.schedule() == best
    kf.fit_batch(dev, opts)
    got = dev.results_numpy()
    for k in ("params", "cov", "status"):
        assert np.array_equal(got[k], ref[k]), k
    kf.close()


@pytest.mark.parametrize("cfg", [1, 3, 5])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_builds_with_and_without_records_agree(torch_cuda, cfg, precision):
    """A fit that asks for track states runs the build that writes them (and
    takes either field kind), one that does not runs the build for the
    batch's field kind without the record writes: the two run the same
    operations and agree to rounding in FP64; in the mixed mode each is
    within the bars of the oracle, so they are within twice the bars of each
    other.  Statuses and measurement counts are equal."""
    from types import SimpleNamespace
    b = G.generate(cfg, n_tracks=2000)
    opts = dict(W, precision=precision)
    plain = run_gpu(b, **opts)
    rec = run_gpu(b, track_states=True, **opts)
    assert np.array_equal(plain["status"], rec["status"])
    assert np.array_equal(plain["n_measured"], rec["n_measured"])
    ref = SimpleNamespace(status=rec["status"], params=rec["params"], cov=rec["cov"],
                          chi2=rec["chi2"])
    pd, cd, c2 = compare_fit(plain, ref)
    if precision == "fp64":
        assert max(pd, cd, c2) <= 1e-10, (pd, cd, c2)
    else:
        assert pd <= 2 * TOL_MIXED_PARAMS and cd <= 2 * TOL_MIXED_COV, (pd, cd)


def _padded_planes(planes, extra):
    """The plane table followed by `extra` planes that no sequence uses
    (copies of the last plane, 10 m further along x, with new geo ids)."""
    n = len(planes)
    out = (abi.kf_plane * (n + extra))()
    ctypes.memmove(out, planes, ctypes.sizeof(abi.kf_plane) * n)
    gmax = max(planes[i].geo_id for i in range(n))
    for j in range(extra):
        p = out[n + j]
        ctypes.memmove(ctypes.byref(p), ctypes.byref(planes[n - 1]), ctypes.sizeof(abi.kf_plane))
        p.center[0] = planes[n - 1].center[0] + 10000.0 + j
        p.geo_id = gmax + 1 + j
    return out


@pytest.mark.parametrize("cfg", [3, 5])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_plane_table_larger_than_lds(torch_cuda, cfg, precision):
    """A plane table of more planes than the kernels copy to LDS (256) is
    read from global memory by the build that takes either field kind: the
    fits match the oracle at the usual bars and the fit of the unpadded table
    to rounding (FP64)."""
    import dataclasses
    b = G.generate(cfg, n_tracks=2000)
    bp = dataclasses.replace(b, planes=_padded_planes(b.planes, 300))
    assert len(bp.planes) > 256
    opts = dict(W, precision=precision)
    g = run_gpu(bp, **opts)
    o, _, _, _ = oracle.fit(bp, c_options(**W))
    if precision == "fp64":
        check_parity(g, o)
        ref = run_gpu(b, **opts)
        from types import SimpleNamespace
        pd, cd, c2 = compare_fit(g, SimpleNamespace(status=ref["status"], params=ref["params"],
                                                    cov=ref["cov"], chi2=ref["chi2"]))
        assert max(pd, cd, c2) <= 1e-10, (pd, cd, c2)
    else:
        assert np.array_equal(g["status"], o.status)
        pd, cd, _ = compare_fit(g, o)
        assert pd <= TOL_MIXED_PARAMS and cd <= TOL_MIXED_COV, (pd, cd)
