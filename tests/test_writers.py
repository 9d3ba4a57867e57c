"""Output and validation formats (gpukalmanfitter_amd.writers): the
reference's fitted-parameter CSV, OBJ polylines of hits and states, the
truth/residual/pull table of writeParamsRoot and the Gaussian fits of
perigeeParamResolution.C, on oracle fits of the config-1 workload."""
import math
import types

import numpy as np
import pytest

import oracle
from gpukalmanfitter_amd import abi, configs
from gpukalmanfitter_amd import generate as G
from gpukalmanfitter_amd import writers as W


@pytest.fixture(scope="module")
def fit1(built):
    b = G.generate(1, n_tracks=4000)
    res, _, _, _ = oracle.fit(b, abi.default_options(**configs.WORKLOAD_FIT_OPTIONS),
                              track_states=True)
    return b, res


def test_params_csv_round_trip(tmp_path, fit1):
    b, res = fit1
    status = res.status.copy()
    status[[3, 17]] = 0  # failed tracks are left out
    path = tmp_path / "params.csv"
    W.write_params_csv(path, res.params, status)
    with open(path) as f:
        assert f.readline().rstrip("\n") == W.PARAM_CSV_HEADER
    idx, par = W.read_params_csv(path)
    keep = np.setdiff1d(np.arange(b.n_tracks), [3, 17])
    assert np.array_equal(idx, keep)
    assert np.array_equal(par, res.params[:, keep])  # 17 digits round-trip exactly


def test_obj_polylines_numbering(tmp_path):
    """Two tracks of 3 and 2 hits: vertices 1..5 and segments 1-2, 2-3, 4-5
    (the numbering writeSimHitsObj produces)."""
    rot = np.eye(3)[:, [1, 2, 0]].reshape(-1)  # U = y, V = z, normal = x
    planes = (abi.kf_plane * 2)()
    for i, x in enumerate((10.0, 20.0)):
        planes[i].center[:] = [x, 0.0, 0.0]
        planes[i].rot[:] = list(rot)
        planes[i].geo_id = 7 + i
    wl = types.SimpleNamespace(planes=planes, n_tracks=2, n_hits=5,
                               hit_offset=np.array([0, 3, 5]),
                               hit_geo_id=np.array([7, 8, 7, 7, 8], dtype=np.int32),
                               hit_meas=np.array([[1, 3, 5, 7, 9], [2, 4, 6, 8, 10]],
                                                 dtype=float))
    path = tmp_path / "hits.obj"
    W.write_sim_hits_obj(path, wl)
    lines = open(path).read().splitlines()
    v = [list(map(float, l.split()[1:])) for l in lines if l.startswith("v ")]
    seg = [l for l in lines if l.startswith("l ")]
    assert v == [[10, 1, 2], [20, 3, 4], [10, 5, 6], [10, 7, 8], [20, 9, 10]]
    assert seg == ["l 1 2", "l 2 3", "l 4 5"]


def test_state_positions_lie_on_their_planes(tmp_path, fit1):
    b, res = fit1
    frames = W.plane_frames(b.planes)
    for which in ("predicted", "filtered", "smoothed"):
        pos = W.state_positions(res.track_states[:, :, :50], b.planes, res.n_measured[:50], which)
        for it in range(50):
            geo = res.track_states[:res.n_measured[it], abi.KF_TS_GEO_ID, it].astype(int)
            for x, g in zip(pos[it], geo):
                c, R = frames[g]
                assert abs(np.dot(x - c, R[:, 2])) < 1e-9
    # smoothed positions sit within a few hit resolutions of the hits
    pos = W.state_positions(res.track_states[:, :, :50], b.planes, res.n_measured[:50])
    hits = W.local_to_global(b.planes, b.hit_geo_id, b.hit_meas[0], b.hit_meas[1])
    for it in range(50):
        h = hits[b.hit_offset[it]:b.hit_offset[it + 1]]
        assert len(h) == len(pos[it])
        assert np.abs(pos[it] - h).max() < 0.5
    path = tmp_path / "states.obj"
    W.write_states_obj(path, res.track_states[:, :, :50], res.status[:50], b.planes,
                       res.n_measured[:50])
    text = open(path).read().splitlines()
    assert sum(l.startswith("v ") for l in text) == int(res.n_measured[:50].sum())
    assert sum(l.startswith("l ") for l in text) == int(res.n_measured[:50].sum()) - 50


def test_gaussian_fit_recovers_a_normal():
    rng = np.random.default_rng(4)
    n = 200000
    v = rng.normal(0.3, 1.7, n)
    g = W.gaussian_fit(v, 5.0)
    assert g["entries"] == int(((v >= -5) & (v <= 5)).sum())
    assert abs(g["mean"] - 0.3) < 5 * g["mean_err"]
    assert abs(g["sigma"] - 1.7) < 5 * g["sigma_err"]
    assert 0.5 < g["mean_err"] / (1.7 / math.sqrt(n)) < 2.0
    assert 0.5 < g["sigma_err"] / (1.7 / math.sqrt(2 * n)) < 2.0


def test_pull_report_and_table(tmp_path, fit1):
    b, res = fit1
    tab = W.params_table(res.params, res.cov, res.status, b.truth, b.perigee)
    assert len(tab["trackIdx"]) == b.n_tracks
    # the truth columns restate the truth parameters
    assert np.allclose(tab["t_pT"] / np.sin(tab["t_theta"]), 1.0 / np.abs(b.truth[4]))
    assert np.allclose(np.hypot(tab["t_px"], tab["t_py"]), tab["t_pT"])
    assert np.allclose(tab["pull_eQOP"] * tab["err_eQOP_fit"], res.params[4] - b.truth[4])
    # the reference script's residual windows suit its detector, not this
    # 270 mm, 2 T telescope (q/p resolution ~0.008/GeV vs a +-0.002 window,
    # theta all in one bin): windows from the residuals themselves
    rep = W.resolution_report(tab, residual_range="auto")
    for name in W.PARAM_KEYS:
        pull = rep[name]["pull"]
        assert 0.9 < pull["sigma"] < 1.1, (name, pull)
        assert abs(pull["mean"]) < 0.1, (name, pull)
        r = rep[name]["residual"]
        assert r["entries"] > 0.99 * b.n_tracks, (name, r)
        assert 0.5 < r["sigma"] / tab[f"res_{name}"].std() < 1.5, (name, r)
    assert rep["eQOP"]["residual"]["window"] > 0.01
    # the script's own windows: loc0 fits, theta falls in one bin (no fit)
    ref = W.resolution_report(tab)
    assert ref["eLOC0"]["residual"]["sigma"] > 0 and ref["eLOC0"]["residual"]["window"] == 0.25
    assert np.isnan(ref["eTHETA"]["residual"]["sigma"])
    for suffix in (".npz", ".csv"):
        path = tmp_path / ("table" + suffix)
        W.write_params_table(path, tab)
        if suffix == ".npz":
            with np.load(path) as z:
                assert np.array_equal(z["pull_ePHI"], tab["pull_ePHI"])
        else:
            head = open(path).readline().strip().split(",")
            assert head == list(tab)
