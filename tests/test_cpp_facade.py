"""The C++ side of the boundary: include/kfmi.hpp (the reference's
KalmanFitter::fit, KalmanFitter.hpp:716-768) and the kf_run driver (the
reference's Run/GPU/KalmanFitterGPUTest.cu options)."""
import json
import os
import subprocess

import numpy as np
import pytest

from gpukalmanfitter_amd import abi, configs
from gpukalmanfitter_amd import generate as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KF_RUN = os.path.join(ROOT, "gpukalmanfitter_amd", "kf_run")


def run(args, cwd=None, timeout=300):
    return subprocess.run([KF_RUN] + args, cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


@pytest.mark.parametrize("cfg", [1, 2, 3, 5])
def test_driver_presets_match_python_workloads(built, cfg):
    p = run(["-c", str(cfg), "--print-config"])
    assert p.returncode == 0, p.stderr
    got = json.loads(p.stdout)
    spec = configs.CONFIGS[cfg]
    assert got.pop("grid") == (1 if spec["field"]["kind"] == "grid" else 0)
    for k, v in spec["gen"].items():
        assert got[k] == pytest.approx(v, rel=0, abs=0), k


@pytest.mark.parametrize("args", [["-d", "cpu"], ["-u", "1"], ["-c", "4"],
                                  ["-p", "fp16"], ["--bogus"]])
def test_driver_refuses_what_this_build_does_not_do(built, args):
    p = run(args)
    assert p.returncode == 1
    assert "ERROR" in p.stderr


@pytest.mark.gpu
def test_driver_fit_matches_oracle_and_per_track_facade(built, tmp_path):
    import oracle
    from parity import param_dev
    n = 3000
    p = run(["-c", "1", "-t", str(n), "-o", "1", "-n", "2", "-e", "2", "--per-track", "40"],
            cwd=tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    print(p.stdout)
    assert f"{n} of {n} tracks fitted" in p.stdout
    # the per-track facade call equals the batch fit bit for bit
    line = [ln for ln in p.stdout.splitlines() if "per-track fit of 40 tracks" in ln][0]
    assert "40 with the batch's status" in line and line.endswith("difference 0"), line
    csv = np.loadtxt(tmp_path / f"params-fitted_config1_{n}.csv", delimiter=",", skiprows=1)
    idx = csv[:, 0].astype(np.int64)
    b = G.generate(1, n_tracks=n)
    o, _, _, _ = oracle.fit(b, abi.default_options(**configs.WORKLOAD_FIT_OPTIONS))
    assert np.array_equal(idx, np.nonzero(o.status & abi.KF_STATUS_OK)[0])
    dev = param_dev(csv[:, 1:].T, o.params[:, idx], o.cov[:, idx])
    assert dev.max() <= 1e-9, dev.max()


@pytest.mark.gpu
def test_driver_with_copies_gives_the_same_fit(built, tmp_path):
    """kf_run --with-copies (each stream copies its chunk in, fits it and
    copies its results out inside the timed loop) fits the same tracks to the
    same results as the resident-input run, and every copied-out result array
    (params, cov, chi2, ndf, n_measured, status) equals the device's."""
    n = 20000
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    pa = run(["-c", "3", "-t", str(n), "-o", "1", "-n", "2", "-e", "3", "-p", "mixed"], cwd=a)
    pb = run(["-c", "3", "-t", str(n), "-o", "1", "-n", "2", "-e", "3", "-p", "mixed",
              "--with-copies"], cwd=b)
    assert pa.returncode == 0 and pb.returncode == 0, pa.stderr + pb.stderr
    assert "copied-out result elements differing from the device's: 0" in pb.stdout, pb.stdout
    name = f"params-fitted_config3_{n}.csv"
    assert (a / name).read_bytes() == (b / name).read_bytes()


@pytest.mark.gpu
def test_driver_lane_groups_fit_the_same_tracks(built, tmp_path):
    """kf_run -s 1 (the reference's block-per-track switch) fits every track
    on a group of 4 lanes: the same statuses and CSV rows as one lane per
    track, parameters within 1e-12 of them."""
    n = 5000
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    pa = run(["-c", "3", "-t", str(n), "-o", "1", "-n", "2"], cwd=a)
    pb = run(["-c", "3", "-t", str(n), "-o", "1", "-n", "2", "-s", "1"], cwd=b)
    assert pa.returncode == 0 and pb.returncode == 0, pa.stderr + pb.stderr
    assert "4 lane(s) per track" in pb.stdout, pb.stdout
    name = f"params-fitted_config3_{n}.csv"
    ca = np.loadtxt(a / name, delimiter=",", skiprows=1)
    cb = np.loadtxt(b / name, delimiter=",", skiprows=1)
    assert np.array_equal(ca[:, 0], cb[:, 0])
    assert np.abs(ca[:, 1:] - cb[:, 1:]).max() <= 1e-12 * np.abs(ca[:, 1:]).max()
