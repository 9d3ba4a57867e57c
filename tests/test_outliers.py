"""The outlier finder (kf_options.outlier_chi2_cut) on the oracle.

The reference's finder is void (VoidKalmanComponents.hpp:66-80) and 0, the
default, keeps that.  With a cut, a measurement whose predicted-residual chi2
exceeds it keeps its prediction and leaves chi2 and ndf: the fit must then
equal the fit of the same tracks with that hit deleted (its plane a hole) --
the property that pins the semantics -- up to one difference of the two
paths: the outlier state binds the covariance to its plane (the free ->
bound -> free projection of boundState) where a hole does not, which moves
the result by < 1e-3 of its errors.  Hits displaced by 40 sigma make the
outliers; a cut of 30 on a 2-dof chi2 flags a clean hit with probability
3e-7."""
import numpy as np
import pytest

import oracle
from gpukalmanfitter_amd import abi, configs
from gpukalmanfitter_amd import generate as G

CUT = 30.0


def displaced(cfg, n, every=3, which=4, shift=2.0, seed=0):
    """Batch with hit `which` of every `every`-th track moved by `shift` mm
    in loc0; returns the batch and the mask of moved hits."""
    b = G.generate(cfg, n_tracks=n)
    moved = np.zeros(b.n_hits, dtype=bool)
    cnt = np.diff(b.hit_offset)
    for it in range(0, n, every):
        if cnt[it] > which:
            moved[b.hit_offset[it] + which] = True
    b.hit_meas = b.hit_meas.copy()
    b.hit_meas[0, moved] += shift
    return b, moved


def opts(cut, **kw):
    return abi.default_options(**dict(configs.WORKLOAD_FIT_OPTIONS, outlier_chi2_cut=cut, **kw))


@pytest.mark.parametrize("cfg", [1, 3])
def test_outliers_equal_holes(built, cfg):
    b, moved = displaced(cfg, 1500)
    track_of = np.repeat(np.arange(b.n_tracks), np.diff(b.hit_offset))
    bad = np.zeros(b.n_tracks, dtype=bool)
    bad[track_of[moved]] = True
    r_cut, _, _, _ = oracle.fit(b, opts(CUT), track_states=True)
    r_hole, _, _, _ = oracle.fit(b.drop_hits(moved), opts(0.0))
    r_void, _, _, _ = oracle.fit(b, opts(0.0))
    ok = (r_cut.status & abi.KF_STATUS_OK) != 0
    assert ok.mean() > 0.95
    assert np.array_equal(r_cut.status, r_hole.status)
    # one outlier on each displaced track, none elsewhere
    n_out = r_cut.n_measured - r_cut.ndf // 2
    assert np.array_equal(n_out[ok], bad[ok].astype(np.int32))
    assert np.array_equal(r_cut.n_measured[ok], np.diff(b.hit_offset)[ok])
    assert np.array_equal(r_cut.ndf[ok], r_hole.ndf[ok])
    # the outlier state is the one above the cut, and it kept its prediction
    ts = r_cut.track_states
    for it in np.nonzero(bad & ok)[0][:50]:
        k = r_cut.n_measured[it]
        chi2 = ts[:k, abi.KF_TS_CHI2, it]
        assert (chi2 > CUT).sum() == 1
        j = int(np.argmax(chi2))
        assert np.array_equal(ts[j, abi.KF_TS_FILTERED:abi.KF_TS_FILTERED + 27, it],
                              ts[j, abi.KF_TS_PREDICTED:abi.KF_TS_PREDICTED + 27, it])
    # equal to the fit with the hit deleted: parameters within 1e-3 of their
    # errors, covariance within 1e-3 in correlation units, chi2 within 0.01;
    # tracks without an outlier bit for bit
    sig = np.sqrt(r_hole.cov[abi.DIAG_INDEX][:, ok])
    assert (np.abs(r_cut.params[:, ok] - r_hole.params[:, ok]) / sig).max() < 1e-3
    norm = np.array([sig[i] * sig[j] for i in range(6) for j in range(i, 6)])
    assert (np.abs(r_cut.cov[:, ok] - r_hole.cov[:, ok]) / norm).max() < 1e-3
    assert np.abs(r_cut.chi2[ok] - r_hole.chi2[ok]).max() < 0.01
    clean = ok & ~bad
    assert np.array_equal(r_cut.params[:, clean], r_hole.params[:, clean])
    assert np.array_equal(r_cut.chi2[clean], r_hole.chi2[clean])
    # without the finder the displaced hits pull the fit: chi2 far larger
    assert r_void.chi2[bad & ok].mean() > 20 * r_cut.chi2[bad & ok].mean()


def test_cut_zero_is_the_void_finder(built):
    b, _ = displaced(1, 300)
    a, _, _, _ = oracle.fit(b, opts(0.0))
    c, _, _, _ = oracle.fit(b, abi.default_options(**configs.WORKLOAD_FIT_OPTIONS))
    for k in ("params", "cov", "chi2", "ndf", "status", "n_measured"):
        assert np.array_equal(getattr(a, k), getattr(c, k)), k
    assert np.array_equal(a.ndf, 2 * a.n_measured)


def test_drop_hits_makes_holes(built):
    b = G.generate(1, n_tracks=50)
    drop = np.zeros(b.n_hits, dtype=bool)
    drop[b.hit_offset[:-1] + 2] = True
    h = b.drop_hits(drop)
    assert h.n_hits == b.n_hits - 50
    assert np.array_equal(np.diff(h.hit_offset), np.diff(b.hit_offset) - 1)
    assert np.array_equal(h.seq_count, b.seq_count)
    assert not np.isin(b.hit_meas[0, drop], h.hit_meas[0]).any()
