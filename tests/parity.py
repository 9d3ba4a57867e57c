"""Parity metrics between the HIP fitter and the CPU oracle.

Relative error of a fitted quantity is taken against its natural scale so that
components that are zero by construction (d0 and z0 scatter around 0,
off-diagonal covariances pass through 0) do not divide by ~0:

* parameter k of a track: |x_gpu - x_ora| / max(|x_ora|, sigma_k), sigma_k the
  oracle's standard deviation of that parameter;
* covariance (i, j): |C_gpu - C_ora| / sqrt(C_ora_ii * C_ora_jj).
"""
import numpy as np

from gpukalmanfitter_amd import abi


def param_dev(g, o, cov_o):
    sig = np.sqrt(np.abs(cov_o[abi.DIAG_INDEX]))
    scale = np.maximum(np.abs(o), sig)
    scale = np.where(scale > 0, scale, 1.0)
    return np.abs(g - o) / scale


def cov_dev(g, o):
    d = np.abs(o[abi.DIAG_INDEX])
    out = np.zeros_like(o)
    for (i, j), k in abi.SYM_INDEX.items():
        if i <= j:
            s = np.sqrt(d[i] * d[j])
            s = np.where(s > 0, s, 1.0)
            out[k] = np.abs(g[k] - o[k]) / s
    return out


def compare_fit(gres, ores, mask=None):
    """Max deviations (params, cov, chi2) over tracks in mask (default: all finished)."""
    st = ores.status
    if mask is None:
        mask = (st & abi.KF_STATUS_FINISHED) != 0
    pd = param_dev(gres["params"][:, mask], ores.params[:, mask], ores.cov[:, mask])
    cd = cov_dev(gres["cov"][:, mask], ores.cov[:, mask])
    c2 = np.abs(gres["chi2"][mask] - ores.chi2[mask]) / np.maximum(np.abs(ores.chi2[mask]), 1e-300)
    return (float(pd.max()) if pd.size else 0.0, float(cd.max()) if cd.size else 0.0,
            float(c2.max()) if c2.size else 0.0)


def compare_states(gts, ots, n_meas, smoothed_mask, chi2_floor=1e-12):
    """Max normalised deviation of predicted / filtered / smoothed states,
    jacobians, path and chi2 over the measured states of every track.  A
    state's chi2 deviation is taken against max(chi2, chi2_floor)."""
    worst = {}
    n_states = gts.shape[0]
    for k in range(n_states):
        m = n_meas > k
        if not m.any():
            continue
        for name, off, mk in (("predicted", abi.KF_TS_PREDICTED, m),
                              ("filtered", abi.KF_TS_FILTERED, m),
                              ("smoothed", abi.KF_TS_SMOOTHED, m & smoothed_mask)):
            if not mk.any():
                continue
            gp, op = gts[k, off:off + 6][:, mk], ots[k, off:off + 6][:, mk]
            gc, oc = gts[k, off + 6:off + 27][:, mk], ots[k, off + 6:off + 27][:, mk]
            worst[name + "_par"] = max(worst.get(name + "_par", 0.0), float(param_dev(gp, op, oc).max()))
            worst[name + "_cov"] = max(worst.get(name + "_cov", 0.0), float(cov_dev(gc, oc).max()))
        gj = gts[k, abi.KF_TS_JACOBIAN:abi.KF_TS_JACOBIAN + 36][:, m]
        oj = ots[k, abi.KF_TS_JACOBIAN:abi.KF_TS_JACOBIAN + 36][:, m]
        # jacobian element (a, b) against the row scale max_b |J_ab|
        oj3 = oj.reshape(6, 6, -1)
        rowscale = np.maximum(np.abs(oj3).max(axis=1, keepdims=True), 1e-300)
        jd = np.abs(gj.reshape(6, 6, -1) - oj3) / rowscale
        worst["jacobian"] = max(worst.get("jacobian", 0.0), float(jd.max()))
        for name, f, floor in (("path", abi.KF_TS_PATH, 1e-12), ("chi2", abi.KF_TS_CHI2, chi2_floor)):
            g, o = gts[k, f][m], ots[k, f][m]
            worst[name] = max(worst.get(name, 0.0),
                              float((np.abs(g - o) / np.maximum(np.abs(o), floor)).max()))
        worst["geo_id"] = max(worst.get("geo_id", 0.0),
                              float(np.abs(gts[k, abi.KF_TS_GEO_ID][m] - ots[k, abi.KF_TS_GEO_ID][m]).max()))
    return worst
