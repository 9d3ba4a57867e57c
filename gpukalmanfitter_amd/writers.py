"""Output and validation formats of the fit.

The reference writes (Run/Common/Writer.hpp) the fitted parameters as CSV
(``writeParamsCsv``, :16-229), the simulated hits and the fitted states as
OBJ polylines (``writeSimHitsObj``, ``writeStatesObj``), and a ROOT tree of
truth, fitted parameters, errors, residuals and pulls (``writeParamsRoot``),
which ``Performance/Scripts/perigeeParamResolution.C:27-76`` histograms and
fits with Gaussians.  ROOT is not available here: the tree is written as
NPZ or CSV with the same branch names, and the Gaussian fits are a binned
least-squares fit with Poisson bin errors over the same histogram ranges.
"""
import numpy as np

from . import abi

PARAM_CSV_HEADER = ("trackIdx, fit_BoundLoc0, fit_BoundLoc1, fit_BoundPhi, fit_BoundTheta, "
                    "fit_BoundQOverP, fit_BoundTime")
PARAM_KEYS = ("eLOC0", "eLOC1", "ePHI", "eTHETA", "eQOP", "eT")
# residual histogram half-ranges and the pull half-range of
# perigeeParamResolution.C:41-50
RESIDUAL_RANGE = {"eLOC0": 0.25, "eLOC1": 0.25, "ePHI": 0.08, "eTHETA": 0.08, "eQOP": 0.002,
                  "eT": 5000.0}
PULL_RANGE = 5.0
N_BINS = 100


def _ok(status):
    return (np.asarray(status).view(np.uint32) & abi.KF_STATUS_OK) != 0


def write_params_csv(path, params, status, precision=17):
    """writeParamsCsv: one row per successful track, ``trackIdx`` then the six
    fitted bound parameters ([6][n] array); values with `precision`
    significant digits (17 round-trips a double)."""
    params = np.asarray(params)
    ok = _ok(status)
    fmt = f"%.{int(precision)}g"
    with open(path, "w") as f:
        f.write(PARAM_CSV_HEADER + "\n")
        for it in np.nonzero(ok)[0]:
            f.write(str(int(it)) + "," + ",".join(fmt % v for v in params[:, it]) + "\n")


def read_params_csv(path):
    """(track indices, [6][m] parameters) from a file of write_params_csv
    or of the reference's writer."""
    with open(path) as f:
        header = f.readline().strip()
        if header != PARAM_CSV_HEADER:
            raise ValueError(f"{path}: not a fitted-parameter CSV")
        rows = [line.split(",") for line in f if line.strip()]
    idx = np.array([int(r[0]) for r in rows], dtype=np.int64)
    par = np.array([[float(v) for v in r[1:7]] for r in rows], dtype=np.float64).reshape(-1, 6)
    return idx, np.ascontiguousarray(par.T)


def plane_frames(planes):
    """geo id -> (centre [3], rotation [3][3], columns U, V, normal)."""
    out = {}
    for p in planes:
        out[int(p.geo_id)] = (np.array(p.center[:]), np.array(p.rot[:]).reshape(3, 3))
    return out


def local_to_global(planes, geo_ids, loc0, loc1):
    """Global positions [n][3] of local (loc0, loc1) points on the planes with
    the given geo ids (PlaneSurface::localToGlobal: centre + U loc0 + V loc1)."""
    frames = plane_frames(planes)
    geo_ids = np.asarray(geo_ids).astype(np.int64)
    out = np.empty((len(geo_ids), 3))
    for g in np.unique(geo_ids):
        if int(g) not in frames:
            raise ValueError(f"geo id {int(g)} is not in the plane table")
        c, R = frames[int(g)]
        sel = geo_ids == g
        out[sel] = c + np.outer(np.asarray(loc0)[sel], R[:, 0]) + np.outer(np.asarray(loc1)[sel],
                                                                           R[:, 1])
    return out


def _write_polylines(f, tracks):
    """OBJ vertices and line segments of consecutive vertices, one polyline
    per track (the numbering of Writer.hpp: 1-based, continuous)."""
    v = 0
    for pts in tracks:
        for x in pts:
            f.write("v %.17g %.17g %.17g\n" % (x[0], x[1], x[2]))
        for k in range(1, len(pts)):
            f.write("l %d %d\n" % (v + k, v + k + 1))
        v += len(pts)


def write_sim_hits_obj(path, workload):
    """writeSimHitsObj: the measured hits of every track as an OBJ polyline
    (global positions of the local measurements [2][m] on their planes)."""
    pos = local_to_global(workload.planes, workload.hit_geo_id, workload.hit_meas[0],
                          workload.hit_meas[1]) if workload.n_hits else np.zeros((0, 3))
    off = workload.hit_offset
    with open(path, "w") as f:
        _write_polylines(f, (pos[off[i]:off[i + 1]] for i in range(workload.n_tracks)))


STATE_BLOCKS = {"predicted": abi.KF_TS_PREDICTED, "filtered": abi.KF_TS_FILTERED,
                "smoothed": abi.KF_TS_SMOOTHED}


def state_positions(track_states, planes, n_states, parameters="smoothed"):
    """Global positions of the track states, a list of [k][3] arrays, from a
    ``track_states`` buffer [max_states][KF_TS_STRIDE][n] and the number of
    states of each track (``results["n_measured"]`` or the sequence lengths)."""
    blk = STATE_BLOCKS[parameters]
    ts = np.asarray(track_states)
    out = []
    for it in range(ts.shape[2]):
        k = int(n_states[it])
        s = ts[:k, :, it]
        out.append(local_to_global(planes, s[:, abi.KF_TS_GEO_ID], s[:, blk + 0], s[:, blk + 1])
                   if k else np.zeros((0, 3)))
    return out


def write_states_obj(path, track_states, status, planes, n_states, parameters="smoothed"):
    """writeStatesObj: the predicted, filtered or smoothed positions of the
    states of every successful track as an OBJ polyline."""
    pos = state_positions(track_states, planes, n_states, parameters)
    ok = _ok(status)
    with open(path, "w") as f:
        _write_polylines(f, (pos[i] for i in range(len(pos)) if ok[i]))


def params_table(params, cov, status, truth, vertex):
    """The columns of writeParamsRoot for the successful tracks: truth
    (t_charge, t_time, t_vx..t_pT, t_eta), fitted parameters (eLOC0_fit ..),
    their errors (err_*), residuals against the truth (0, 0, phi, theta,
    q/p, t) (res_*) and pulls (pull_*).  `truth` [6][n] as the workload's
    truth array, `vertex` [3][n] the truth origin."""
    ok = _ok(status)
    par = np.asarray(params)[:, ok]
    cov = np.asarray(cov)[:, ok]
    tr = np.asarray(truth)[:, ok]
    vx = np.asarray(vertex)[:, ok]
    phi, theta, qop, t0 = tr[2], tr[3], tr[4], tr[5]
    p = 1.0 / np.abs(qop)
    ux, uy, uz = np.cos(phi) * np.sin(theta), np.sin(phi) * np.sin(theta), np.cos(theta)
    tab = {"trackIdx": np.nonzero(ok)[0], "t_charge": np.sign(qop).astype(np.int32),
           "t_time": t0, "t_vx": vx[0], "t_vy": vx[1], "t_vz": vx[2], "t_px": p * ux,
           "t_py": p * uy, "t_pz": p * uz, "t_theta": theta, "t_phi": phi,
           "t_eta": -np.log(np.tan(theta / 2.0)), "t_pT": p * np.sin(theta)}
    truth_par = np.stack([np.zeros_like(phi), np.zeros_like(phi), phi, theta, qop, t0])
    for k, name in enumerate(PARAM_KEYS):
        err = np.sqrt(cov[abi.SYM_INDEX[(k, k)]])
        res = par[k] - truth_par[k]
        tab[f"{name}_fit"] = par[k]
        tab[f"err_{name}_fit"] = err
        tab[f"res_{name}"] = res
        tab[f"pull_{name}"] = res / err
    return tab


def write_params_table(path, table):
    """The table as NPZ (path ending .npz) or CSV (header = branch names)."""
    if str(path).endswith(".npz"):
        np.savez(path, **table)
        return
    names = list(table)
    with open(path, "w") as f:
        f.write(",".join(names) + "\n")
        cols = [np.asarray(table[n]) for n in names]
        for i in range(len(cols[0])):
            f.write(",".join(repr(c[i].item()) for c in cols) + "\n")


def gaussian_fit(values, half_range, bins=N_BINS):
    """Fit of A exp(-(x - mu)^2 / (2 sigma^2)) to the histogram of `values` in
    [-half_range, half_range] (`bins` bins), least squares with the bin
    errors sqrt(N) and the empty bins left out, as ROOT's TH1::Fit("gaus").
    Returns mean, sigma, their errors and the entries."""
    from scipy.optimize import curve_fit
    v = np.asarray(values, dtype=np.float64)
    v = v[np.isfinite(v)]
    counts, edges = np.histogram(v, bins=bins, range=(-half_range, half_range))
    x = 0.5 * (edges[1:] + edges[:-1])
    use = counts > 0
    out = {"entries": int(counts.sum()), "mean": np.nan, "sigma": np.nan, "mean_err": np.nan,
           "sigma_err": np.nan}
    if use.sum() < 4:
        return out
    inside = v[(v >= -half_range) & (v <= half_range)]
    p0 = [counts.max(), inside.mean(), max(inside.std(), edges[1] - edges[0])]

    def gaus(z, a, mu, sig):
        return a * np.exp(-0.5 * ((z - mu) / sig) ** 2)

    popt, pcov = curve_fit(gaus, x[use], counts[use], p0=p0, sigma=np.sqrt(counts[use]),
                           absolute_sigma=True, maxfev=10000)
    err = np.sqrt(np.diag(pcov))
    out.update(mean=float(popt[1]), sigma=float(abs(popt[2])), mean_err=float(err[1]),
               sigma_err=float(err[2]))
    return out


def resolution_report(table, residual_range=None):
    """perigeeParamResolution.C: per parameter, the Gaussian fit of the
    residuals and of the pulls (in +-5).  Residual windows: the script's
    (tuned to the reference's detector), overridden per parameter by the
    dict `residual_range`, or "auto": +-5 robust standard deviations
    (1.4826 x median absolute deviation) of each residual."""
    if residual_range == "auto":
        ranges = {}
        for name in PARAM_KEYS:
            r = np.asarray(table[f"res_{name}"])
            r = r[np.isfinite(r)]
            mad = np.median(np.abs(r - np.median(r))) if len(r) else 0.0
            ranges[name] = 5.0 * 1.4826 * mad if mad > 0 else RESIDUAL_RANGE[name]
    else:
        ranges = dict(RESIDUAL_RANGE, **(residual_range or {}))
    rep = {}
    for name in PARAM_KEYS:
        rep[name] = {"residual": dict(gaussian_fit(table[f"res_{name}"], ranges[name]),
                                      window=float(ranges[name])),
                     "pull": gaussian_fit(table[f"pull_{name}"], PULL_RANGE)}
    return rep
