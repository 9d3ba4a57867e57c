"""ctypes mirror of include/kf_abi.h (the C ABI of libkfmi.so).

Only plain ctypes structures and constants live here; loading the library and
calling it is done by :mod:`gpukalmanfitter_amd.fitter`.
"""
import ctypes as C

KF_ABI_VERSION = 4
KF_NOALLOC_WORKSPACE_BYTES = 256 << 20

KF_OK = 0
KF_E_INVALID = -1
KF_E_HIP = -2
KF_E_STATE = -3
KF_E_NOMEM = -4

KF_STATUS_OK = 0x001
KF_STATUS_SMOOTHED = 0x002
KF_STATUS_FINISHED = 0x004
KF_STATUS_RESULT_FAIL = 0x008
KF_STATUS_STEP_FAIL = 0x010
KF_STATUS_MAX_STEPS = 0x020
KF_STATUS_NAV_BREAK = 0x040
KF_STATUS_SMOOTHER_NAN = 0x080
KF_STATUS_BAD_INPUT = 0x100

KF_PRECISION_FP64 = 0
KF_PRECISION_MIXED = 1

KF_TS_PREDICTED = 0
KF_TS_FILTERED = 27
KF_TS_SMOOTHED = 54
KF_TS_JACOBIAN = 81
KF_TS_PATH = 117
KF_TS_CHI2 = 118
KF_TS_GEO_ID = 119
KF_TS_STRIDE = 120

# packed upper-triangle index of (i, j), i <= j
SYM_INDEX = {}
_k = 0
for _i in range(6):
    for _j in range(_i, 6):
        SYM_INDEX[(_i, _j)] = _k
        SYM_INDEX[(_j, _i)] = _k
        _k += 1
DIAG_INDEX = [SYM_INDEX[(i, i)] for i in range(6)]


class kf_plane(C.Structure):
    _fields_ = [
        ("center", C.c_double * 3),
        ("rot", C.c_double * 9),
        ("thickness", C.c_double),
        ("x0", C.c_double),
        ("l0", C.c_double),
        ("ar", C.c_double),
        ("z", C.c_double),
        ("molar_rho", C.c_double),
        ("half_u", C.c_double),
        ("half_v", C.c_double),
        ("geo_id", C.c_int32),
        ("reserved", C.c_int32),
    ]


class kf_options(C.Structure):
    _fields_ = [
        ("smoothing", C.c_int32),
        ("multiple_scattering", C.c_int32),
        ("energy_loss", C.c_int32),
        ("precision", C.c_int32),
        ("abs_pdg", C.c_int32),
        ("max_steps", C.c_int32),
        ("max_rk_trials", C.c_int32),
        ("max_states", C.c_int32),
        ("mass", C.c_double),
        ("tolerance", C.c_double),
        ("step_size_cutoff", C.c_double),
        ("max_step_size", C.c_double),
        ("target_tolerance", C.c_double),
        ("overstep_limit", C.c_double),
        ("outlier_chi2_cut", C.c_double),
    ]


class kf_tracks(C.Structure):
    _fields_ = [
        ("n_tracks", C.c_int64),
        ("n_hits", C.c_int64),
        ("start_params", C.c_void_p),
        ("start_cov", C.c_void_p),
        ("perigee", C.c_void_p),
        ("seq_first", C.c_void_p),
        ("seq_count", C.c_void_p),
        ("hit_offset", C.c_void_p),
        ("hit_geo_id", C.c_void_p),
        ("hit_meas", C.c_void_p),
        ("hit_cov", C.c_void_p),
    ]


class kf_schedule(C.Structure):
    _fields_ = [
        ("sort_by_length", C.c_int32),
        ("persistent", C.c_int32),
        ("waves_per_cu", C.c_int32),
        ("chunks_per_grab", C.c_int32),
        ("lanes_per_track", C.c_int32),
        ("graph", C.c_int32),
        ("store_by_plane", C.c_int32),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class kf_results(C.Structure):
    _fields_ = [
        ("params", C.c_void_p),
        ("cov", C.c_void_p),
        ("chi2", C.c_void_p),
        ("ndf", C.c_void_p),
        ("n_measured", C.c_void_p),
        ("status", C.c_void_p),
        ("track_states", C.c_void_p),
    ]


def default_options(**kw):
    """Reference defaults (Propagator.hpp:66-95, KalmanFitter.hpp:72)."""
    o = kf_options()
    o.smoothing = 1
    o.multiple_scattering = 1
    o.energy_loss = 1
    o.precision = KF_PRECISION_FP64
    o.abs_pdg = 13
    o.max_steps = 1000
    o.max_rk_trials = 10000
    o.max_states = 0
    o.mass = 105.6583755 * 1e-3
    o.tolerance = 1e-4
    o.step_size_cutoff = 0.0
    o.max_step_size = 1.7976931348623157e308
    o.target_tolerance = 1e-4
    o.overstep_limit = 0.01
    o.outlier_chi2_cut = 0.0
    for k, v in kw.items():
        setattr(o, k, v)
    return o
