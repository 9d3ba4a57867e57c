"""CPU oracle of the fitter -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers of liboracle.so (the FP64 C++ restatement of the reference fit
in kfo_fitter.hpp) and liboracle_count.so (the same source on a flop-counting
scalar).  Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of
bench.py use this package, always as the checker or the baseline, never as
the thing measured or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from gpukalmanfitter_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


class kfo_field(C.Structure):
    _fields_ = [("grid", C.c_int32), ("n", C.c_int32 * 3), ("b", C.c_double * 3),
                ("values", C.c_void_p), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("scale", C.c_double)]


def build():
    """Compiles liboracle.so / liboracle_count.so (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(counting=False):
    name = "liboracle_count.so" if counting else "liboracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(_HERE, name)
    if not counting and os.environ.get("KFO_LIBRARY"):
        # another build of the same sources (the sanitizer build)
        path = os.environ["KFO_LIBRARY"]
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    P = C.c_void_p
    lib.kfo_fit_batch.argtypes = [P, C.c_int32, C.POINTER(kfo_field), C.POINTER(abi.kf_tracks),
                                  C.POINTER(abi.kf_options), C.POINTER(abi.kf_results),
                                  C.c_int32, C.c_int32, P, P, P, P]
    lib.kfo_fit_batch.restype = C.c_int
    lib.kfo_time_fit_batch.argtypes = [P, C.c_int32, C.POINTER(kfo_field), C.POINTER(abi.kf_tracks),
                                       C.POINTER(abi.kf_options), C.POINTER(abi.kf_results), C.c_int32]
    lib.kfo_time_fit_batch.restype = C.c_double
    lib.kfo_propagate.argtypes = [P, C.c_int32, C.POINTER(kfo_field), P, P, P,
                                  C.POINTER(abi.kf_options), P, P, P, P, P, P]
    lib.kfo_propagate.restype = C.c_int
    lib.kfo_gain_update.argtypes = [P, P, P, P, P, P, P]
    lib.kfo_rts_smooth.argtypes = [C.c_int32, P, P, P, P, P, P, P]
    lib.kfo_rts_smooth.restype = C.c_int
    lib.kfo_material.argtypes = [C.c_int32] + [C.c_double] * 9
    lib.kfo_material.restype = C.c_double
    lib.kfo_field_get.argtypes = [C.POINTER(kfo_field), P, P]
    lib.kfo_material_interact.argtypes = [C.POINTER(abi.kf_plane), P, P, C.c_double, C.c_int32,
                                          C.c_double, P, C.c_int32, C.c_int32, C.c_int32,
                                          C.c_int32, P, P]
    lib.kfo_intersect.argtypes = [P, C.c_int64, P, P, C.c_int64, C.c_int32, P, C.c_int64,
                                  C.c_int32]
    lib.kfo_intersect.restype = C.c_int64
    lib.kfo_time_intersect.argtypes = [P, C.c_int64, P, P, C.c_int64, C.c_int32, C.c_int32]
    lib.kfo_time_intersect.restype = C.c_double
    _LIBS[name] = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def make_field(spec):
    """kfo_field from a generate.FieldSpec (keeps a reference to the grid)."""
    f = kfo_field()
    f.scale = float(spec.scale)
    if spec.kind == "grid":
        vals = np.ascontiguousarray(spec.values, dtype=np.float32)
        f.grid = 1
        for k in range(3):
            f.n[k] = vals.shape[k]
            f.lo[k] = spec.lo[k]
            f.hi[k] = spec.hi[k]
        f.values = vals.ctypes.data
        f._keep = vals
    else:
        for k in range(3):
            f.b[k] = spec.b[k]
    return f


class HostResults:
    """Host output buffers in the kf_results layout."""

    def __init__(self, n, max_states=0):
        self.params = np.zeros((6, n))
        self.cov = np.zeros((21, n))
        self.chi2 = np.zeros(n)
        self.ndf = np.zeros(n, dtype=np.int32)
        self.n_measured = np.zeros(n, dtype=np.int32)
        self.status = np.zeros(n, dtype=np.uint32)
        self.track_states = (np.zeros((max_states, abi.KF_TS_STRIDE, n))
                             if max_states > 0 else None)

    def c(self):
        r = abi.kf_results()
        for name in ("params", "cov", "chi2", "ndf", "n_measured", "status"):
            setattr(r, name, getattr(self, name).ctypes.data)
        r.track_states = self.track_states.ctypes.data if self.track_states is not None else None
        return r


def host_tracks(batch):
    t = abi.kf_tracks()
    t.n_tracks = batch.n_tracks
    t.n_hits = batch.n_hits
    keep = []
    for name in ("start_params", "start_cov", "perigee", "seq_first", "seq_count", "hit_offset",
                 "hit_geo_id", "hit_meas", "hit_cov"):
        arr = np.ascontiguousarray(getattr(batch, name))
        keep.append(arr)
        setattr(t, name, arr.ctypes.data)
    t._keep = keep
    return t


def fit(batch, options=None, track_states=False, n_threads=0, counting=False, planes=None,
        field=None):
    """Fits ``batch`` with the oracle.  Returns (HostResults, steps, flops, trans);
    with ``counting`` the results also carry ``flops32``, the part of each
    track's flops in the float class of the mixed mode, and ``flops_bwd``,
    the part from the smoother on (the HIP fitter's second kernel), and
    ``flops32_bwd``, the float-class part of the latter."""
    lib = _lib(counting)
    opts = options if options is not None else abi.default_options()
    n = batch.n_tracks
    max_states = int(opts.max_states) if opts.max_states > 0 else int(batch.max_states)
    res = HostResults(n, max_states if track_states else 0)
    steps = np.zeros(n, dtype=np.int32)
    flops = np.zeros(n, dtype=np.uint64)
    trans = np.zeros(n, dtype=np.uint64)
    classes = np.zeros((3, n), dtype=np.uint64)
    pl = planes if planes is not None else batch.planes
    fl = make_field(field if field is not None else batch.field)
    tr = host_tracks(batch)
    rc = lib.kfo_fit_batch(pl, len(pl), C.byref(fl), C.byref(tr), C.byref(opts), C.byref(res.c()),
                           max_states, n_threads, _p(steps), _p(flops), _p(trans), _p(classes))
    if rc != 0:
        raise RuntimeError(f"kfo_fit_batch failed ({rc})")
    res.flops32, res.flops_bwd, res.flops32_bwd = classes[0], classes[1], classes[2]
    return res, steps, flops, trans


def time_fit(batch, options=None, n_threads=0):
    """Seconds spent in the oracle's fit loop (OpenMP over tracks)."""
    lib = _lib(False)
    opts = options if options is not None else abi.default_options()
    res = HostResults(batch.n_tracks)
    fl = make_field(batch.field)
    tr = host_tracks(batch)
    return lib.kfo_time_fit_batch(batch.planes, len(batch.planes), C.byref(fl), C.byref(tr),
                                  C.byref(opts), C.byref(res.c()), n_threads)


def propagate(planes, field, start_par, start_cov, perigee, options=None):
    """Propagation without fitting; bound state at every plane reached."""
    lib = _lib(False)
    opts = options if options is not None else abi.default_options()
    n = len(planes)
    out = dict(params=np.zeros((n, 6)), cov=np.zeros((n, 6, 6)), jac=np.zeros((n, 6, 6)),
               pos=np.zeros((n, 3)), dir=np.zeros((n, 3)), path=np.zeros(n))
    sp = np.ascontiguousarray(start_par, dtype=np.float64)
    sc = np.ascontiguousarray(start_cov, dtype=np.float64).reshape(36)
    pc = np.ascontiguousarray(perigee, dtype=np.float64)
    fl = make_field(field)
    k = lib.kfo_propagate(planes, n, C.byref(fl), _p(sp), _p(sc), _p(pc), C.byref(opts),
                          _p(out["params"]), _p(out["cov"]), _p(out["jac"]), _p(out["pos"]),
                          _p(out["dir"]), _p(out["path"]))
    return {key: v[:k] for key, v in out.items()}


def gain_update(pred, pcov, meas, V):
    lib = _lib(False)
    f = np.zeros(6)
    fc = np.zeros((6, 6))
    chi2 = C.c_double()
    lib.kfo_gain_update(_p(np.ascontiguousarray(pred, dtype=np.float64)),
                        _p(np.ascontiguousarray(pcov, dtype=np.float64)),
                        _p(np.ascontiguousarray(meas, dtype=np.float64)),
                        _p(np.ascontiguousarray(V, dtype=np.float64)), _p(f), _p(fc),
                        C.byref(chi2))
    return f, fc, chi2.value


def rts_smooth(fpar, fcov, ppar, pcov, jac):
    lib = _lib(False)
    n = len(fpar)
    sp = np.zeros((n, 6))
    sc = np.zeros((n, 6, 6))
    args = [np.ascontiguousarray(x, dtype=np.float64) for x in (fpar, fcov, ppar, pcov, jac)]
    rc = lib.kfo_rts_smooth(n, *[_p(a) for a in args], _p(sp), _p(sc))
    return rc, sp, sc


def material(which, slab, mass, qop, q):
    """which: 'bethe' | 'highland' | 'landau_qop'; slab = (thickness, x0, l0, ar, z, molar_rho)."""
    w = {"bethe": 0, "highland": 1, "landau_qop": 2}[which]
    return _lib(False).kfo_material(w, *map(float, slab), float(mass), float(qop), float(q))


def material_interact(plane, pos, direction, p, q, mass, cov, nav=1, stage=0,
                      multiple_scattering=True, energy_loss=True):
    """One material-interactor call on ``plane`` (an abi.kf_plane): returns the
    momentum and the 6x6 bound covariance after it."""
    c = np.ascontiguousarray(cov, dtype=np.float64).reshape(36)
    out = np.zeros(36)
    p_out = C.c_double()
    _lib(False).kfo_material_interact(
        C.byref(plane), _p(np.ascontiguousarray(pos, dtype=np.float64)),
        _p(np.ascontiguousarray(direction, dtype=np.float64)), float(p), int(q), float(mass),
        _p(c), int(nav), int(stage), int(bool(multiple_scattering)), int(bool(energy_loss)),
        C.byref(p_out), _p(out))
    return p_out.value, out.reshape(6, 6)


def field_get(field, pos):
    fl = make_field(field)
    b = np.zeros(3)
    _lib(False).kfo_field_get(C.byref(fl), _p(np.ascontiguousarray(pos, dtype=np.float64)), _p(b))
    return b


def intersect(surfaces, ray_pos, ray_dir, bcheck=True, n_threads=0):
    """Every kept ray x surface intersection, in (ray, surface) order
    (kfo_intersect, oracle/kfo_intersect.cpp); surfaces is an array of
    gpukalmanfitter_amd.intersect.POLYGON_DTYPE, rays [3][n]."""
    from gpukalmanfitter_amd.intersect import HIT_DTYPE, POLYGON_DTYPE
    s = np.ascontiguousarray(surfaces, dtype=POLYGON_DTYPE)
    p = np.ascontiguousarray(ray_pos, dtype=np.float64)
    d = np.ascontiguousarray(ray_dir, dtype=np.float64)
    n_rays = p.shape[1]
    lib = _lib()
    total = lib.kfo_intersect(_p(s), len(s), _p(p), _p(d), n_rays, int(bool(bcheck)), None, 0,
                              n_threads)
    out = np.zeros(total, dtype=HIT_DTYPE)
    lib.kfo_intersect(_p(s), len(s), _p(p), _p(d), n_rays, int(bool(bcheck)), _p(out), total,
                      n_threads)
    return out


def time_intersect(surfaces, ray_pos, ray_dir, bcheck=True, n_threads=0):
    """Wall time (s) of the oracle intersection on n_threads threads."""
    from gpukalmanfitter_amd.intersect import POLYGON_DTYPE
    s = np.ascontiguousarray(surfaces, dtype=POLYGON_DTYPE)
    p = np.ascontiguousarray(ray_pos, dtype=np.float64)
    d = np.ascontiguousarray(ray_dir, dtype=np.float64)
    return _lib().kfo_time_intersect(_p(s), len(s), _p(p), _p(d), p.shape[1], int(bool(bcheck)),
                                     n_threads)
