"""Seeded synthetic workloads (libkfgen.so on the host, libkfgen_gpu.so on the GPU).

``generate(config_id, n_tracks=None)`` returns a :class:`Batch` holding the
plane table, the field and every per-track / per-hit array in the layout of
``kf_tracks`` (structure of arrays, CSR hits), plus the truth parameters at the
perigee for pull studies.
"""
import ctypes as C
import dataclasses
import math
import os

import numpy as np

from . import abi
from .configs import CONFIGS
from ._lib import load_library

_GEN_FIELDS = [
    ("n_tracks", C.c_int64), ("seed", C.c_uint64), ("n_planes", C.c_int32),
    ("min_planes", C.c_int32), ("max_planes", C.c_int32), ("material", C.c_int32),
    ("first_x", C.c_double), ("pitch", C.c_double), ("half_u", C.c_double),
    ("half_v", C.c_double), ("thickness", C.c_double), ("hole_prob", C.c_double),
    ("p_min", C.c_double), ("p_max", C.c_double), ("phi_min", C.c_double),
    ("phi_max", C.c_double), ("eta_min", C.c_double), ("eta_max", C.c_double),
    ("vertex_sigma", C.c_double), ("hit_sigma_u", C.c_double), ("hit_sigma_v", C.c_double),
    ("sigma_d0", C.c_double), ("sigma_z0", C.c_double), ("sigma_phi", C.c_double),
    ("sigma_theta", C.c_double), ("sigma_p_rel", C.c_double), ("sigma_t", C.c_double),
    ("mass", C.c_double), ("align_shift_sigma", C.c_double), ("align_rot_sigma", C.c_double),
    ("true_field_scale", C.c_double),
]


class kfg_config(C.Structure):
    _fields_ = _GEN_FIELDS


class kfg_field(C.Structure):
    _fields_ = [("grid", C.c_int32), ("n", C.c_int32 * 3), ("b", C.c_double * 3),
                ("values", C.c_void_p), ("lo", C.c_double * 3), ("hi", C.c_double * 3)]


@dataclasses.dataclass
class FieldSpec:
    """Constant field ``b`` or a trilinear grid ``values[nx, ny, nz, 3]`` (float32)."""
    kind: str
    b: tuple = (0.0, 0.0, 0.0)
    values: np.ndarray = None
    lo: tuple = (0.0, 0.0, 0.0)
    hi: tuple = (0.0, 0.0, 0.0)
    scale: float = 1.0

    @property
    def n(self):
        return tuple(self.values.shape[:3]) if self.values is not None else (0, 0, 0)


@dataclasses.dataclass
class Batch:
    planes: object            # ctypes array of abi.kf_plane (nominal geometry)
    true_planes: object       # ctypes array used by the simulation
    field: FieldSpec
    start_params: np.ndarray  # [6, n]
    start_cov: np.ndarray     # [21, n]
    perigee: np.ndarray       # [3, n]
    seq_first: np.ndarray     # [n] int32
    seq_count: np.ndarray     # [n] int32
    hit_offset: np.ndarray    # [n+1] int64
    hit_geo_id: np.ndarray    # [m] int32
    hit_meas: np.ndarray      # [2, m]
    hit_cov: np.ndarray       # [3, m]
    truth: np.ndarray         # [6, n] truth (0, 0, phi, theta, q/p, t0) at the perigee
    config_id: int = 0
    true_field_scale: float = 1.0

    @property
    def n_tracks(self):
        return int(self.start_params.shape[1])

    @property
    def n_hits(self):
        return int(self.hit_geo_id.shape[0])

    @property
    def max_states(self):
        return int(self.seq_count.max()) if self.n_tracks else 0

    def subset(self, idx):
        """Batch restricted to the tracks ``idx`` (a sorted index array)."""
        idx = np.asarray(idx, dtype=np.int64)
        counts = self.hit_offset[idx + 1] - self.hit_offset[idx]
        off = np.zeros(len(idx) + 1, dtype=np.int64)
        np.cumsum(counts, out=off[1:])
        sel = (np.repeat(self.hit_offset[idx] - off[:-1], counts) +
               np.arange(off[-1], dtype=np.int64))
        return dataclasses.replace(
            self,
            start_params=np.ascontiguousarray(self.start_params[:, idx]),
            start_cov=np.ascontiguousarray(self.start_cov[:, idx]),
            perigee=np.ascontiguousarray(self.perigee[:, idx]),
            seq_first=np.ascontiguousarray(self.seq_first[idx]),
            seq_count=np.ascontiguousarray(self.seq_count[idx]),
            hit_offset=off,
            hit_geo_id=np.ascontiguousarray(self.hit_geo_id[sel]),
            hit_meas=np.ascontiguousarray(self.hit_meas[:, sel]),
            hit_cov=np.ascontiguousarray(self.hit_cov[:, sel]),
            truth=np.ascontiguousarray(self.truth[:, idx]),
        )

    def drop_hits(self, drop):
        """Batch without the hits where the boolean mask ``drop`` [m] is set:
        their planes stay in the surface sequences and become holes."""
        keep = ~np.asarray(drop, dtype=bool)
        track = np.repeat(np.arange(self.n_tracks), np.diff(self.hit_offset))
        counts = np.bincount(track[keep], minlength=self.n_tracks)
        off = np.zeros(self.n_tracks + 1, dtype=np.int64)
        np.cumsum(counts, out=off[1:])
        return dataclasses.replace(
            self, hit_offset=off,
            hit_geo_id=np.ascontiguousarray(self.hit_geo_id[keep]),
            hit_meas=np.ascontiguousarray(self.hit_meas[:, keep]),
            hit_cov=np.ascontiguousarray(self.hit_cov[:, keep]))


def _lib():
    lib = load_library("libkfgen.so")
    if not getattr(lib, "_kfg_typed", False):
        P = C.c_void_p
        lib.kfg_planes.argtypes = [C.POINTER(kfg_config), P]
        lib.kfg_misalign.argtypes = [C.POINTER(kfg_config), P, C.c_uint64]
        lib.kfg_simulate.argtypes = [C.POINTER(kfg_config), P, C.POINTER(kfg_field)] + [P] * 10 + [C.c_int32]
        lib.kfg_simulate.restype = C.c_int
        lib.kfg_pack_hits.argtypes = [C.POINTER(kfg_config), P, P, P, P, P, C.c_double, C.c_double,
                                      P, P, P, C.c_int32]
        lib.kfg_fill_grid.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, P, P, C.c_double, C.c_double]
        lib.kfg_philox.argtypes = [P, P]
        lib.kfg_landau.argtypes = [C.c_uint64, C.c_int64, P]
        lib.kfg_gamma.argtypes = [C.c_double, C.c_uint64, C.c_int64, P]
        lib._kfg_typed = True
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def make_field(spec):
    """FieldSpec from a configs.py field description."""
    if spec["kind"] == "const":
        return FieldSpec(kind="const", b=tuple(spec["b"]))
    nx, ny, nz = spec["n"]
    vals = np.zeros((nx, ny, nz, 3), dtype=np.float32)
    lo = np.array(spec["lo"], dtype=np.float64)
    hi = np.array(spec["hi"], dtype=np.float64)
    _lib().kfg_fill_grid(_ptr(vals), nx, ny, nz, _ptr(lo), _ptr(hi), spec["b0"], spec["br"])
    return FieldSpec(kind="grid", values=vals, lo=tuple(lo), hi=tuple(hi))


def _setup(config_id, n_tracks, seed, overrides):
    """Generator configuration, nominal and true plane tables, field."""
    spec = CONFIGS[config_id]
    gen = dict(spec["gen"])
    if n_tracks is not None:
        gen["n_tracks"] = int(n_tracks)
    if seed is not None:
        gen["seed"] = int(seed)
    gen.update(overrides)
    field = make_field(spec["field"])
    if "field_scale_sigma" in spec and "true_field_scale" not in overrides:
        # seeded true field scale 1 +- sigma (config 4)
        rs = np.random.default_rng(gen["seed"])
        gen["true_field_scale"] = 1.0 + spec["field_scale_sigma"] * float(rs.standard_normal())
    cfg = kfg_config(**gen)
    lib = _lib()
    planes = (abi.kf_plane * cfg.n_planes)()
    lib.kfg_planes(C.byref(cfg), planes)
    true_planes = (abi.kf_plane * cfg.n_planes)()
    C.memmove(true_planes, planes, C.sizeof(planes))
    if cfg.align_shift_sigma > 0 or cfg.align_rot_sigma > 0:
        lib.kfg_misalign(C.byref(cfg), true_planes, C.c_uint64(cfg.seed ^ 0xA11C))

    fld = kfg_field()
    fld.grid = 1 if field.kind == "grid" else 0
    if field.kind == "grid":
        for k in range(3):
            fld.n[k] = field.n[k]
            fld.lo[k] = field.lo[k]
            fld.hi[k] = field.hi[k]
        fld.values = field.values.ctypes.data
    else:
        for k in range(3):
            fld.b[k] = field.b[k]
    return cfg, planes, true_planes, field, fld


def generate(config_id, n_tracks=None, seed=None, n_threads=0, **overrides):
    """Generates the workload of config ``config_id`` (1..5)."""
    cfg, planes, true_planes, field, fld = _setup(config_id, n_tracks, seed, overrides)
    lib = _lib()
    n = cfg.n_tracks
    maxp = cfg.max_planes
    start_params = np.zeros((6, n))
    start_cov = np.zeros((21, n))
    perigee = np.zeros((3, n))
    seq_first = np.zeros(n, dtype=np.int32)
    seq_count = np.zeros(n, dtype=np.int32)
    n_hits = np.zeros(n, dtype=np.int32)
    su = np.zeros((n, maxp))
    sv = np.zeros((n, maxp))
    sg = np.zeros((n, maxp), dtype=np.int32)
    truth = np.zeros((6, n))
    rc = lib.kfg_simulate(C.byref(cfg), true_planes, C.byref(fld), _ptr(start_params),
                          _ptr(start_cov), _ptr(perigee), _ptr(seq_first), _ptr(seq_count),
                          _ptr(n_hits), _ptr(su), _ptr(sv), _ptr(sg), _ptr(truth), n_threads)
    if rc != 0:
        raise RuntimeError("kfg_simulate failed")
    hit_offset = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(n_hits, out=hit_offset[1:])
    m = int(hit_offset[-1])
    hit_geo = np.zeros(m, dtype=np.int32)
    hit_meas = np.zeros((2, m))
    hit_cov = np.zeros((3, m))
    lib.kfg_pack_hits(C.byref(cfg), _ptr(n_hits), _ptr(hit_offset), _ptr(su), _ptr(sv), _ptr(sg),
                      cfg.hit_sigma_u, cfg.hit_sigma_v, _ptr(hit_geo), _ptr(hit_meas),
                      _ptr(hit_cov), n_threads)
    del su, sv, sg
    return Batch(planes=planes, true_planes=true_planes, field=field,
                 start_params=start_params, start_cov=start_cov, perigee=perigee,
                 seq_first=seq_first, seq_count=seq_count, hit_offset=hit_offset,
                 hit_geo_id=hit_geo, hit_meas=hit_meas, hit_cov=hit_cov, truth=truth,
                 config_id=config_id, true_field_scale=cfg.true_field_scale)


def philox4x32_10(counter, key):
    """Philox4x32-10 of the generator (for its known-answer test)."""
    c = np.array(counter, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    _lib().kfg_philox(_ptr(c), _ptr(k))
    return [int(x) for x in c]


def landau_samples(n, seed=1):
    """n standard Landau variates of the generator's sampler (the energy-loss
    straggling of the material simulation), for its distribution test."""
    out = np.empty(int(n), dtype=np.float64)
    _lib().kfg_landau(C.c_uint64(seed), int(n), _ptr(out))
    return out


def gamma_samples(k, n, seed=1):
    """n Gamma(k, 1) variates of the generator's sampler (the Bethe-Heitler
    loss of the material simulation), for its distribution test."""
    out = np.empty(int(n), dtype=np.float64)
    _lib().kfg_gamma(float(k), C.c_uint64(seed), int(n), _ptr(out))
    return out


def _gpu_lib():
    lib = load_library("libkfgen_gpu.so")
    if not getattr(lib, "_kfg_typed", False):
        P = C.c_void_p
        lib.kfg_simulate_device.argtypes = ([C.POINTER(kfg_config), P, C.POINTER(kfg_field)] +
                                            [P] * 11)
        lib.kfg_simulate_device.restype = C.c_int
        lib.kfg_offsets_scratch_bytes.argtypes = [C.c_int64]
        lib.kfg_offsets_scratch_bytes.restype = C.c_int64
        lib.kfg_offsets_device.argtypes = [C.c_int64, P, P, P, P]
        lib.kfg_offsets_device.restype = C.c_int
        lib.kfg_pack_hits_device.argtypes = ([C.POINTER(kfg_config), P, P, C.c_int64, P, P, P,
                                              C.c_double, C.c_double, P, P, P, P])
        lib.kfg_pack_hits_device.restype = C.c_int
        lib._kfg_typed = True
    return lib


@dataclasses.dataclass
class DeviceWorkload:
    """A workload generated on the GPU: the host-side geometry and field, and
    the per-track / per-hit arrays as device tensors in the layout of
    ``kf_tracks`` (``tensors``), plus the truth parameters."""
    planes: object
    true_planes: object
    field: FieldSpec
    tensors: dict
    n_tracks: int
    n_hits: int
    max_states: int
    config_id: int = 0
    true_field_scale: float = 1.0

    def device_batch(self, track_states=False):
        from .fitter import DeviceBatch
        return DeviceBatch.from_tensors(self.tensors, self.n_tracks, self.n_hits,
                                        self.max_states, track_states=track_states)

    def to_host(self):
        """The same workload as a host :class:`Batch` (copies)."""
        h = {k: v.cpu().numpy() for k, v in self.tensors.items()}
        return Batch(planes=self.planes, true_planes=self.true_planes, field=self.field,
                     start_params=h["start_params"], start_cov=h["start_cov"],
                     perigee=h["perigee"], seq_first=h["seq_first"], seq_count=h["seq_count"],
                     hit_offset=h["hit_offset"], hit_geo_id=h["hit_geo_id"],
                     hit_meas=h["hit_meas"], hit_cov=h["hit_cov"], truth=h["truth"],
                     config_id=self.config_id, true_field_scale=self.true_field_scale)


def generate_device(config_id, n_tracks=None, seed=None, device="cuda:0", stream=None,
                    **overrides):
    """Generates the workload of config ``config_id`` on the GPU
    (libkfgen_gpu.so): the same tracks as :func:`generate` (identical random
    streams; values agree to the rounding of the device math library), with
    every array left in device memory.  Synchronises once to size the hit
    arrays."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("generate_device needs a HIP device")
    cfg, planes, true_planes, field, fld = _setup(config_id, n_tracks, seed, overrides)
    lib = _gpu_lib()
    dev = torch.device(device)
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    sp = C.c_void_p(s.cuda_stream)
    n = cfg.n_tracks
    maxp = cfg.max_planes
    f64 = dict(dtype=torch.float64, device=dev)
    i32 = dict(dtype=torch.int32, device=dev)
    nn = max(n, 1)
    t = {"start_params": torch.empty((6, nn), **f64), "start_cov": torch.empty((21, nn), **f64),
         "perigee": torch.empty((3, nn), **f64), "seq_first": torch.empty(nn, **i32),
         "seq_count": torch.empty(nn, **i32), "truth": torch.empty((6, nn), **f64)}
    n_hits = torch.zeros(nn, **i32)
    su = torch.empty((nn, maxp), **f64)
    sv = torch.empty((nn, maxp), **f64)
    sg = torch.empty((nn, maxp), **i32)
    planes_d = torch.frombuffer(bytearray(bytes(true_planes)), dtype=torch.uint8).to(dev)
    grid_d = None
    if field.kind == "grid":
        grid_d = torch.from_numpy(np.ascontiguousarray(field.values)).to(dev)
        fld.values = grid_d.data_ptr()
    with torch.cuda.stream(s):
        rc = lib.kfg_simulate_device(
            C.byref(cfg), planes_d.data_ptr(), C.byref(fld),
            *[t[k].data_ptr() for k in ("start_params", "start_cov", "perigee", "seq_first",
                                         "seq_count")],
            n_hits.data_ptr(), su.data_ptr(), sv.data_ptr(), sg.data_ptr(),
            t["truth"].data_ptr(), sp)
        if rc != 0:
            raise RuntimeError(f"kfg_simulate_device failed ({rc})")
        hit_offset = torch.empty(n + 1, dtype=torch.int64, device=dev)
        scratch = torch.empty(max(int(lib.kfg_offsets_scratch_bytes(n)), 8), dtype=torch.uint8,
                              device=dev)
        rc = lib.kfg_offsets_device(n, n_hits.data_ptr(), hit_offset.data_ptr(),
                                    scratch.data_ptr(), sp)
        if rc != 0:
            raise RuntimeError(f"kfg_offsets_device failed ({rc})")
        m = int(hit_offset[n].item())
        mm = max(m, 1)
        hit_geo = torch.empty(mm, **i32)
        hit_meas = torch.empty((2, mm), **f64)
        hit_cov = torch.empty((3, mm), **f64)
        rc = lib.kfg_pack_hits_device(C.byref(cfg), n_hits.data_ptr(), hit_offset.data_ptr(), m,
                                      su.data_ptr(), sv.data_ptr(), sg.data_ptr(),
                                      cfg.hit_sigma_u, cfg.hit_sigma_v, hit_geo.data_ptr(),
                                      hit_meas.data_ptr(), hit_cov.data_ptr(), sp)
        if rc != 0:
            raise RuntimeError(f"kfg_pack_hits_device failed ({rc})")
    s.synchronize()
    del su, sv, sg, scratch, planes_d, grid_d
    t.update(hit_offset=hit_offset, hit_geo_id=hit_geo[:m] if m else hit_geo,
             hit_meas=hit_meas[:, :m].contiguous() if m else hit_meas,
             hit_cov=hit_cov[:, :m].contiguous() if m else hit_cov)
    max_states = int(t["seq_count"].max().item()) if n else 0
    return DeviceWorkload(planes=planes, true_planes=true_planes, field=field, tensors=t,
                          n_tracks=n, n_hits=m, max_states=max_states, config_id=config_id,
                          true_field_scale=cfg.true_field_scale)
