"""Host side of the MI355X Kalman fitter, mirroring the reference's fit API.

The reference fits one track with
``KalmanFitter::fit(sourcelinks, startParameters, kfOptions, kfResult,
surfaceSequence, surfaceSequenceSize) -> bool``
(include/Fitter/KalmanFitter.hpp:716-768) and batches it in a CUDA kernel
(Run/GPU/Kernels.cu:6-94).  Here :class:`KalmanFitter` does both through the C
ABI of ``libkfmi.so`` (include/kf_abi.h): :meth:`KalmanFitter.fit` keeps the
per-track signature and meaning, :meth:`KalmanFitter.fit_batch` fits a whole
:class:`~gpukalmanfitter_amd.generate.Batch` already resident on the GPU.

Device memory and streams come from PyTorch; every compute call goes to the
HIP library and raises :class:`FitError` when it fails.  There is no CPU
fallback: without ``libkfmi.so`` or a GPU the fitter refuses to run.
"""
import ctypes as C
import dataclasses
import math

import numpy as np

from . import abi
from ._lib import load_library


class FitError(RuntimeError):
    """A C-ABI call of libkfmi.so returned an error code."""


def _lib():
    lib = load_library("libkfmi.so")
    if not getattr(lib, "_kf_typed", False):
        P = C.c_void_p
        lib.kf_abi_version.restype = C.c_int
        lib.kf_options_default.argtypes = [C.POINTER(abi.kf_options)]
        lib.kf_ctx_create.argtypes = [C.c_int, C.POINTER(P)]
        lib.kf_ctx_create.restype = C.c_int
        lib.kf_ctx_destroy.argtypes = [P]
        lib.kf_ctx_destroy.restype = C.c_int
        lib.kf_last_error.argtypes = [P]
        lib.kf_last_error.restype = C.c_char_p
        lib.kf_set_geometry.argtypes = [P, P, C.c_int32]
        lib.kf_set_geometry.restype = C.c_int
        lib.kf_set_field_constant.argtypes = [P, C.c_double, C.c_double, C.c_double]
        lib.kf_set_field_constant.restype = C.c_int
        lib.kf_set_field_grid.argtypes = [P, P, C.c_int32, C.c_int32, C.c_int32, P, P]
        lib.kf_set_field_grid.restype = C.c_int
        lib.kf_set_field_scale.argtypes = [P, C.c_double]
        lib.kf_set_field_scale.restype = C.c_int
        lib.kf_fit.argtypes = [P, C.POINTER(abi.kf_tracks), C.POINTER(abi.kf_options),
                               C.POINTER(abi.kf_results), P]
        lib.kf_fit.restype = C.c_int
        lib.kf_last_kernel_ms.argtypes = [P]
        lib.kf_last_kernel_ms.restype = C.c_double
        lib.kf_last_phase_ms.argtypes = [P, P, P]
        lib.kf_last_phase_ms.restype = C.c_int
        lib.kf_fit_noalloc.argtypes = [P, C.POINTER(abi.kf_tracks), C.POINTER(abi.kf_options),
                                       C.POINTER(abi.kf_results), P]
        lib.kf_fit_noalloc.restype = C.c_int
        lib.kf_graph_captures.argtypes = [P]
        lib.kf_graph_captures.restype = C.c_int64
        lib.kf_chi2_gradient.argtypes = [P, C.POINTER(abi.kf_tracks), C.POINTER(abi.kf_options),
                                         P, P, P, P, P, P]
        lib.kf_chi2_gradient.restype = C.c_int
        lib.kf_schedule_default.argtypes = [C.POINTER(abi.kf_schedule)]
        lib.kf_schedule_default.restype = None
        lib.kf_set_schedule.argtypes = [P, C.POINTER(abi.kf_schedule)]
        lib.kf_set_schedule.restype = C.c_int
        lib.kf_get_schedule.argtypes = [P, C.POINTER(abi.kf_schedule)]
        lib.kf_get_schedule.restype = C.c_int
        lib.kf_autotune.argtypes = [P, C.POINTER(abi.kf_tracks), C.POINTER(abi.kf_options),
                                    C.POINTER(abi.kf_results), P, C.c_int32,
                                    C.POINTER(abi.kf_schedule), C.POINTER(C.c_double)]
        lib.kf_autotune.restype = C.c_int
        lib.kf_set_memory_budget.argtypes = [P, C.c_uint64]
        lib.kf_set_memory_budget.restype = C.c_int
        lib._kf_typed = True
    return lib


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise FitError("no GPU: the MI355X fitter runs only on a HIP device")
    return torch


@dataclasses.dataclass
class KalmanFitterOptions:
    """KalmanFitterOptions (KalmanFitter.hpp:48-96) + the PropagatorOptions
    fields the fit uses (Propagator.hpp:52-111).  ``precision`` selects the
    plain FP64 fitter or the FP32-mixed one."""
    smoothing: bool = True
    multiple_scattering: bool = True
    energy_loss: bool = True
    precision: str = "fp64"  # "fp64" | "mixed"
    abs_pdg: int = 13
    mass: float = 105.6583755e-3
    max_steps: int = 1000
    max_rk_trials: int = 10000
    tolerance: float = 1e-4
    step_size_cutoff: float = 0.0
    max_step_size: float = 1.7976931348623157e308
    target_tolerance: float = 1e-4
    overstep_limit: float = 0.01
    max_states: int = 0
    outlier_chi2_cut: float = 0.0  # 0 = VoidOutlierFinder



    def to_c(self):
        o = abi.default_options()
        o.smoothing = int(self.smoothing)
        o.multiple_scattering = int(self.multiple_scattering)
        o.energy_loss = int(self.energy_loss)
        if self.precision == "fp64":
            o.precision = abi.KF_PRECISION_FP64
        elif self.precision == "mixed":
            o.precision = abi.KF_PRECISION_MIXED
        else:
            raise ValueError(f"precision must be 'fp64' or 'mixed', not {self.precision!r}")
        o.abs_pdg = self.abs_pdg
        o.mass = self.mass
        o.max_steps = self.max_steps
        o.max_rk_trials = self.max_rk_trials
        o.tolerance = self.tolerance
        o.step_size_cutoff = self.step_size_cutoff
        o.max_step_size = self.max_step_size
        o.target_tolerance = self.target_tolerance
        o.overstep_limit = self.overstep_limit
        o.max_states = self.max_states
        o.outlier_chi2_cut = self.outlier_chi2_cut
        return o


@dataclasses.dataclass
class PixelSourceLink:
    """2-D measurement on a plane (include/EventData/PixelSourceLink.hpp:22-75)."""
    values: tuple          # (loc0, loc1)
    cov: tuple             # ((V00, V01), (V10, V11))
    geometry_id: int


@dataclasses.dataclass
class BoundParameters:
    """Bound parameters (loc0, loc1, phi, theta, q/p, t) + 6x6 covariance."""
    parameters: np.ndarray
    covariance: np.ndarray


@dataclasses.dataclass
class KalmanFitterResult:
    """KalmanFitterResult (KalmanFitter.hpp:98-120)."""
    fitted_parameters: BoundParameters = None
    measurement_states: int = 0
    smoothed: bool = False
    finished: bool = False
    result: bool = True
    chi2: float = 0.0
    ndf: int = 0
    status: int = 0
    fitted_states: np.ndarray = None  # [n_states, KF_TS_STRIDE]


def unpack_sym(packed):
    """[21, ...] packed upper triangle -> [..., 6, 6]."""
    packed = np.asarray(packed)
    out = np.empty(packed.shape[1:] + (6, 6), dtype=packed.dtype)
    for (i, j), k in abi.SYM_INDEX.items():
        out[..., i, j] = packed[k]
    return out


def pack_sym(m):
    """[..., 6, 6] -> [21, ...] upper triangle."""
    m = np.asarray(m)
    out = np.empty((21,) + m.shape[:-2], dtype=m.dtype)
    k = 0
    for i in range(6):
        for j in range(i, 6):
            out[k] = m[..., i, j]
            k += 1
    return out


class DeviceBatch:
    """A Batch copied to GPU memory (torch tensors) plus output buffers."""

    _INPUTS = ("start_params", "start_cov", "perigee", "seq_first", "seq_count",
               "hit_offset", "hit_geo_id", "hit_meas", "hit_cov")

    def __init__(self, batch, device="cuda:0", track_states=False, max_states=None):
        torch = _torch()
        self.device = torch.device(device)
        self.n_tracks = batch.n_tracks
        self.n_hits = batch.n_hits
        self.t = {}
        for name in self._INPUTS:
            arr = np.ascontiguousarray(getattr(batch, name))
            if arr.size == 0:
                arr = np.zeros(1, dtype=arr.dtype)
            self.t[name] = torch.from_numpy(arr).to(self.device)
        self._outputs(track_states, max_states if max_states else max(batch.max_states, 1))

    @classmethod
    def from_tensors(cls, tensors, n_tracks, n_hits, max_states, track_states=False):
        """A batch whose inputs are already device tensors (the GPU generator's
        output), in the layout of kf_tracks; no copy."""
        self = cls.__new__(cls)
        self.t = {name: tensors[name] for name in cls._INPUTS}
        self.device = self.t["start_params"].device
        self.n_tracks = int(n_tracks)
        self.n_hits = int(n_hits)
        self._outputs(track_states, max(int(max_states), 1))
        return self

    def _outputs(self, track_states, max_states):
        torch = _torch()
        n = max(self.n_tracks, 1)
        f64 = dict(dtype=torch.float64, device=self.device)
        self.out = {
            "params": torch.zeros((6, n), **f64),
            "cov": torch.zeros((21, n), **f64),
            "chi2": torch.zeros(n, **f64),
            "ndf": torch.zeros(n, dtype=torch.int32, device=self.device),
            "n_measured": torch.zeros(n, dtype=torch.int32, device=self.device),
            "status": torch.zeros(n, dtype=torch.int32, device=self.device),
        }
        self.max_states = int(max_states)
        if track_states:
            self.out["track_states"] = torch.zeros(
                (self.max_states, abi.KF_TS_STRIDE, n), **f64)

    def c_tracks(self):
        t = abi.kf_tracks()
        t.n_tracks = self.n_tracks
        t.n_hits = self.n_hits
        for name in self._INPUTS:
            setattr(t, name, self.t[name].data_ptr())
        return t

    def c_results(self):
        r = abi.kf_results()
        for name, tensor in self.out.items():
            setattr(r, name, tensor.data_ptr())
        return r

    def clear_results(self):
        """Zero-fills every output buffer (fill kernels on the current stream)."""
        for v in self.out.values():
            v.zero_()

    def results_numpy(self):
        """Copies the outputs to host numpy arrays (synchronising)."""
        res = {k: v.cpu().numpy() for k, v in self.out.items()}
        res["status"] = res["status"].view(np.uint32)
        return res


class KalmanFitter:
    """The fitter: geometry + field on one HIP device, fits through kf_fit().

    ``KalmanFitter(planes, field)`` binds the plane table (a ctypes array of
    :class:`abi.kf_plane`) and a :class:`generate.FieldSpec`.
    """

    def __init__(self, planes, field, device=0):
        _torch()
        self._lib = _lib()
        if self._lib.kf_abi_version() != abi.KF_ABI_VERSION:
            raise FitError("libkfmi.so ABI version mismatch")
        ctx = C.c_void_p()
        rc = self._lib.kf_ctx_create(int(device), C.byref(ctx))
        if rc != abi.KF_OK:
            raise FitError(self._lib.kf_last_error(None).decode())
        self._ctx = ctx
        self.device = device
        self.n_planes = len(planes)
        self._check(self._lib.kf_set_geometry(self._ctx, planes, len(planes)))
        self.set_field(field)

    def _check(self, rc):
        if rc != abi.KF_OK:
            raise FitError(f"libkfmi error {rc}: {self._lib.kf_last_error(self._ctx).decode()}")

    def set_field(self, field):
        if field.kind == "const":
            self._check(self._lib.kf_set_field_constant(self._ctx, *map(float, field.b)))
        else:
            vals = np.ascontiguousarray(field.values, dtype=np.float32)
            lo = (C.c_double * 3)(*field.lo)
            hi = (C.c_double * 3)(*field.hi)
            nx, ny, nz = vals.shape[:3]
            self._check(self._lib.kf_set_field_grid(self._ctx, vals.ctypes.data, nx, ny, nz, lo, hi))
        self._check(self._lib.kf_set_field_scale(self._ctx, float(field.scale)))

    def set_field_scale(self, s):
        self._check(self._lib.kf_set_field_scale(self._ctx, float(s)))

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.kf_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ batch API --
    def fit_batch(self, dev_batch, options=None, stream=None):
        """Fits every track of a :class:`DeviceBatch` (asynchronous on ``stream``)."""
        torch = _torch()
        opts = (options or KalmanFitterOptions()).to_c()
        if opts.max_states == 0:
            opts.max_states = dev_batch.max_states
        tracks = dev_batch.c_tracks()
        results = dev_batch.c_results()
        s = stream if stream is not None else torch.cuda.current_stream(dev_batch.device)
        self._check(self._lib.kf_fit(self._ctx, C.byref(tracks), C.byref(opts), C.byref(results),
                                     C.c_void_p(s.cuda_stream)))

    def fit_batch_noalloc(self, dev_batch, options=None, stream=None):
        """fit_batch through kf_fit_noalloc: no device allocation, the library's
        fixed per-device workspace instead (one thread per track; calls that
        overlap on one GPU race)."""
        torch = _torch()
        opts = (options or KalmanFitterOptions()).to_c()
        if opts.max_states == 0:
            opts.max_states = dev_batch.max_states
        tracks = dev_batch.c_tracks()
        results = dev_batch.c_results()
        s = stream if stream is not None else torch.cuda.current_stream(dev_batch.device)
        self._check(self._lib.kf_fit_noalloc(self._ctx, C.byref(tracks), C.byref(opts),
                                             C.byref(results), C.c_void_p(s.cuda_stream)))

    def set_schedule(self, **fields):
        """Sets kf_schedule fields (unnamed ones keep the library defaults)."""
        s = abi.kf_schedule()
        self._lib.kf_schedule_default(C.byref(s))
        for k, v in fields.items():
            setattr(s, k, int(v))
        self._check(self._lib.kf_set_schedule(self._ctx, C.byref(s)))

    def set_memory_budget(self, nbytes):
        """Device-memory budget of the per-batch working buffers (0: the
        library default, 16 GiB); larger batches run in slices."""
        self._check(self._lib.kf_set_memory_budget(self._ctx, int(nbytes)))

    def schedule(self):
        s = abi.kf_schedule()
        self._check(self._lib.kf_get_schedule(self._ctx, C.byref(s)))
        return s.as_dict()

    def autotune(self, dev_batch, options=None, stream=None, repeats=2):
        """Runs the library's autotuner on this batch; returns (schedule dict, ms)."""
        torch = _torch()
        opts = (options or KalmanFitterOptions()).to_c()
        if opts.max_states == 0:
            opts.max_states = dev_batch.max_states
        tracks = dev_batch.c_tracks()
        results = dev_batch.c_results()
        s = stream if stream is not None else torch.cuda.current_stream(dev_batch.device)
        best = abi.kf_schedule()
        ms = C.c_double()
        self._check(self._lib.kf_autotune(self._ctx, C.byref(tracks), C.byref(opts),
                                          C.byref(results), C.c_void_p(s.cuda_stream),
                                          int(repeats), C.byref(best), C.byref(ms)))
        return best.as_dict(), ms.value

    def set_geometry(self, planes):
        """Replaces the plane table (same number of planes or not)."""
        self.n_planes = len(planes)
        self._check(self._lib.kf_set_geometry(self._ctx, planes, len(planes)))

    def chi2_gradient(self, dev_batch, options=None, stream=None, synchronize=True):
        """Gradient of the summed forward-filter chi2 of the batch with respect
        to the alignment of every plane (local shifts du, dv, dw and rotations
        a, b, c about the local axes, at zero) and to the field scale, by the
        library's adjoint kernels (FP64).

        Returns ``(grad_planes, grad_field_scale, chi2_sum, track_flags,
        grad_raw)``: ``grad_planes`` is (n_planes, 6); ``track_flags`` holds
        the KF_ADJ_* bits of tracks left out (0 = included); ``grad_raw`` is
        (n_planes, 12), the derivatives with respect to each plane's centre (3)
        and rotation matrix (9, row-major).  Device tensors when
        ``synchronize`` is False, else numpy / floats.
        """
        torch = _torch()
        opts = (options or KalmanFitterOptions()).to_c()
        if opts.max_states == 0:
            opts.max_states = dev_batch.max_states
        dev = dev_batch.device
        gp = torch.zeros((self.n_planes, 6), dtype=torch.float64, device=dev)
        gs = torch.zeros(1, dtype=torch.float64, device=dev)
        cs = torch.zeros(1, dtype=torch.float64, device=dev)
        fl = torch.zeros(max(dev_batch.n_tracks, 1), dtype=torch.int32, device=dev)
        raw = torch.zeros((self.n_planes, 12), dtype=torch.float64, device=dev)
        tracks = dev_batch.c_tracks()
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        self._check(self._lib.kf_chi2_gradient(self._ctx, C.byref(tracks), C.byref(opts),
                                               gp.data_ptr(), gs.data_ptr(), cs.data_ptr(),
                                               raw.data_ptr(), fl.data_ptr(),
                                               C.c_void_p(s.cuda_stream)))
        if not synchronize:
            return gp, gs, cs, fl, raw
        s.synchronize()
        return (gp.cpu().numpy(), float(gs.item()), float(cs.item()),
                fl[:dev_batch.n_tracks].cpu().numpy().view(np.uint32), raw.cpu().numpy())

    def last_kernel_ms(self):
        """Duration of the last fit's kernels (both launches), ms."""
        return float(self._lib.kf_last_kernel_ms(self._ctx))

    def last_phase_ms(self):
        """(forward filter, smoother + back-propagation) durations of the last fit, ms."""
        f, b = C.c_double(), C.c_double()
        self._check(self._lib.kf_last_phase_ms(self._ctx, C.byref(f), C.byref(b)))
        return f.value, b.value

    def graph_captures(self):
        """Launch sequences captured as hipGraphs so far (a replay of unchanged
        arguments captures nothing)."""
        return int(self._lib.kf_graph_captures(self._ctx))

    # ------------------------------------------------------- per-track API --
    def fit(self, sourcelinks, start_parameters, kf_options, perigee,
            surface_sequence_first=0, surface_sequence_size=None):
        """Per-track fit with the reference's argument roles.

        ``sourcelinks``: list of :class:`PixelSourceLink`; ``start_parameters``:
        :class:`BoundParameters` on the z-parallel perigee line through
        ``perigee`` (the reference's LineSurface target, which is also
        ``kfOptions.referenceSurface``); the surface sequence is the run of
        ``surface_sequence_size`` planes of the bound plane table starting at
        ``surface_sequence_first``.  Returns ``(ok, KalmanFitterResult)`` where
        ``ok`` is the reference's bool.
        """
        from .generate import Batch
        if surface_sequence_size is None:
            surface_sequence_size = self.n_planes - surface_sequence_first
        m = len(sourcelinks)
        batch = Batch(
            planes=None, true_planes=None, field=None,
            start_params=np.asarray(start_parameters.parameters, dtype=np.float64).reshape(6, 1),
            start_cov=pack_sym(np.asarray(start_parameters.covariance, dtype=np.float64)).reshape(21, 1),
            perigee=np.asarray(perigee, dtype=np.float64).reshape(3, 1),
            seq_first=np.array([surface_sequence_first], dtype=np.int32),
            seq_count=np.array([surface_sequence_size], dtype=np.int32),
            hit_offset=np.array([0, m], dtype=np.int64),
            hit_geo_id=np.array([sl.geometry_id for sl in sourcelinks], dtype=np.int32),
            hit_meas=np.array([[sl.values[0] for sl in sourcelinks],
                               [sl.values[1] for sl in sourcelinks]], dtype=np.float64).reshape(2, m),
            hit_cov=np.array([[sl.cov[0][0] for sl in sourcelinks],
                              [sl.cov[0][1] for sl in sourcelinks],
                              [sl.cov[1][1] for sl in sourcelinks]], dtype=np.float64).reshape(3, m),
            truth=np.zeros((6, 1)))
        dev = DeviceBatch(batch, device=f"cuda:{self.device}", track_states=True,
                          max_states=max(surface_sequence_size, 1))
        self.fit_batch(dev, kf_options)
        r = dev.results_numpy()
        st = int(r["status"][0])
        n_meas = int(r["n_measured"][0])
        res = KalmanFitterResult(
            fitted_parameters=BoundParameters(r["params"][:, 0], unpack_sym(r["cov"][:, 0])),
            measurement_states=n_meas,
            smoothed=bool(st & abi.KF_STATUS_SMOOTHED),
            finished=bool(st & abi.KF_STATUS_FINISHED),
            result=not bool(st & abi.KF_STATUS_RESULT_FAIL),
            chi2=float(r["chi2"][0]), ndf=int(r["ndf"][0]), status=st,
            fitted_states=r["track_states"][:n_meas, :, 0])
        return bool(st & abi.KF_STATUS_OK), res
