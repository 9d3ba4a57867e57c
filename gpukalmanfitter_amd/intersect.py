"""Batched ray x bounded-plane intersection (kf_intersect of libkfmi.so).

The reference tests the intersection of one ray with every surface of a
triangular mesh of the TrackML detector (Run/GPU/IntersectGPUTest.cu): each
triangle becomes a plane surface with ConvexPolygonBounds<3>, its centre the
centroid, its normal (v1 - v2) x (v1 - v3) and its local frame the
curvilinear one (U = z x T, or x x T near the z axis; V = T x U).  This
module builds such surfaces (``polygon_planes``), a synthetic barrel mesh of
the same size (``barrel_mesh``: the mesh file the reference reads is not
shipped), rays (``make_rays``), and runs the GPU kernel (``intersect``).
"""
import ctypes as C

import numpy as np

from . import abi
from .fitter import FitError, _lib, _torch

KF_POLY_MAX = 8
S_CURVILINEAR_PROJ_TOLERANCE = 0.999995  # Definitions.hpp:46

POLYGON_DTYPE = np.dtype([("center", "<f8", 3), ("rot", "<f8", 9),
                          ("vertices", "<f8", 2 * KF_POLY_MAX), ("n_vertices", "<i4"),
                          ("pad", "<i4")])
HIT_DTYPE = np.dtype([("ray", "<i4"), ("surface", "<i4"), ("path", "<f8"),
                      ("position", "<f8", 3)])


def _typed():
    lib = _lib()
    if not getattr(lib, "_kf_ix_typed", False):
        P = C.c_void_p
        lib.kf_intersect_scratch_bytes.argtypes = [C.c_int64, C.c_int64]
        lib.kf_intersect_scratch_bytes.restype = C.c_int64
        lib.kf_intersect.argtypes = [P, C.c_int64, P, P, C.c_int64, C.c_int32, P, C.c_int64, P,
                                     P, P]
        lib.kf_intersect.restype = C.c_int
        lib._kf_ix_typed = True
    return lib


def polygon_planes(polygons):
    """Plane surfaces from polygons given as global vertices [n, k, 3]
    (k in 3..8, convex, in order), built as the reference driver builds its
    triangles (IntersectGPUTest.cu:116-160): centre = vertex mean, normal =
    (v1 - v2) x (v1 - v3), curvilinear U, V; the vertices in (U, V)."""
    v = np.asarray(polygons, dtype=np.float64)
    n, k, _ = v.shape
    if not 3 <= k <= KF_POLY_MAX:
        raise ValueError(f"polygons need 3..{KF_POLY_MAX} vertices")
    center = v.mean(axis=1)
    normal = np.cross(v[:, 0] - v[:, 1], v[:, 0] - v[:, 2])
    T = normal / np.linalg.norm(normal, axis=1, keepdims=True)
    z = np.array([0.0, 0.0, 1.0])
    x = np.array([1.0, 0.0, 0.0])
    use_z = np.abs(T @ z) < S_CURVILINEAR_PROJ_TOLERANCE
    U = np.where(use_z[:, None], np.cross(z, T), np.cross(x, T))
    U /= np.linalg.norm(U, axis=1, keepdims=True)
    V = np.cross(T, U)
    out = np.zeros(n, dtype=POLYGON_DTYPE)
    out["center"] = center
    rot = np.stack([U, V, T], axis=2)  # rot[i][r][c]: column c is U, V, T
    out["rot"] = rot.reshape(n, 9)
    rel = v - center[:, None, :]
    out["vertices"][:, 0:2 * k:2] = np.einsum("nkj,nj->nk", rel, U)
    out["vertices"][:, 1:2 * k:2] = np.einsum("nkj,nj->nk", rel, V)
    out["n_vertices"] = k
    return out


def barrel_mesh(n_triangles=37456, radii=(32.0, 72.0, 116.0, 172.0, 260.0, 360.0, 500.0,
                                           660.0, 820.0, 1020.0), half_z=1000.0):
    """A triangular mesh of concentric cylinders (a pixel/strip barrel),
    about ``n_triangles`` triangles split over the layers in proportion to
    their radius; each cylinder is n_phi x n_z quads of two triangles."""
    radii = np.asarray(radii, dtype=np.float64)
    share = radii / radii.sum()
    tris = []
    for r, f in zip(radii, share):
        quads = max(int(round(n_triangles * f / 2)), 4)
        n_z = max(int(round(np.sqrt(quads * (2 * half_z) / (2 * np.pi * r)))), 2)
        n_phi = max(quads // n_z, 3)
        phi = np.linspace(0.0, 2 * np.pi, n_phi + 1)
        zz = np.linspace(-half_z, half_z, n_z + 1)
        P, Z = np.meshgrid(phi, zz, indexing="ij")
        pts = np.stack([r * np.cos(P), r * np.sin(P), Z], axis=-1)
        a = pts[:-1, :-1].reshape(-1, 3)
        b = pts[1:, :-1].reshape(-1, 3)
        c = pts[1:, 1:].reshape(-1, 3)
        d = pts[:-1, 1:].reshape(-1, 3)
        tris.append(np.stack([a, b, c], axis=1))
        tris.append(np.stack([a, c, d], axis=1))
    return np.concatenate(tris, axis=0)


def make_rays(n_rays, seed=7, eta_max=1.0, vertex_sigma=0.01):
    """Rays from a smeared origin, phi uniform, eta uniform in +-eta_max
    ([3][n] positions and unit directions)."""
    rng = np.random.default_rng(seed)
    phi = rng.uniform(-np.pi, np.pi, n_rays)
    eta = rng.uniform(-eta_max, eta_max, n_rays)
    theta = 2.0 * np.arctan(np.exp(-eta))
    d = np.stack([np.cos(phi) * np.sin(theta), np.sin(phi) * np.sin(theta), np.cos(theta)])
    p = rng.normal(0.0, vertex_sigma, (3, n_rays))
    return np.ascontiguousarray(p), np.ascontiguousarray(d)


class Intersector:
    """Device copies of a surface set for repeated kf_intersect calls."""

    def __init__(self, surfaces, device="cuda:0"):
        torch = _torch()
        self.device = torch.device(device)
        s = np.ascontiguousarray(surfaces, dtype=POLYGON_DTYPE)
        if len(s) and not ((s["n_vertices"] >= 3) & (s["n_vertices"] <= KF_POLY_MAX)).all():
            raise ValueError(f"surfaces need 3..{KF_POLY_MAX} vertices")
        self.n_surfaces = len(s)
        raw = np.frombuffer(s.tobytes() if len(s) else b"\0" * 8, dtype=np.uint8).copy()
        self.surfaces = torch.from_numpy(raw).to(self.device)

    def run(self, ray_pos, ray_dir, bcheck=True, capacity=None, stream=None):
        """Kept intersections, (ray, surface) order, as a HIT_DTYPE array;
        with the capacity reached, the first `capacity` and the total."""
        torch = _torch()
        pos = torch.as_tensor(np.ascontiguousarray(ray_pos, dtype=np.float64), device=self.device)
        dirs = torch.as_tensor(np.ascontiguousarray(ray_dir, dtype=np.float64),
                               device=self.device)
        n_rays = int(pos.shape[1]) if pos.dim() == 2 else 0
        out, total = self.run_device(pos, dirs, n_rays, bcheck, capacity, stream)
        got = min(total, out.numel() // HIT_DTYPE.itemsize)
        hits = np.frombuffer(out[:got * HIT_DTYPE.itemsize].cpu().numpy().tobytes(),
                             dtype=HIT_DTYPE)
        return hits, total

    def run_device(self, pos, dirs, n_rays, bcheck=True, capacity=None, stream=None,
                   out=None, scratch=None, n_hits=None, synchronize=True):
        """kf_intersect on device tensors ([3][n_rays] each); returns the hit
        buffer (bytes) and, when synchronize, the total number of hits."""
        torch = _torch()
        lib = _typed()
        if capacity is None:
            capacity = 64 * max(n_rays, 1)
        if out is None:
            out = torch.empty(max(capacity, 1) * HIT_DTYPE.itemsize, dtype=torch.uint8,
                              device=self.device)
        if scratch is None:
            scratch = torch.empty(max(int(lib.kf_intersect_scratch_bytes(n_rays,
                                                                         self.n_surfaces)), 8),
                                  dtype=torch.uint8, device=self.device)
        if n_hits is None:
            n_hits = torch.zeros(1, dtype=torch.int64, device=self.device)
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        rc = lib.kf_intersect(self.surfaces.data_ptr(), self.n_surfaces, pos.data_ptr(),
                              dirs.data_ptr(), n_rays, int(bool(bcheck)), out.data_ptr(),
                              capacity, n_hits.data_ptr(), scratch.data_ptr(),
                              C.c_void_p(s.cuda_stream))
        if rc != abi.KF_OK:
            raise FitError(f"kf_intersect failed ({rc}): {lib.kf_last_error(None).decode()}")
        if not synchronize:
            return out, None
        return out, int(n_hits.item())


def intersect(surfaces, ray_pos, ray_dir, bcheck=True, capacity=None, device="cuda:0"):
    """One-shot kf_intersect: see :class:`Intersector`."""
    return Intersector(surfaces, device).run(ray_pos, ray_dir, bcheck, capacity)
