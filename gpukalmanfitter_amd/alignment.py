"""Alignment parameterisation of the plane table used by the chi2 adjoint.

A plane with centre c and rotation R (columns U, V, normal) is moved by six
parameters (du, dv, dw, a, b, c_rot): centre' = c + R (du, dv, dw) and
R' = R Rx(a) Ry(b) Rz(c_rot), i.e. shifts along and rotations about the
plane's own local axes.  kf_chi2_gradient() returns the derivatives with
respect to these parameters at zero, and the raw derivatives with respect to
(centre, R) for other parameterisations.
"""
import ctypes as C

import numpy as np

from . import abi


def rotations(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    rx = np.array([[1, 0, 0], [0, ca, -sa], [0, sa, ca]])
    ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    rz = np.array([[cc, -sc, 0], [sc, cc, 0], [0, 0, 1]])
    return rx @ ry @ rz


def apply(planes, params):
    """New plane table: ``params`` is (n_planes, 6) = (du, dv, dw, a, b, c)."""
    params = np.asarray(params, dtype=np.float64).reshape(len(planes), 6)
    out = (abi.kf_plane * len(planes))()
    C.memmove(out, planes, C.sizeof(out))
    for i in range(len(planes)):
        R = np.array(out[i].rot).reshape(3, 3)
        c = np.array(out[i].center) + R @ params[i, :3]
        Rn = R @ rotations(*params[i, 3:])
        for k in range(3):
            out[i].center[k] = c[k]
        for k in range(9):
            out[i].rot[k] = Rn.reshape(-1)[k]
    return out


def centres_and_rotations(planes):
    """(n, 3) centres and (n, 3, 3) rotations of a plane table."""
    c = np.array([list(p.center) for p in planes], dtype=np.float64)
    r = np.array([list(p.rot) for p in planes], dtype=np.float64).reshape(-1, 3, 3)
    return c, r


def with_geometry(planes, centres, rots):
    """Copy of ``planes`` with the given centres and rotation matrices."""
    out = (abi.kf_plane * len(planes))()
    C.memmove(out, planes, C.sizeof(out))
    centres = np.asarray(centres, dtype=np.float64)
    rots = np.asarray(rots, dtype=np.float64).reshape(len(planes), 9)
    for i in range(len(planes)):
        for k in range(3):
            out[i].center[k] = centres[i, k]
        for k in range(9):
            out[i].rot[k] = rots[i, k]
    return out
