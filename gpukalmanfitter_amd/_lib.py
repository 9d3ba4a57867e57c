"""Locates and loads the in-tree shared libraries of the package."""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def lib_path(name):
    return os.path.join(_HERE, name)


def load_library(name):
    """Loads ``name`` from the package directory; raises if it was not built."""
    if name in _LIBS:
        return _LIBS[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{name} is not built ({path} missing); run __graft_entry__.build() "
            "or `make -C gpukalmanfitter_amd/csrc`")
    lib = C.CDLL(path)
    _LIBS[name] = lib
    return lib
