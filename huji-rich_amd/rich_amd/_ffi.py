"""Locating and loading the native libraries built in-tree (lib/)."""
import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(_PKG_ROOT, "lib")

_cache = {}


def load(name):
    """Load lib/<name>.so; raise if it has not been built (no fallback)."""
    if name in _cache:
        return _cache[name]
    path = os.path.join(LIB_DIR, name + ".so")
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run __graft_entry__.build() (make -C huji-rich_amd)")
    lib = ctypes.CDLL(path)
    _cache[name] = lib
    return lib


def ptr(a):
    """ctypes pointer to the data of a contiguous numpy array (or None)."""
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)
