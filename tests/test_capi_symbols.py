"""The C-ABI libraries load without a GPU and export every function their
headers declare (no compute calls here)."""
import ctypes
import os
import re

import pytest

from rich_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HEADERS = {"richgpu.h": "librichgpu", "richmesh.h": "librichmesh", "richio.h": "librichio"}


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if n.startswith(("rg_", "rm_", "rio_"))))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_exports(header):
    names = declared_functions(header)
    assert len(names) >= 5
    if header == "richio.h" and not os.path.exists(os.path.join(_ffi.LIB_DIR, "librichio.so")):
        pytest.skip("built without HDF5")
    lib = _ffi.load(HEADERS[header])
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_richgpu_entry_points_cover_the_plugins():
    names = declared_functions("richgpu.h")
    for n in ("rg_time_step", "rg_reconstruct", "rg_flux", "rg_update_extensives",
              "rg_update_cells", "rg_init_extensives", "rg_advance", "rg_set_mesh"):
        assert n in names


def test_kernel_names_without_gpu():
    lib = _ffi.load("librichgpu")
    lib.rg_kernel_count.restype = ctypes.c_int32
    lib.rg_kernel_name.restype = ctypes.c_char_p
    lib.rg_kernel_name.argtypes = [ctypes.c_int32]
    names = [lib.rg_kernel_name(k).decode() for k in range(lib.rg_kernel_count())]
    assert "k_slopes" in names and "k_flux2" in names and "k_update" in names
