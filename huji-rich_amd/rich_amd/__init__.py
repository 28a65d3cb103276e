"""MI355X-native per-timestep finite-volume update of RICH (huji-rich).

Python side of the package: ctypes bindings over the C ABI in
include/richgpu.h (lib/librichgpu.so, hand-written HIP for gfx950) and
include/richmesh.h (lib/librichmesh.so, the lattice mesh builder)."""
from . import mesh  # noqa: F401
