"""glass_amd -- MI355X-native resident stacking path of GLASS.

calibrate (bias/dark/flat) -> affine warp (bilinear, Lanczos3) -> winsorized
sigma-clip rejection with weighted-mean integration, plus output maps, as
hand-written HIP kernels for gfx950 behind a C ABI (include/glass_hip.h).

The public operator API mirrors the reference's resident backend
(``glass_cuda.ResidentCalibratedStack``, src/glass_cuda.py:2653).
"""

from __future__ import annotations

from ._native import LIBRARY_PATH, EXPORTED_SYMBOLS, load_library
from .resident import ResidentCalibratedStack, dark_scale, invert_affine

__all__ = [
    "LIBRARY_PATH",
    "EXPORTED_SYMBOLS",
    "load_library",
    "ResidentCalibratedStack",
    "dark_scale",
    "invert_affine",
]


def cuda_available() -> bool:
    """True when libglass_hip.so loads and sees at least one HIP device."""
    import ctypes

    try:
        lib = load_library()
    except (OSError, RuntimeError):
        return False
    count = ctypes.c_int(0)
    if lib.gh_device_count(ctypes.byref(count)) != 0:
        return False
    return count.value > 0
