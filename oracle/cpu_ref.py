"""Build and call oracle/cpu_ref.c (test infrastructure; CPU baseline timing).

Only tests/, __graft_entry__ and bench.py's cpu_baseline leg use this.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SOURCE = HERE / "cpu_ref.c"
LIBRARY = HERE / "_build" / "libcpu_ref.so"

_lib = None


def build(force: bool = True) -> Path:
    LIBRARY.parent.mkdir(exist_ok=True)
    if LIBRARY.exists() and not force:
        return LIBRARY
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared", "-fPIC", str(SOURCE),
           "-o", str(tmp), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(tmp, LIBRARY)
    return LIBRARY


def load():
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            build()
        lib = ctypes.CDLL(str(LIBRARY))
        f32p = ctypes.POINTER(ctypes.c_float)
        lib.cpu_set_threads.argtypes = [ctypes.c_int]
        lib.cpu_calibrate_u16be.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_int,
                                            ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
        lib.cpu_warp.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int,
                                 ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]
        lib.cpu_warp_map64.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                       ctypes.c_double, ctypes.c_void_p]
        lib.cpu_integrate_winsorized.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                                 ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p]
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def set_threads(threads: int) -> None:
    """OpenMP worker threads of the restated loops (<= 0: runtime default)."""
    load().cpu_set_threads(int(threads))


def calibrate_u16be(raw, npix, bias, dark, flat, dark_scale, dark_includes_bias=True, flat_floor=1e-6, pedestal=0.0):
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    n = raw.size // (2 * npix)
    out = np.empty((n, npix), dtype=np.float32)
    load().cpu_calibrate_u16be(_p(raw), n, npix, _p(bias), _p(dark), _p(flat), float(dark_scale),
                               int(dark_includes_bias), float(flat_floor), float(pedestal), _p(out))
    return out


def warp(src, inverse, interpolation="lanczos3", fill=np.nan):
    src = np.ascontiguousarray(src, dtype=np.float32)
    inv = np.ascontiguousarray(inverse, dtype=np.float64)
    out = np.empty_like(src)
    cov = np.empty_like(src)
    load().cpu_warp(_p(src), src.shape[0], src.shape[1], _p(inv), 3 if interpolation == "lanczos3" else 1,
                    float(fill), _p(out), _p(cov))
    return out, cov


def warp_map64(src, source_map, fill=0.0):
    """FP64 CPU-path Lanczos3 through an affine source map (2x3)."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    b = np.ascontiguousarray(np.asarray(source_map, dtype=np.float64).reshape(6))
    out = np.empty(src.shape, dtype=np.float64)
    load().cpu_warp_map64(_p(src), src.shape[0], src.shape[1], _p(b), float(fill), _p(out))
    return out


def integrate_winsorized(stack, weights, low_sigma=3.0, high_sigma=3.0, min_samples=3, max_frac=0.5):
    stack = np.ascontiguousarray(stack, dtype=np.float32)
    f = stack.shape[0]
    npix = int(np.prod(stack.shape[1:]))
    w = np.ascontiguousarray(weights, dtype=np.float32)
    outs = [np.empty(stack.shape[1:], dtype=np.float32) for _ in range(5)]
    load().cpu_integrate_winsorized(_p(stack), f, npix, _p(w), float(low_sigma), float(high_sigma),
                                    int(min_samples), float(max_frac), *[_p(o) for o in outs])
    return tuple(outs)
