"""Module facade with the attribute surface of the reference's native module.

The reference loads its pybind11 extension ``_glass_cuda_native`` in
``glass_cuda._native()`` (src/glass_cuda.py:31-43) and every public helper of
``glass_cuda`` then reads attributes off that module object: plain calls for
the device queries and the resident stack, ``hasattr`` probes for optional
capabilities (src/glass_cuda.py:228-260, :381-390).  Returning this module from
``_native()`` therefore switches the resident path to libglass_hip.so without
touching any caller.

Attributes provided (each replaces the reference binding of the same name):

* ``cuda_available()``            native_bindings.cpp:15570-15574
* ``list_devices()``              native_bindings.cpp:15576-15584
* ``get_device_info(device_id)``  native_bindings.cpp:15586-15594, keys of
  ``device_info_dict`` (:1932-1944)
* ``host_pinned_empty_f32(h, w)`` / ``host_pinned_empty_u8(n)``
  native_bindings.cpp:1946-1990
* ``ResidentCalibratedStack``     native_bindings.cpp:20624-21213
* ``triangle_asterism_descriptors_f32`` / ``_batch_f32``,
  ``estimate_similarity_from_triangle_descriptors_f32`` / ``_batch_f32``,
  ``estimate_similarity_from_catalogs_f32``  native_bindings.cpp:17044-18678
  (the resident default registration route, resident_cuda.py:9985-11399)

Deliberately absent, so the reference's ``hasattr`` probes take their Python
routes: ``RawFitsReadQueue`` (``raw_fits_read_queue_available()`` is then
False and the driver reads files itself, resident_cuda.py:1159-1897),
``read_simple_fits_raw_batch_into_u8`` and the host DQ helpers
``resident_dq_map_host_f32`` / ``resident_dq_map_count_maps_i16`` (the
reference's default DQ route is its vectorised Python, glass_cuda.py:240-252).
"""

from __future__ import annotations

import ctypes
import weakref

import numpy as np

from . import _native
from .resident import ResidentCalibratedStack
from .triangles import (
    estimate_similarity_from_catalogs_f32,
    estimate_similarity_from_triangle_descriptors_batch_f32,
    estimate_similarity_from_triangle_descriptors_f32,
    triangle_asterism_descriptors_batch_f32,
    triangle_asterism_descriptors_f32,
)

__all__ = [
    "cuda_available",
    "list_devices",
    "get_device_info",
    "host_pinned_empty_f32",
    "host_pinned_empty_u8",
    "ResidentCalibratedStack",
    "triangle_asterism_descriptors_f32",
    "triangle_asterism_descriptors_batch_f32",
    "estimate_similarity_from_triangle_descriptors_f32",
    "estimate_similarity_from_triangle_descriptors_batch_f32",
    "estimate_similarity_from_catalogs_f32",
]


def cuda_available() -> bool:
    """True when libglass_hip.so loads and sees at least one HIP device
    (never raises, like the reference's ``cudaGetDeviceCount`` probe)."""
    try:
        lib = _native.load_library()
    except (OSError, RuntimeError):
        return False
    count = ctypes.c_int(0)
    if lib.gh_device_count(ctypes.byref(count)) != _native.GH_OK:
        return False
    return count.value > 0


def _device_count() -> int:
    count = ctypes.c_int(0)
    _native.call("gh_device_count", ctypes.byref(count))
    return int(count.value)


def get_device_info(device_id: int) -> dict:
    """Properties of one device, with the reference's keys.  IndexError for an
    id outside [0, device count), as ``std::out_of_range`` maps in pybind11."""
    props = _native.DeviceProperties()
    _native.call("gh_device_info", int(device_id), ctypes.byref(props))
    return {
        "device_id": int(device_id),
        "name": props.name.decode("utf-8", errors="replace"),
        "compute_capability": f"{props.major}.{props.minor}",
        "memory_total_mib": int(props.memory_total_bytes) // (1024 * 1024),
        "multi_processor_count": int(props.multiprocessor_count),
        "native_backend": True,
        "available_to_glass": True,
        "arch": props.arch.decode("utf-8", errors="replace"),
        "wavefront_size": int(props.wavefront_size),
    }


def list_devices() -> list[dict]:
    return [get_device_info(i) for i in range(_device_count())]


def _pinned_array(nbytes: int, dtype, shape) -> np.ndarray:
    ptr = ctypes.c_void_p()
    _native.call("gh_host_alloc", ctypes.c_size_t(nbytes), ctypes.byref(ptr))
    raw = (ctypes.c_uint8 * nbytes).from_address(ptr.value)
    weakref.finalize(raw, _native.load_library().gh_host_free, ctypes.c_void_p(ptr.value))
    return np.frombuffer(raw, dtype=dtype).reshape(shape)


def host_pinned_empty_f32(height: int, width: int) -> np.ndarray:
    """Uninitialised (height, width) float32 array in pinned host memory."""
    height, width = int(height), int(width)
    if height <= 0 or width <= 0:
        raise ValueError("pinned host array dimensions must be non-empty")
    return _pinned_array(height * width * 4, np.float32, (height, width))


def host_pinned_empty_u8(byte_count: int) -> np.ndarray:
    """Uninitialised 1-D uint8 array in pinned host memory."""
    byte_count = int(byte_count)
    if byte_count <= 0:
        raise ValueError("pinned host byte array must be non-empty")
    return _pinned_array(byte_count, np.uint8, (byte_count,))
