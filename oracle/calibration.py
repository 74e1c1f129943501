"""Light calibration oracle (test infrastructure; see oracle/__init__.py).

Follows src/glass/cpu/calibration.py:51-76 (``calibrate_light``) and the raw
sample decode of cpp/cuda/calibration_kernels.cu:110-153 (FITS BITPIX=16,
BZERO=32768, big-endian).
"""

from __future__ import annotations

import numpy as np


def dark_scale_factor(light_exposure_s, dark_exposure_s, dark_scaling_enabled=True):
    """Scale applied to the master dark.

    src/glass/cpu/calibration.py:63-65: ``float(light)/float(dark)`` in FP64
    when scaling is enabled and the dark exposure is neither None nor 0,
    otherwise 1.0.  numpy then multiplies the f32 dark by it in f32
    (the Python float is a weak scalar), so the effective factor is
    ``np.float32(scale)``.
    """
    scale = 1.0
    if dark_scaling_enabled and dark_exposure_s not in (None, 0):
        scale = float(light_exposure_s) / float(dark_exposure_s)
    return np.float32(scale)


def calibrate_light(
    light,
    master_bias,
    master_dark,
    master_flat,
    light_exposure_s,
    dark_exposure_s,
    *,
    master_dark_includes_bias=True,
    dark_scaling_enabled=True,
    flat_floor=1.0e-6,
    pedestal=0.0,
):
    """src/glass/cpu/calibration.py:51-76, operation for operation in f32.

    x = light; x -= dark*s (then bias if the dark excludes it) or x -= bias;
    x /= max(flat, floor); x += pedestal (only if pedestal != 0).
    """
    calibrated = np.asarray(light).astype(np.float32, copy=True)
    if master_dark is not None:
        scale = dark_scale_factor(light_exposure_s, dark_exposure_s, dark_scaling_enabled)
        calibrated = calibrated - np.asarray(master_dark, dtype=np.float32) * scale
        if not master_dark_includes_bias and master_bias is not None:
            calibrated = calibrated - np.asarray(master_bias, dtype=np.float32)
    elif master_bias is not None:
        calibrated = calibrated - np.asarray(master_bias, dtype=np.float32)
    if master_flat is not None:
        flat = np.maximum(np.asarray(master_flat, dtype=np.float32), np.float32(flat_floor))
        calibrated = calibrated / flat
    if pedestal:
        calibrated = calibrated + np.float32(pedestal)
    return calibrated.astype(np.float32)


def decode_fits_u16be_bzero(raw, shape=None):
    """Big-endian int16 FITS samples with BZERO=32768 -> ADU as float32.

    cpp/cuda/calibration_kernels.cu:133-137: ``((b0<<8)|b1) ^ 0x8000``.
    """
    data = np.frombuffer(np.asarray(raw, dtype=np.uint8).tobytes(), dtype=">u2")
    adu = (data.astype(np.uint32) ^ 0x8000).astype(np.float32)
    return adu if shape is None else adu.reshape(shape)


def decode_u16le(raw, shape=None):
    """Native little-endian uint16 ADU -> float32 (exact)."""
    data = np.asarray(raw)
    if data.dtype != np.uint16:
        data = np.frombuffer(np.asarray(raw, dtype=np.uint8).tobytes(), dtype="<u2")
    out = data.astype(np.float32)
    return out if shape is None else out.reshape(shape)


def encode_fits_u16be_bzero(adu):
    """Inverse of :func:`decode_fits_u16be_bzero` (ADU uint16 -> BE bytes)."""
    adu = np.asarray(adu, dtype=np.uint16)
    stored = (adu.astype(np.uint32) ^ 0x8000).astype(">u2")
    return np.frombuffer(stored.tobytes(), dtype=np.uint8).copy()
