"""Star detection and star-list registration on the MI355X path.

Mirrors the reference operator surface that produces the warp matrices:

* ``detect_stars(data, threshold_sigma=5.0, max_stars=500, *, background=None,
  noise=None, window_radius=4, min_separation_px=2.0, origin_x=0, origin_y=0)``
  -- src/glass/cpu/star_detect.py:121-168.  ``data`` is a host array or a GPU
  tensor; ``ResidentCalibratedStack.detect_stars(index, ...)`` runs on a
  resident frame without a copy.
* ``estimate_star_transform(reference, moving, transform_model="translation",
  min_inliers=6, tolerance_px=3.0, max_stars=80)`` --
  src/glass/cpu/registration.py:476-566 (translation, similarity, affine).

Detection runs the background selection and the 3x3 local-maximum candidate
pass on the GPU (libglass_hip.so) and measures the candidate windows in FP64
on the host inside the same library; matching is host FP64 in the library.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

_MODELS = {"translation": N.GH_MODEL_TRANSLATION, "similarity": N.GH_MODEL_SIMILARITY,
           "affine": N.GH_MODEL_AFFINE}


@dataclass(slots=True)
class Star:
    """src/glass/cpu/star_detect.py:8-18."""

    x: float
    y: float
    flux: float
    peak: float = 0.0
    fwhm_px: float | None = None
    eccentricity: float | None = None
    snr: float | None = None
    sharpness: float | None = None
    npix: int = 0


@dataclass(slots=True)
class StarRegistration:
    """src/glass/cpu/registration.py:11-20."""

    transform_model: str
    matrix: list
    matched_stars: int
    inliers: int
    rms_px: float
    status: str
    warnings: list = field(default_factory=list)
    pairs: list = field(default_factory=list)


def _stars_from_records(records, n) -> list[Star]:
    out = []
    for k in range(n):
        r = records[k]
        out.append(Star(x=r.x, y=r.y, flux=r.flux, peak=r.peak, fwhm_px=r.fwhm_px if r.has_fwhm else None,
                        eccentricity=r.eccentricity if r.has_eccentricity else None, snr=r.snr,
                        sharpness=r.sharpness, npix=int(r.npix)))
    return out


def detect_resident(handle, index, threshold_sigma=5.0, max_stars=500, background=None, noise=None,
                    window_radius=4, min_separation_px=2.0, origin_x=0, origin_y=0):
    """Stars of resident frame `index` of a stack handle (list[Star], and the
    (median, mad, noise) background triple)."""
    radius = int(window_radius)
    if radius > 8:
        raise ValueError("window_radius must be at most 8")
    cap = max(int(max_stars), 0) if max_stars is not None else 0
    records = (N.StarRecord * max(cap, 1))()
    count = ctypes.c_int64(0)
    bg = (ctypes.c_double * 3)()
    N.call("gh_detect_stars", handle, int(index), float(threshold_sigma), int(max_stars),
           math.nan if background is None else float(background), math.nan if noise is None else float(noise),
           radius, float(min_separation_px), int(origin_x), int(origin_y), records, cap, ctypes.byref(count), bg)
    n = min(int(count.value), cap)
    return _stars_from_records(records, n), (bg[0], bg[1], bg[2])


def detect_stars(data, threshold_sigma: float = 5.0, max_stars: int = 500, *, background=None, noise=None,
                 window_radius: int = 4, min_separation_px: float = 2.0, origin_x: int = 0,
                 origin_y: int = 0) -> list[Star]:
    """star_detect.py:121-168 on a host array or GPU tensor (H, W)."""
    from .resident import ResidentCalibratedStack

    shape = tuple(data.shape)
    if len(shape) != 2:
        raise ValueError(f"detect_stars expects a 2-D image, got shape {shape}")
    h, w = shape
    if h < 3 or w < 3:
        return []
    with ResidentCalibratedStack(1, h, w) as stack:
        if type(data).__module__.startswith("torch") and data.is_cuda:
            stack.upload_calibrated_frame(0, data.float().contiguous())
        else:
            stack.upload_calibrated_frame(0, np.ascontiguousarray(np.asarray(data, dtype=np.float32)))
        stars, _bg = detect_resident(stack._h, 0, threshold_sigma, max_stars, background, noise, window_radius,
                                     min_separation_px, origin_x, origin_y)
    return stars


def _xy(stars) -> np.ndarray:
    return np.ascontiguousarray(np.asarray([(float(s.x), float(s.y)) for s in stars], dtype=np.float64).reshape(-1))


def estimate_star_transform(reference, moving, transform_model: str = "translation", min_inliers: int = 6,
                            tolerance_px: float = 3.0, max_stars: int = 80) -> StarRegistration:
    """registration.py:476-566 for the translation, similarity and affine models."""
    if transform_model not in _MODELS:
        raise ValueError(f"unsupported star registration model: {transform_model}")
    ref, mov = _xy(reference), _xy(moving)
    keep = max(int(max_stars), 0)
    matrix = (ctypes.c_double * 9)()
    pairs = (ctypes.c_int64 * (2 * max(keep, 1)))()
    n_pairs, rms, ok = ctypes.c_int64(0), ctypes.c_double(0.0), ctypes.c_int32(0)
    dp = ctypes.POINTER(ctypes.c_double)
    N.call("gh_estimate_star_transform", ref.ctypes.data_as(dp), len(reference), mov.ctypes.data_as(dp), len(moving),
           _MODELS[transform_model], int(min_inliers), float(tolerance_px), int(max_stars), matrix, pairs,
           ctypes.byref(n_pairs), ctypes.byref(rms), ctypes.byref(ok))
    m = [[matrix[3 * r + c] for c in range(3)] for r in range(3)]
    n = int(n_pairs.value)
    minimum = 3 if transform_model == "affine" else 2 if transform_model == "similarity" else 1
    warnings = []
    if min(len(reference), keep) < minimum or min(len(moving), keep) < minimum:
        warnings.append(f"not enough stars for {transform_model} registration")
    elif not ok.value:
        warnings.append(f"registration has {n} inliers, below min_inliers={min_inliers}")
    return StarRegistration(transform_model=transform_model, matrix=m, matched_stars=n, inliers=n,
                            rms_px=float(rms.value), status="ok" if ok.value else "failed", warnings=warnings,
                            pairs=[(int(pairs[2 * k]), int(pairs[2 * k + 1])) for k in range(n)])
