"""Winsorized sigma statistics oracle (test infrastructure; see oracle/__init__.py).

Restates src/glass/engine/rejection.py:31-366: ``center_and_scale`` with its two
arithmetic variants.  Which variant runs is decided once for the whole array
passed in (rejection.py:39-44, ``_all_valid_finite``):

* all samples valid and finite  -> ``_winsorized_sigma_all_valid``
  (rejection.py:219-239): percentile lerp ``a + (b-a)*t`` only when the
  position is fractional (rejection.py:352-366); np.std with the unrounded
  FP64 mean; fallback when scale <= 0.
* otherwise -> ``_winsorized_sigma_valid_deterministic`` (rejection.py:242-296):
  per-pixel count of valid samples, lerp ``a*(1-t) + b*t`` (rejection.py:
  299-320), FP64 sums with the f32-rounded mean (rejection.py:266-291,
  323-349), fallback when scale is not (>0 and finite).
"""

from __future__ import annotations

import numpy as np

IQR_TO_SIGMA = np.float32(1.349)


def _all_valid_finite(stack, valid):
    return bool(np.all(valid)) and bool(np.all(np.isfinite(stack)))


def _percentile_all_valid(sorted_stack, percentile):
    """rejection.py:352-366 (frame count fixed by the array's first axis)."""
    n = int(sorted_stack.shape[0])
    if n <= 0:
        raise ValueError("cannot compute percentile of an empty stack")
    if n == 1:
        return np.asarray(sorted_stack[0], dtype=np.float32)
    position = (np.float32(percentile) / np.float32(100.0)) * np.float32(n - 1)
    lower = int(np.floor(position))
    upper = int(np.ceil(position))
    frac = np.float32(position - lower)
    a = np.asarray(sorted_stack[lower], dtype=np.float32)
    if upper == lower:
        return a
    b = np.asarray(sorted_stack[upper], dtype=np.float32)
    return (a + (b - a) * frac).astype(np.float32)


def _percentile_valid(sorted_stack, count, fraction):
    """rejection.py:299-320 (per-pixel valid count; invalid sorted to +inf)."""
    safe = np.maximum(count, 1).astype(np.int64)
    position = (safe - 1).astype(np.float32) * np.float32(fraction)
    lower = np.floor(position).astype(np.int64)
    upper = np.minimum(lower + 1, safe - 1).astype(np.int64)
    t = (position - lower.astype(np.float32)).astype(np.float32)
    a = np.take_along_axis(sorted_stack, lower[None], axis=0)[0].astype(np.float32)
    b = np.take_along_axis(sorted_stack, upper[None], axis=0)[0].astype(np.float32)
    value = (a * (np.float32(1.0) - t) + b * t).astype(np.float32)
    return np.where(count > 0, value, np.float32(0.0)).astype(np.float32)


def _frame_axis_std(data, valid, count):
    """rejection.py:323-349: FP64 sums, mean rounded to f32 before the deltas."""
    n64 = count.astype(np.float64)
    d64 = data.astype(np.float64)
    s = np.sum(np.where(valid, d64, 0.0), axis=0, dtype=np.float64)
    mean = np.divide(s, n64, out=np.zeros_like(s), where=count > 0).astype(np.float32)
    delta = d64 - mean[None].astype(np.float64)
    ss = np.sum(np.where(valid, delta * delta, 0.0), axis=0, dtype=np.float64)
    var = np.divide(ss, n64, out=np.zeros_like(ss), where=count > 0)
    return np.sqrt(var).astype(np.float32)


def _winsorized_all_valid(stack, low_sigma, high_sigma):
    """rejection.py:219-239."""
    data = np.asarray(stack, dtype=np.float32)
    ordered = np.sort(data, axis=0)
    center0 = _percentile_all_valid(ordered, 50.0)
    q25 = _percentile_all_valid(ordered, 25.0)
    q75 = _percentile_all_valid(ordered, 75.0)
    scale0 = ((q75 - q25) / IQR_TO_SIGMA).astype(np.float32)
    fallback = np.std(data, axis=0, dtype=np.float64).astype(np.float32)
    scale0 = np.where(scale0 > 0, scale0, fallback)
    lo = center0 - np.float32(low_sigma) * scale0
    hi = center0 + np.float32(high_sigma) * scale0
    wins = np.clip(data, lo[None], hi[None])
    center = np.mean(wins, axis=0, dtype=np.float64).astype(np.float32)
    scale = np.std(wins, axis=0, dtype=np.float64).astype(np.float32)
    scale = np.where(scale > 0, scale, scale0)
    return center.astype(np.float32), scale.astype(np.float32)


def _winsorized_valid(stack, valid, low_sigma, high_sigma):
    """rejection.py:242-296."""
    data = np.asarray(stack, dtype=np.float32)
    ok = np.asarray(valid, dtype=bool) & np.isfinite(data)
    count = np.sum(ok, axis=0).astype(np.int64)
    ordered = np.sort(np.where(ok, data, np.inf), axis=0)
    center0 = _percentile_valid(ordered, count, 0.5)
    q25 = _percentile_valid(ordered, count, 0.25)
    q75 = _percentile_valid(ordered, count, 0.75)
    fallback = _frame_axis_std(data, ok, count)
    scale0 = ((q75 - q25) / IQR_TO_SIGMA).astype(np.float32)
    scale0 = np.where((scale0 > 0) & np.isfinite(scale0), scale0, fallback).astype(np.float32)
    lo = (center0 - np.float32(low_sigma) * scale0).astype(np.float32)
    hi = (center0 + np.float32(high_sigma) * scale0).astype(np.float32)
    wins = np.clip(data, lo[None], hi[None])
    n64 = count.astype(np.float64)
    wsum = np.sum(np.where(ok, wins, np.float32(0.0)).astype(np.float64), axis=0, dtype=np.float64)
    wmean = np.divide(wsum, n64, out=np.zeros_like(wsum), where=count > 0).astype(np.float32)
    delta = wins.astype(np.float64) - wmean[None].astype(np.float64)
    vsum = np.sum(np.where(ok, delta * delta, 0.0), axis=0, dtype=np.float64)
    var = np.divide(vsum, n64, out=np.zeros_like(vsum), where=count > 0)
    scale = np.sqrt(var).astype(np.float32)
    scale = np.where((scale > 0) & np.isfinite(scale), scale, scale0).astype(np.float32)
    has = count > 0
    return (
        np.where(has, wmean, np.float32(0.0)).astype(np.float32),
        np.where(has, scale, np.float32(0.0)).astype(np.float32),
    )


def center_and_scale(stack, valid, *, method="winsorized_sigma", low_sigma=3.0, high_sigma=3.0, all_valid=None):
    """rejection.py:31-65 (winsorized_sigma and sigma_clip methods).

    ``all_valid`` None decides the arithmetic from this array, as the
    reference does (rejection.py:39-44); True / False impose the decision
    taken over a whole frame when this array is a block of it."""
    if all_valid is None:
        all_valid = _all_valid_finite(stack, valid)
    if method == "winsorized_sigma" and all_valid:
        return _winsorized_all_valid(stack, low_sigma, high_sigma)
    masked = np.where(valid, stack, np.nan)
    if method in {"sigma", "sigma_clip"}:
        center = np.nanmedian(masked, axis=0)
        scale = np.nanstd(masked, axis=0)
    elif method in {"mad", "median_sigma"}:
        center = np.nanmedian(masked, axis=0)
        scale = np.float32(1.4826) * np.nanmedian(np.abs(masked - center[None]), axis=0)
    elif method == "winsorized_sigma":
        center, scale = _winsorized_valid(stack, valid, low_sigma, high_sigma)
    else:
        raise ValueError(f"unsupported rejection statistics method: {method}")
    return (
        np.nan_to_num(center, nan=0.0).astype(np.float32),
        np.nan_to_num(scale, nan=0.0).astype(np.float32),
    )


def uses_all_valid_variant(stack, valid):
    """True when ``center_and_scale`` takes the all-valid arithmetic."""
    return _all_valid_finite(stack, valid)
