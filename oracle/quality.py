"""Per-frame statistics and frame weights oracle (test infrastructure).

* ``scan_quality_stats``: src/glass/engine/quality.py:126-200 -- exact median
  by np.partition (even counts average the two middle values in FP64), MAD of
  |x - f32(median)|, FP64 sum / sum of squares, rms = sqrt(max(E[x^2]-mean^2,0)),
  noise_mad = 1.4826*MAD (rms when 0).
* ``quality_weights``: src/glass/engine/integration.py:80-106.
* ``frame_global_stats``: the resident FP64 mean/std of
  cpp/src/native_bindings.cpp:10569-10643.
* ``simple_snr_weight``: src/glass/engine/resident_cuda.py:1948-1955.
"""

from __future__ import annotations

import numpy as np


def _partition_median(work):
    n = int(work.size)
    mid = n // 2
    work.partition(mid)
    value = float(work[mid])
    if n % 2 == 0:
        value = (float(np.max(work[:mid])) + value) / 2.0
    return value


def scan_quality_stats(frame):
    values = np.asarray(frame, dtype=np.float32).ravel()
    finite = values[np.isfinite(values)]
    n = int(finite.size)
    if n == 0:
        raise ValueError("cannot measure quality: no finite pixels")
    f64 = finite.astype(np.float64)
    total = float(np.sum(f64))
    total_sq = float(np.sum(f64 * f64))
    work = finite.copy()
    median = _partition_median(work)
    np.subtract(work, np.float32(median), out=work)
    np.abs(work, out=work)
    mad = _partition_median(work)
    mean = total / n
    variance = max(total_sq / n - mean * mean, 0.0)
    rms = float(np.sqrt(variance))
    noise_mad = float(1.4826 * mad)
    if noise_mad <= 0.0:
        noise_mad = rms
    return {
        "median": median,
        "mean": float(mean),
        "rms": rms,
        "mad": mad,
        "noise_mad": noise_mad,
        "pixel_count": n,
    }


def frame_global_stats(frame):
    values = np.asarray(frame, dtype=np.float32).ravel()
    finite = values[np.isfinite(values)].astype(np.float64)
    n = int(finite.size)
    mean = float(np.sum(finite)) / n if n else 0.0
    var = 0.0
    if n:
        var = max(float(np.sum(finite * finite)) / n - mean * mean, 0.0)
    return {
        "mean": mean,
        "std": float(np.sqrt(var)),
        "valid_pixels": n,
        "total_pixels": int(values.size),
        "nonfinite_pixels": int(values.size) - n,
    }


def simple_snr_weight(stats, eps=1.0e-6):
    if int(stats.get("valid_pixels") or 0) <= 0:
        return 0.0
    std = float(stats.get("std") or 0.0)
    mean = abs(float(stats.get("mean") or 0.0))
    if std <= eps:
        return 1.0 / eps
    return max(mean / std, eps)


def quality_weights(qualities, weighting):
    """engine/integration.py:80-106 on a list of per-frame quality dicts."""
    if weighting == "none":
        return [1.0 for _ in qualities]
    if weighting not in {"simple_snr", "combined", "variance_aware"}:
        raise ValueError(f"unsupported integration weighting mode: {weighting}")
    raw = []
    for q in qualities:
        if weighting == "simple_snr":
            value = float(q.get("snr") or q.get("weight") or 1.0)
        elif weighting == "combined":
            value = float(q.get("quality_score") or q.get("weight") or 1.0)
        else:
            noise = q.get("noise_sigma") or q.get("background_rms")
            if noise is None:
                value = float(q.get("weight") or 1.0)
            else:
                value = 1.0 / max(float(noise) ** 2, 1.0e-12)
        raw.append(max(value, 1.0e-6))
    positive = np.asarray([v for v in raw if np.isfinite(v) and v > 0.0], dtype=np.float32)
    scale = float(np.median(positive)) if positive.size else 1.0
    scale = max(scale, 1.0e-6)
    return [max(float(v) / scale, 1.0e-6) for v in raw]
