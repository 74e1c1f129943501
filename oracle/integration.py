"""Rejection + weighted-mean integration oracle (test infrastructure).

* ``weighted_integrate_stack`` restates src/glass/cpu/integration.py:36-118 --
  the compact CPU baseline the reference resident tests compare against
  (tests/test_cuda_resident_stack.py:2871-2932).
* ``stack_engine_combine`` restates the tile loop of CPUStackEngine.stack
  (src/glass/engine/stack_engine.py:39-120) with ``_apply_rejection``
  (:542-575), ``_rejection_pass`` (:578-593) and the mean branch of
  ``_combine_tile`` (:616-663).  The CPU path builds masters with it
  (src/glass/engine/pipeline.py:529-610, CombinePolicy(mean, float64)).
* ``mean_integrate`` restates src/glass/cpu/integration.py:8-18.
"""

from __future__ import annotations

import numpy as np

from .rejection import center_and_scale


def mean_integrate(frames, weights=None):
    """src/glass/cpu/integration.py:8-18."""
    if not len(frames):
        raise ValueError("cannot integrate an empty frame list")
    weights = weights if weights is not None else [1.0] * len(frames)
    acc = np.zeros_like(np.asarray(frames[0]), dtype=np.float64)
    wsum = np.zeros_like(acc)
    for frame, weight in zip(frames, weights):
        acc += np.asarray(frame, dtype=np.float64) * float(weight)
        wsum += float(weight)
    master = np.divide(acc, wsum, out=np.zeros_like(acc), where=wsum != 0).astype(np.float32)
    return master, wsum.astype(np.float32)


def _guard(valid, low, high, min_samples, max_reject_fraction):
    """Per-pixel rejection guard (cpu/integration.py:85-99)."""
    rejected = low | high
    if np.any(rejected):
        total = np.sum(valid, axis=0)
        rejected_count = np.sum(rejected, axis=0)
        remaining = total - rejected_count
        fraction = np.divide(
            rejected_count,
            np.maximum(total, 1),
            out=np.zeros_like(total, dtype=np.float32),
            where=total > 0,
        )
        allowed = (remaining >= int(min_samples)) & (fraction <= np.float32(max_reject_fraction))
        low = low & allowed[None]
        high = high & allowed[None]
    return low, high


def weighted_integrate_stack(
    stack,
    coverage=None,
    weights=None,
    rejection="none",
    low_sigma=3.0,
    high_sigma=3.0,
    min_samples=3,
    max_reject_fraction=0.5,
    all_valid=None,
):
    """src/glass/cpu/integration.py:36-118.

    Returns (master, weight_map, coverage_map, low_reject, high_reject), all f32.
    ``all_valid``: see ``center_and_scale`` (None: decided from this stack).
    """
    frames = np.asarray(stack, dtype=np.float32)
    if frames.ndim != 3:
        raise ValueError("stack must have shape (frame_count, height, width)")
    if min_samples < 1:
        raise ValueError("min_samples must be at least 1")
    if not 0.0 <= max_reject_fraction <= 1.0:
        raise ValueError("max_reject_fraction must be between 0 and 1")
    count, height, width = frames.shape
    if weights is None:
        w = np.ones(count, dtype=np.float32)
    else:
        w = np.asarray(weights, dtype=np.float32)
        if w.shape != (count,):
            raise ValueError("weights must have shape (frame_count,)")
    if coverage is None:
        valid = np.isfinite(frames)
    else:
        cov = np.asarray(coverage, dtype=np.float32)
        if cov.shape != frames.shape:
            raise ValueError("coverage must have the same shape as stack")
        valid = (cov > 0.5) & np.isfinite(frames)
    valid = valid & np.isfinite(w)[:, None, None] & (w[:, None, None] > 0)
    low_map = np.zeros((height, width), dtype=np.float32)
    high_map = np.zeros((height, width), dtype=np.float32)
    if rejection in {"sigma_clip", "winsorized_sigma"}:
        center, scale = center_and_scale(
            frames, valid, method=rejection, low_sigma=low_sigma, high_sigma=high_sigma, all_valid=all_valid
        )
        lo_t = center - np.float32(low_sigma) * scale
        hi_t = center + np.float32(high_sigma) * scale
        low = valid & (frames < lo_t[None])
        high = valid & (frames > hi_t[None])
        low, high = _guard(valid, low, high, min_samples, max_reject_fraction)
        low_map = np.sum(low, axis=0).astype(np.float32)
        high_map = np.sum(high, axis=0).astype(np.float32)
        valid = valid & ~(low | high)
    elif rejection != "none":
        raise ValueError(f"unsupported rejection mode: {rejection}")
    eff = np.where(valid, w[:, None, None], 0.0).astype(np.float64)
    wsum = np.sum(np.where(valid, frames, 0.0).astype(np.float64) * eff, axis=0)
    wmap = np.sum(eff, axis=0)
    cov_map = np.sum(valid, axis=0).astype(np.float32)
    master = np.divide(wsum, wmap, out=np.zeros_like(wsum), where=wmap > 0)
    return master.astype(np.float32), wmap.astype(np.float32), cov_map, low_map, high_map


def _tiles(width, height, tile_size):
    """Row-major tile order of src/glass/gpu/tile_scheduler.py (iter_tiles)."""
    for y0 in range(0, height, tile_size):
        for x0 in range(0, width, tile_size):
            yield y0, min(y0 + tile_size, height), x0, min(x0 + tile_size, width)


def stack_engine_combine(
    frames,
    weights=None,
    *,
    low_sigma=3.0,
    high_sigma=3.0,
    min_samples=3,
    max_reject_fraction=0.5,
    iterations=1,
    method="winsorized_sigma",
    tile_size=512,
    accumulator=np.float64,
):
    """CPUStackEngine.stack, mean combine with one rejection method.

    Per tile (stack_engine.py:83-87): valid = finite samples; rejection runs
    ``iterations`` passes (stack_engine.py:542-575) where a sample is rejected
    only if the scale is finite and > 0 (``usable``, stack_engine.py:586-589);
    combine (stack_engine.py:616-663) divides FP64 weighted sums.  The
    all-valid/valid arithmetic choice of center_and_scale is per tile.
    Weights <= 0 or non-finite become 0 (stack_engine.py:522-525) but those
    frames still take part in the rejection statistics.

    Returns (master, weight, coverage, low, high) f32 maps.
    """
    data = np.asarray(frames, dtype=np.float32)
    count, height, width = data.shape
    w = np.ones(count, dtype=np.float32) if weights is None else np.asarray(weights, dtype=np.float64)
    w = np.where(np.isfinite(w) & (w > 0), w, 0.0).astype(np.float32)
    master = np.zeros((height, width), dtype=np.float32)
    weight_map = np.zeros_like(master)
    coverage_map = np.zeros_like(master)
    low_map = np.zeros_like(master)
    high_map = np.zeros_like(master)
    for y0, y1, x0, x1 in _tiles(width, height, tile_size):
        tile = data[:, y0:y1, x0:x1]
        valid = np.isfinite(tile)
        low_acc = np.zeros(tile.shape[1:], dtype=np.float32)
        high_acc = np.zeros(tile.shape[1:], dtype=np.float32)
        if method != "none" and iterations > 0:
            work = valid.copy()
            for _ in range(max(1, int(iterations))):
                center, scale = center_and_scale(
                    tile, work, method=method, low_sigma=low_sigma, high_sigma=high_sigma
                )
                lo_t = center - np.float32(low_sigma) * scale
                hi_t = center + np.float32(high_sigma) * scale
                usable = work & np.isfinite(scale)[None] & (scale[None] > 0)
                rl = usable & (tile < lo_t[None])
                rh = usable & (tile > hi_t[None])
                rejected = rl | rh
                if not np.any(rejected):
                    break
                remaining = np.sum(work & ~rejected, axis=0)
                total = np.sum(work, axis=0)
                fraction = np.divide(
                    np.sum(rejected, axis=0),
                    np.maximum(total, 1),
                    out=np.zeros_like(total, dtype=np.float32),
                    where=total > 0,
                )
                allowed = (remaining >= min_samples) & (fraction <= max_reject_fraction)
                rl &= allowed[None]
                rh &= allowed[None]
                rejected = rl | rh
                low_acc += np.sum(rl, axis=0).astype(np.float32)
                high_acc += np.sum(rh, axis=0).astype(np.float32)
                work &= ~rejected
            valid = work
        cov = np.sum(valid, axis=0).astype(np.float32)
        eff = np.where(valid, w[:, None, None], 0.0).astype(accumulator)
        weighted = np.where(valid, tile, 0.0).astype(accumulator) * eff
        wsum = np.sum(eff, axis=0)
        m = np.divide(np.sum(weighted, axis=0), wsum, out=np.zeros_like(wsum, dtype=accumulator), where=wsum > 0)
        m = np.where(np.isfinite(m), m, 0.0)
        master[y0:y1, x0:x1] = m.astype(np.float32)
        weight_map[y0:y1, x0:x1] = wsum.astype(np.float32)
        coverage_map[y0:y1, x0:x1] = cov
        low_map[y0:y1, x0:x1] = low_acc
        high_map[y0:y1, x0:x1] = high_acc
    return master, weight_map, coverage_map, low_map, high_map
