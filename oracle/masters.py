"""Master calibration frame oracle (test infrastructure).

Restates the CPU-path master build of src/glass/engine/pipeline.py:
* bias:  ``_stack_mean_master`` (:647) -> ``_stack_mean_master_with_engine``
  (:529-610): CPUStackEngine per output tile, mean combine with FP64
  accumulator and the master rejection policy (:517-526; defaults
  src/glass/models.py:67-88: winsorized_sigma, 1 iteration, 3/3 sigma,
  min 3 samples, max fraction 0.5);
* dark:  same; if the dark does not include bias, the bias master is
  subtracted from the stacked dark (f32, :589-593, :1236-1246);
* flat:  ``_stack_normalized_flat_master`` (:828-955): every flat is reduced by
  the bias master, divided by its own exact median (:721-775, np.partition
  median; even counts average the two middle values in FP64), stacked, then
  floored with ``np.maximum(master, f32(flat_floor))``.
"""

from __future__ import annotations

import numpy as np

from .integration import stack_engine_combine


def exact_median(values):
    """pipeline.py:758-765 / :983-990: exact median of the finite values."""
    flat = np.asarray(values, dtype=np.float32).ravel()
    work = flat[np.isfinite(flat)].copy()
    n = int(work.size)
    if n == 0:
        raise ValueError("cannot compute median: no finite pixels")
    mid = n // 2
    work.partition(mid)
    upper = float(work[mid])
    if n % 2 == 1:
        return upper
    return (float(np.max(work[:mid])) + upper) / 2.0


def _master_stack(frames, policy):
    policy = policy or {}
    master, *_ = stack_engine_combine(
        np.asarray(frames, dtype=np.float32),
        None,
        low_sigma=float(policy.get("master_rejection_low_sigma", 3.0)),
        high_sigma=float(policy.get("master_rejection_high_sigma", 3.0)),
        min_samples=int(policy.get("master_rejection_min_samples", 3)),
        max_reject_fraction=float(policy.get("master_rejection_max_fraction", 0.5)),
        iterations=int(policy.get("master_rejection_iterations", 1)),
        method=str(policy.get("master_rejection", "winsorized_sigma")),
        tile_size=int(policy.get("tile_size", 512)),
    )
    return master


def build_master_bias(frames, policy=None):
    return _master_stack(frames, policy)


def build_master_dark(frames, master_bias=None, policy=None):
    policy = policy or {}
    master = _master_stack(frames, policy)
    if not policy.get("master_dark_includes_bias", True) and master_bias is not None:
        master = (master - np.asarray(master_bias, dtype=np.float32)).astype(np.float32)
    return master


def build_master_flat(frames, master_bias=None, policy=None):
    """Returns (master_flat, per-flat normalisation scalars)."""
    policy = policy or {}
    floor = float(policy.get("flat_floor", 1.0e-6))
    normalized = []
    scalars = []
    for flat in np.asarray(frames, dtype=np.float32):
        data = flat if master_bias is None else flat - np.asarray(master_bias, dtype=np.float32)
        scalar = exact_median(data)
        if abs(scalar) < floor:
            raise ValueError("flat normalization is below flat_floor")
        scalars.append(scalar)
        normalized.append((data / np.float32(scalar)).astype(np.float32))
    master = _master_stack(np.stack(normalized, axis=0), policy)
    return np.maximum(master, np.float32(floor)).astype(np.float32), scalars
