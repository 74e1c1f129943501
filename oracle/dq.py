"""Resident DQ map oracle (test infrastructure; see oracle/__init__.py).

Restates ``_resident_dq_map_python`` (src/glass/engine/resident_cuda.py:
3110-3297) with the DQFlag bits of src/glass/engine/contracts.py:77-88:
* NO_DATA   if master or weight is non-finite, weight <= 0, post-rejection
  coverage is non-finite or <= 0.5, or geometric coverage is non-finite or
  <= 0.5;
* WARP_EDGE if post-rejection coverage <= 0.5 (or non-finite), geometric
  coverage <= 0.5 (or non-finite), or 0.5 < geometric < active_frames - 0.5;
* LOW_REJECTED / HIGH_REJECTED if the finite count map is > 0.
"""

from __future__ import annotations

import numpy as np

DQ_NO_DATA = 1 << 0
DQ_WARP_EDGE = 1 << 6
DQ_LOW_REJECTED = 1 << 8
DQ_HIGH_REJECTED = 1 << 9


def resident_dq_map(
    master,
    weight_map,
    coverage_map=None,
    low_rejection_map=None,
    high_rejection_map=None,
    geometric_warp_coverage_map=None,
    active_frame_count=0,
    dtype=np.uint32,
):
    master = np.asarray(master, dtype=np.float32)
    weights = np.asarray(weight_map, dtype=np.float32)
    dq = np.zeros(master.shape, dtype=dtype)
    invalid = (~np.isfinite(master)) | (~np.isfinite(weights)) | (weights <= 0.0)
    if coverage_map is not None:
        cov = np.asarray(coverage_map, dtype=np.float32)
        bad = (~np.isfinite(cov)) | (cov <= 0.5)
        invalid |= bad
        dq[bad] |= DQ_WARP_EDGE
    if geometric_warp_coverage_map is not None:
        geo = np.asarray(geometric_warp_coverage_map, dtype=np.float32)
        finite = np.isfinite(geo)
        bad = (~finite) | (geo <= 0.5)
        invalid |= bad
        dq[bad] |= DQ_WARP_EDGE
        if int(active_frame_count) > 0:
            full = float(active_frame_count) - 0.5
            partial = finite & (geo > 0.5) & (geo < full)
            dq[partial] |= DQ_WARP_EDGE
    dq[invalid] |= DQ_NO_DATA
    if low_rejection_map is not None:
        low = np.asarray(low_rejection_map, dtype=np.float32)
        dq[np.isfinite(low) & (low > 0.0)] |= DQ_LOW_REJECTED
    if high_rejection_map is not None:
        high = np.asarray(high_rejection_map, dtype=np.float32)
        dq[np.isfinite(high) & (high > 0.0)] |= DQ_HIGH_REJECTED
    return dq
