"""The oracle's integration of one output block of a full-size resident job.

For every frame the caller supplies the calibrated source crop that holds the
block's taps (``fetch(frame, (x0, y0, x1, y1)) -> float32 array``); the block
is warped with the oracle's windowed CPU-path warp (identity frames are taken
as they are, as the resident path bypasses them) and integrated with the
oracle's weighted_integrate_stack under the frame-level statistics choice.
Returns the oracle maps, the warped samples and their coverage, for
tests/contract.py.
"""

from __future__ import annotations

import numpy as np

import oracle
from contract import oracle_block
from oracle.warp import source_box, warp_matrix_window


def is_identity(matrix) -> bool:
    return bool(np.array_equal(np.asarray(matrix, np.float64), np.eye(3)))


def warped_block(fetch, matrices, block, full_shape, interpolation="lanczos3", fill=np.nan):
    """(samples [F, h, w] f32, coverage [F, h, w] f32) of the block."""
    x0, y0, w, h = block
    samples = np.empty((len(matrices), h, w), np.float32)
    cov = np.empty((len(matrices), h, w), np.float32)
    for f, m in enumerate(matrices):
        if is_identity(m):
            samples[f] = fetch(f, (x0, y0, x0 + w, y0 + h))
            cov[f] = 1.0
            continue
        bx0, by0, bx1, by1 = source_box(m, block, full_shape, interpolation)
        crop = fetch(f, (bx0, by0, bx1, by1))
        samples[f], cov[f] = warp_matrix_window(crop, m, block, origin=(bx0, by0), full_shape=full_shape,
                                                fill=fill, interpolation=interpolation)
    return samples, cov


def oracle_maps(samples, cov, weights, all_valid=False, geometric_frames=None, active=None):
    """Oracle maps (master, weight, coverage, low, high) and DQ of a block."""
    exp = oracle_block(samples, cov, weights, all_valid=all_valid)
    geo = np.sum(cov if geometric_frames is None else cov[geometric_frames], axis=0).astype(np.float32)
    n_active = int(np.count_nonzero(np.asarray(weights) > 0)) if active is None else active
    dq = oracle.resident_dq_map(exp[0], exp[1], exp[2], exp[3], exp[4], geo, active_frame_count=n_active)
    return exp, dq, geo
