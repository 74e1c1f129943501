"""CPU checks of the tolerance-contract comparator (tests/contract.py) and of
the oracle pieces the full-size GPU tests use on blocks of a frame: the
windowed warp and the frame-level statistics-arithmetic choice."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from contract import check_contract, near_threshold, oracle_block
from oracle.warp import source_box, warp_matrix_window


def _frame(seed, h=40, w=52):
    rng = np.random.default_rng(seed)
    return rng.normal(100.0, 5.0, (h, w)).astype(np.float32)


@pytest.mark.parametrize("interpolation", ["bilinear", "lanczos3"])
@pytest.mark.parametrize("window", [(0, 0, 52, 40), (3, 5, 20, 11), (40, 30, 12, 10), (0, 33, 52, 7)])
def test_window_warp_equals_full_frame_slice(interpolation, window):
    img = _frame(1)
    a = np.deg2rad(0.9)
    m = np.array([[np.cos(a), -np.sin(a), 2.3], [np.sin(a), np.cos(a), -3.1], [0, 0, 1.0]])
    full, fcov = oracle.warp_matrix(img, m, fill=np.nan, interpolation=interpolation)
    x0, y0, w, h = window
    bx0, by0, bx1, by1 = source_box(m, window, img.shape, interpolation)
    got, cov = warp_matrix_window(img[by0:by1, bx0:bx1], m, window, origin=(bx0, by0), full_shape=img.shape,
                                  fill=np.nan, interpolation=interpolation)
    assert np.array_equal(cov, fcov[y0:y0 + h, x0:x0 + w])
    assert np.array_equal(got, full[y0:y0 + h, x0:x0 + w], equal_nan=True)


def test_window_warp_refuses_a_short_crop():
    img = _frame(2)
    m = np.array([[1, 0, 0.5], [0, 1, 0.5], [0, 0, 1.0]])
    with pytest.raises(ValueError, match="crop"):
        warp_matrix_window(img[10:20, 10:20], m, (10, 10, 10, 10), origin=(10, 10), full_shape=img.shape)


def _stack(seed, f=24, h=9, w=13):
    rng = np.random.default_rng(seed)
    x = rng.normal(1.0, 0.01, (f, h, w)).astype(np.float32)
    hot = rng.random(x.shape) < 0.03
    x[hot] += np.float32(0.5)
    return x


def test_frame_level_arithmetic_choice():
    x = _stack(3)
    cov = np.ones_like(x)
    auto = oracle.weighted_integrate_stack(x, coverage=cov, rejection="winsorized_sigma")
    forced = oracle.weighted_integrate_stack(x, coverage=cov, rejection="winsorized_sigma", all_valid=True)
    for a, b in zip(auto, forced):
        assert np.array_equal(a, b)
    # a block of a frame whose other pixels hold invalid samples: the frame's
    # choice (per-pixel valid-count arithmetic) is imposed on the block
    mixed = x.copy()
    mixed[0, 0, 0] = np.nan
    valid_variant = oracle.weighted_integrate_stack(mixed, rejection="winsorized_sigma")
    block = oracle.weighted_integrate_stack(x[:, 1:, :], rejection="winsorized_sigma", all_valid=False)
    for a, b in zip(valid_variant, block):
        assert np.array_equal(a[1:, :], b)


def test_contract_accepts_identical_and_rejects_planted_errors():
    x = _stack(4)
    w = np.ones(x.shape[0], np.float32)
    cov = np.ones_like(x)
    exp = oracle_block(x, cov, w)
    assert check_contract(exp, exp).count_mismatches == 0
    # a master error of 1e-4 relative everywhere breaks the RMS bound
    bad = (exp[0] * np.float32(1.0001),) + tuple(exp[1:])
    with pytest.raises(AssertionError):
        check_contract(bad, exp)
    # a count flip at a pixel whose samples are far from both thresholds
    flipped = list(exp)
    flipped[4] = exp[4].copy()
    far = np.argwhere(~near_threshold(x, cov > 0.5, w))[0]
    flipped[4][tuple(far)] += 1
    with pytest.raises(AssertionError, match="away from a threshold"):
        check_contract(tuple(flipped), exp, x, cov > 0.5, w)


def test_contract_exempts_a_sample_on_a_threshold():
    from oracle.rejection import center_and_scale

    x = _stack(5)
    w = np.ones(x.shape[0], np.float32)
    cov = np.ones_like(x)
    # move frame 0's sample of pixel (2, 3) onto the high threshold; the
    # threshold depends (weakly) on that sample, so iterate to a fixed point
    x2 = x.copy()
    for _ in range(50):
        center, scale = center_and_scale(x2, cov > 0.5)
        hi = np.float32(center[2, 3] + np.float32(3.0) * scale[2, 3])
        if x2[0, 2, 3] == hi:
            break
        x2[0, 2, 3] = hi
    assert near_threshold(x2, cov > 0.5, w)[2, 3]
    exp = oracle_block(x2, cov, w)
    flipped = list(exp)
    flipped[4] = exp[4].copy()
    flipped[4][2, 3] += 1
    report = check_contract(tuple(flipped), exp, x2, cov > 0.5, w)
    assert report.count_mismatches == 1 and report.exempt == 1 and report.exempt_pixels == [(2, 3)]
