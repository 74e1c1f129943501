"""Pins the mean / sigma-clip oracle (oracle/sigma_clip.py) on the CPU: the
reference's sigma-clip fixtures with their values worked out by hand, an FP64
brute force and the warp-then-integrate composition."""

from __future__ import annotations

import math
import statistics

import numpy as np
import pytest

from oracle import sigma_clip as SC
from oracle.warp import warp_matrix_resident


def test_reference_fixture_sigma_clip_by_hand():
    # tests/test_cuda_resident_stack.py:2744-2770, sigma 1/1, no winsorize
    frames = np.stack([
        np.array([[1, 5], [10, 2]], np.float32), np.array([[1, 6], [10, 2]], np.float32),
        np.array([[1, 7], [10, 2]], np.float32), np.array([[100, 8], [10, -50]], np.float32)])
    out = SC.clip_integrate(frames, None, "sigma_clip", 1.0, 1.0, False)
    # (0,0): 1,1,1,100 -> mean 25.75, std 42.87: 100 high
    # (0,1): 5..8 -> mean 6.5, std 1.118: 5 low, 8 high
    # (1,0): constant 10 -> std 0, nothing rejected
    # (1,1): 2,2,2,-50 -> mean -11, std 22.5: -50 low
    assert out["master"].tolist() == [[1.0, 6.5], [10.0, 2.0]]
    assert out["coverage"].tolist() == [[3, 2], [4, 3]]
    assert out["weight"].tolist() == [[3, 2], [4, 3]]
    assert out["low"].tolist() == [[0, 1], [0, 1]]
    assert out["high"].tolist() == [[1, 1], [0, 0]]


def test_reference_fixture_winsorized_and_zero_weight():
    # tests/test_cuda_resident_stack.py:2797-2826: 1,1,1,100 with one
    # winsorized re-estimate (clip to 68.62 -> mean 17.9, std 29.3): 100 high
    frames = np.stack([np.full((2, 2), v, np.float32) for v in (1, 1, 1, 100)])
    out = SC.clip_integrate(frames, None, "sigma_clip", 1.0, 1.0, True)
    assert np.all(out["master"] == 1.0) and np.all(out["high"] == 1.0) and np.all(out["coverage"] == 3.0)
    # :2773-2794: a zero-weight frame takes no part
    frames = np.stack([np.full((2, 2), v, np.float32) for v in (1, 1, 100)])
    out = SC.clip_integrate(frames, np.array([1, 1, 0], np.float32), "sigma_clip", 1.0, 1.0, False)
    assert np.all(out["master"] == 1.0) and np.all(out["coverage"] == 2.0) and np.all(out["high"] == 0.0)


def _brute(values, weights, lo_k, hi_k, winsorize):
    pairs = [(v, w) for v, w in zip(values, weights) if w > 0 and math.isfinite(v)]
    if not pairs:
        return 0.0, 0, 0, 0
    xs = [v for v, _ in pairs]
    mean, std = statistics.fmean(xs), statistics.pstdev(xs)
    lo, hi = mean - lo_k * std, mean + hi_k * std
    if winsorize:
        cl = [min(max(v, lo), hi) for v in xs]
        mean, std = statistics.fmean(cl), statistics.pstdev(cl)
        lo, hi = mean - lo_k * std, mean + hi_k * std
    keep = [(v, w) for v, w in pairs if lo <= v <= hi]
    nlo = sum(v < lo for v in xs)
    nhi = sum(v > hi for v in xs)
    ws = sum(w for _, w in keep)
    return (sum(v * w for v, w in keep) / ws if ws > 0 else 0.0), len(keep), nlo, nhi


@pytest.mark.parametrize("winsorize", [False, True])
def test_sigma_clip_matches_fp64_brute_force(winsorize):
    rng = np.random.default_rng(31 + winsorize)
    f, h, w = 23, 6, 7
    x = rng.normal(50.0, 4.0, (f, h, w)).astype(np.float32)
    x[rng.random((f, h, w)) < 0.05] = np.nan
    x[3, 2, 2] = 400.0
    x[7, 1, 5] = -300.0
    weights = rng.uniform(0.2, 2.0, f).astype(np.float32)
    weights[[4, 11]] = 0.0
    out = SC.clip_integrate(x, weights, "sigma_clip", 2.0, 2.5, winsorize)
    margin = 0
    for i in range(h):
        for j in range(w):
            m, cov, nlo, nhi = _brute([float(v) for v in x[:, i, j]], [float(v) for v in weights], 2.0, 2.5,
                                      winsorize)
            if (out["coverage"][i, j], out["low"][i, j], out["high"][i, j]) != (cov, nlo, nhi):
                margin += 1  # a sample within float32 rounding of a threshold
                continue
            assert out["master"][i, j] == pytest.approx(m, rel=2e-6)
    assert margin <= 1
    assert out["high"][2, 2] >= 1 and out["low"][1, 5] >= 1


def test_mean_matches_the_reference_warp_then_integrate_composition():
    # tests/test_cuda_resident_stack.py:649-735 (bilinear, translations, a NaN)
    yy, xx = np.indices((12, 14), dtype=np.float32)
    frames = [(10.0 + xx + yy * 0.25).astype(np.float32), (20.0 + xx * 0.5 + yy).astype(np.float32),
              (30.0 + xx * 0.1 - yy * 0.2).astype(np.float32)]
    frames[1][3, 4] = np.nan
    mats = np.stack([np.array([[1, 0, tx], [0, 1, ty], [0, 0, 1]], np.float64)
                     for tx, ty in ((0, 0), (1.25, -0.5), (-2.0, 1.0))])
    weights = np.array([1.0, 0.5, 0.0], np.float32)
    out = SC.matrix_warped_integrate(frames, mats, weights, "bilinear", rejection="mean")
    sums = np.zeros((12, 14), np.float64)
    wmap = np.zeros((12, 14), np.float64)
    cov = np.zeros((12, 14))
    geo = np.zeros((12, 14))
    for frame, m, wt in zip(frames, mats, weights):
        if wt <= 0:
            continue
        warped, fp = warp_matrix_resident(frame, m, fill=np.nan, interpolation="bilinear")
        fin = np.isfinite(warped)
        geo += fp > 0.5
        cov += fin
        sums[fin] += warped[fin] * wt
        wmap[fin] += wt
    exp = np.where(wmap > 0, sums / np.where(wmap > 0, wmap, 1), 0)
    np.testing.assert_allclose(out["master"], exp, rtol=2e-6, atol=2e-5)
    assert np.array_equal(out["coverage"], cov) and np.array_equal(out["geometric"], geo)
    np.testing.assert_allclose(out["weight"], wmap, rtol=1e-7)
    # the mean of the unwarped stack differs (the warp is applied)
    assert not np.allclose(out["master"], SC.clip_integrate(np.stack(frames), weights, "mean")["master"])


def _tile_fixture(sigma: bool):
    """Inputs of the reference's tile-local tests
    (tests/test_cuda_resident_stack.py:470-603)."""
    yy, xx = np.indices((5, 6), dtype=np.float32)
    if sigma:
        frames = [10.0 + xx * 0.1, 12.0 + yy * 0.2, 50.0 + xx * 0.05]
    else:
        frames = [10.0 + xx, 20.0 + yy * 2.0, 30.0 + xx * 0.25 + yy * 0.5]
    frames = [f.astype(np.float32) for f in frames]
    if sigma:
        frames[2][0, 0] = 11.0
    frames[1][2, 3] = np.nan
    return (np.stack(frames), np.array([1.0, 1.0, 0.5], np.float32), np.array([0, 1, 0], np.uint8),
            np.array([[1, 1, 5, 4]], np.int32), np.array([2.0], np.float32))


def test_tile_local_mean_reference_fixture_by_hand():
    stack, w, target, ext, mult = _tile_fixture(False)
    out = SC.tile_local_integrate(stack, w, target, ext, mult, "mean")
    # outside the tile: (10 + 20 + 0.5 * 30) / 2.5
    assert out["master"][0, 0] == pytest.approx(18.0, rel=1e-6) and out["weight"][0, 0] == 2.5
    # (2, 2) inside: frame 1 has weight 2; samples 12, 24, 31.5
    assert out["master"][2, 2] == pytest.approx((12 + 2 * 24 + 0.5 * 31.5) / 3.5, rel=1e-6)
    assert out["weight"][2, 2] == 3.5
    # (2, 3) inside with frame 1 NaN: samples 13, 31.75
    assert out["master"][2, 3] == pytest.approx((13 + 0.5 * 31.75) / 1.5, rel=1e-6)
    assert out["weight"][2, 3] == 1.5
    # row 4 is outside the tile (y1 = 4 is exclusive)
    assert out["weight"][4, 2] == 2.5 and out["weight"][3, 4] == 3.5 and out["weight"][3, 5] == 2.5


def test_tile_local_sigma_reference_fixture_by_hand():
    stack, w, target, ext, mult = _tile_fixture(True)
    out = SC.tile_local_integrate(stack, w, target, ext, mult, "sigma_clip", 3.0, 1.0, True)
    # (0, 0): samples 10, 12, 11; mean 11, std sqrt(2/3); winsorized at
    # 11 + 0.8165 the clipped mean 10.939 and std 0.743 put the high
    # threshold at 11.68: 12 is rejected, 10 (w 1) and 11 (w 0.5) kept
    assert out["master"][0, 0] == pytest.approx(15.5 / 1.5, rel=1e-6)
    assert (out["weight"][0, 0], out["coverage"][0, 0], out["low"][0, 0], out["high"][0, 0]) == (1.5, 2, 0, 1)
    # (1, 1) inside the tile: samples 10.1, 12.2, 50.05 with weights 1, 2, 0.5;
    # the winsorized high threshold (about 36.4) rejects 50.05
    assert out["master"][1, 1] == pytest.approx((10.1 + 2 * 12.2) / 3.0, rel=1e-6)
    assert (out["weight"][1, 1], out["coverage"][1, 1], out["high"][1, 1]) == (3.0, 2, 1)
    # (0, 1) outside: samples 10.1, 12.0, 50.05 with weights 1, 1, 0.5
    assert out["master"][0, 1] == pytest.approx(11.05, rel=1e-6) and out["weight"][0, 1] == 2.0


def _random_tiles(seed, f=7, h=9, w=11):
    rng = np.random.default_rng(seed)
    stack = rng.normal(100, 10, (f, h, w)).astype(np.float32)
    stack[rng.random(stack.shape) < 0.05] = np.nan
    stack[2, 4, 5] = 400.0
    weights = rng.uniform(0.5, 2.0, f).astype(np.float32)
    weights[1] = 0.0
    target = (rng.random(f) < 0.5).astype(np.uint8)
    target[0] = 1
    ext = np.array([[0, 0, 4, 3], [4, 0, 11, 5], [2, 5, 9, 9]], np.int32)
    return stack, weights, target, ext


@pytest.mark.parametrize("rejection", ["mean", "sigma_clip"])
def test_tile_local_invariants(rejection):
    stack, weights, target, ext = _random_tiles(5)
    args = (rejection, 2.5, 2.0, True)
    plain = SC.clip_integrate(stack, weights, *args)
    ones = SC.tile_local_integrate(stack, weights, target, ext, np.ones(3, np.float32), *args)
    for key in plain:
        assert np.array_equal(ones[key], plain[key]), key
    # multiplier 0: the target frames drop out of the tile's statistics
    zero = SC.tile_local_integrate(stack, weights, target, ext, np.array([1.0, 0.0, 1.0], np.float32), *args)
    without = SC.clip_integrate(stack[target == 0], weights[target == 0], *args)
    inside = np.zeros(stack.shape[1:], bool)
    inside[0:5, 4:11] = True
    for key in plain:
        assert np.array_equal(zero[key][inside], without[key][inside]), key
        assert np.array_equal(zero[key][~inside], plain[key][~inside]), key
    # multiplier 2 on every frame: the same master, twice the weight
    twice = SC.tile_local_integrate(stack, weights, np.ones(7, np.uint8), ext, np.full(3, 2.0, np.float32), *args)
    inside_any = SC.tile_multiplier_map(stack.shape[1:], ext, [2, 2, 2]) == 2
    assert np.array_equal(twice["master"], plain["master"])
    assert np.array_equal(twice["weight"][inside_any], 2 * plain["weight"][inside_any])
    assert np.array_equal(twice["weight"][~inside_any], plain["weight"][~inside_any])


def test_tile_multiplier_map_first_tile_wins():
    m = SC.tile_multiplier_map((4, 5), [[0, 0, 3, 2], [1, 1, 5, 4]], [3.0, 7.0])
    assert m[0, 0] == 3 and m[1, 2] == 3 and m[1, 3] == 7 and m[3, 0] == 1 and m[3, 4] == 7
