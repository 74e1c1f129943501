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
