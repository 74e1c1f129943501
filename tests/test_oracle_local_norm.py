"""Pins the local-normalisation oracle (oracle/local_norm.py) without a GPU:
the reference's own fixtures, brute force with exact sums, the explicit
edge-tile partition and closed forms."""

from __future__ import annotations

import math
import statistics

import numpy as np
import pytest

from oracle import local_norm as LN
from oracle.quality import frame_global_stats


def _brute_tile_stats(src, ref, y0, y1, x0, x1):
    pairs = [(float(src[y, x]), float(ref[y, x])) for y in range(y0, y1) for x in range(x0, x1)
             if math.isfinite(float(src[y, x])) and math.isfinite(float(ref[y, x]))]
    if not pairs:
        return 0, 0.0, 0.0, 0.0, 0.0
    s = [p[0] for p in pairs]
    r = [p[1] for p in pairs]
    return len(pairs), statistics.fmean(s), statistics.pstdev(s), statistics.fmean(r), statistics.pstdev(r)


@pytest.mark.parametrize("shape,tile", [((7, 9), (4, 5)), ((13, 11), (3, 4)), ((5, 5), (5, 5)), ((6, 10), (8, 3))])
def test_pair_grid_stats_match_brute_force_with_exact_sums(shape, tile):
    rng = np.random.default_rng(sum(shape) * 7 + tile[0])
    src = rng.normal(300.0, 20.0, shape).astype(np.float32)
    ref = (src * np.float32(1.3) + rng.normal(0, 2.0, shape)).astype(np.float32)
    src[rng.random(shape) < 0.1] = np.nan
    ref[rng.random(shape) < 0.1] = np.inf
    got = LN.pair_grid_stats(src, ref, *tile)
    rows, cols = -(-shape[0] // tile[0]), -(-shape[1] // tile[1])
    assert (got["grid_rows"], got["grid_cols"]) == (rows, cols)
    for gy in range(rows):
        for gx in range(cols):
            # explicit partition: tile gy covers [gy*th, min(H, (gy+1)*th))
            y0, x0 = gy * tile[0], gx * tile[1]
            n, sm, sd, rm, rd = _brute_tile_stats(src, ref, y0, min(shape[0], y0 + tile[0]), x0,
                                                  min(shape[1], x0 + tile[1]))
            assert int(got["valid_pixels"][gy, gx]) == n
            for key, want in (("source_mean", sm), ("source_std", sd), ("reference_mean", rm),
                              ("reference_std", rd)):
                assert got[key][gy, gx] == pytest.approx(want, rel=2e-6, abs=1e-9), (key, gy, gx)


def test_pair_grid_stats_reference_fixture_two_sources_with_nan():
    # tests/test_cuda_resident_stack.py:1439-1478 (6x6, tiles 3 x 2)
    reference = (np.arange(36, dtype=np.float32).reshape(6, 6) + np.float32(25.0)) / np.float32(3.0)
    moving_a = reference * np.float32(1.5) + np.float32(4.0)
    moving_b = reference * np.float32(0.75) - np.float32(2.0)
    moving_b[0, 0] = np.nan
    moving_b[5, 5] = np.nan
    a = LN.pair_grid_stats(moving_a, reference, 3, 2)
    b = LN.pair_grid_stats(moving_b, reference, 3, 2)
    assert (a["grid_rows"], a["grid_cols"]) == (2, 3)
    assert np.all(a["valid_pixels"] == 6)
    assert b["valid_pixels"][0, 0] == 5 and b["valid_pixels"][1, 2] == 5 and b["valid_pixels"].sum() == 34
    # affine relation: std scales by |a|, mean maps affinely
    np.testing.assert_allclose(a["source_std"], 1.5 * a["reference_std"], rtol=1e-5)
    np.testing.assert_allclose(a["source_mean"], 1.5 * a["reference_mean"] + 4.0, rtol=1e-6)


def test_grid_stats_drive_exact_tile_recovery():
    # tests/test_cuda_resident_stack.py:1394-1436: per-tile affine moving frame
    reference = np.arange(16, dtype=np.float32).reshape(4, 4) + np.float32(10.0)
    scales_true = np.array([[2.0, 0.5], [1.5, 3.0]], dtype=np.float32)
    offsets_true = np.array([[5.0, -2.0], [8.0, 1.0]], dtype=np.float32)
    moving = np.empty_like(reference)
    for row in range(2):
        for col in range(2):
            sl = np.s_[row * 2:row * 2 + 2, col * 2:col * 2 + 2]
            moving[sl] = reference[sl] * scales_true[row, col] + offsets_true[row, col]
    stats = LN.pair_grid_stats(moving, reference, 2, 2)
    assert np.all(stats["valid_pixels"] == 4)
    coeffs = LN.grid_coefficients(stats)
    assert coeffs["ok_tiles"] == 4
    np.testing.assert_allclose(coeffs["scales"], 1.0 / scales_true, rtol=1e-6)
    out = LN.apply_grid(moving, coeffs["scales"], coeffs["offsets"], 2, 2)
    np.testing.assert_allclose(out, reference, rtol=1e-5, atol=1e-5)


def test_grid_fixture_from_cpu_local_norm_tests():
    # tests/test_cpu_local_norm.py:55-71 and tests/test_gpu_local_norm_vs_cpu.py:68-88
    yy, xx = np.indices((7, 9), dtype=np.float32)
    data = 10.0 + xx + yy * 0.5
    reference = data.copy()
    reference[:4, :5] = reference[:4, :5] * 1.5 + 7.0
    reference[:4, 5:] = reference[:4, 5:] * 0.75 - 3.0
    reference[4:, :5] = reference[4:, :5] * 1.2 + 2.0
    reference[4:, 5:] = reference[4:, 5:] * 0.9 + 11.0
    coeffs = LN.grid_coefficients(LN.pair_grid_stats(data, reference, 4, 5))
    out = LN.apply_grid(data, coeffs["scales"], coeffs["offsets"], 4, 5)
    for y0, y1 in [(0, 4), (4, 7)]:
        for x0, x1 in [(0, 5), (5, 9)]:
            assert np.allclose(np.mean(out[y0:y1, x0:x1]), np.mean(reference[y0:y1, x0:x1]), atol=1.0e-5)
            assert np.allclose(np.std(out[y0:y1, x0:x1]), np.std(reference[y0:y1, x0:x1]), atol=1.0e-5)


def test_apply_grid_edge_tiles_literal():
    # tests/test_cpu_local_norm.py:74-84
    data = np.arange(15, dtype=np.float32).reshape(3, 5)
    scales = np.array([[1.0, 2.0], [3.0, 4.0]], dtype=np.float32)
    offsets = np.array([[0.0, 10.0], [20.0, 30.0]], dtype=np.float32)
    out = LN.apply_grid(data, scales, offsets, 2, 3)
    assert np.array_equal(out[:2, :3], data[:2, :3])
    assert np.array_equal(out[:2, 3:], data[:2, 3:] * 2.0 + 10.0)
    assert np.array_equal(out[2:, :3], data[2:, :3] * 3.0 + 20.0)
    assert np.array_equal(out[2:, 3:], data[2:, 3:] * 4.0 + 30.0)
    # a grid shaped off the frame is refused
    with pytest.raises(ValueError):
        LN.apply_grid(data, scales[:1], offsets[:1], 2, 3)


def test_apply_grid_rounds_product_and_sum_separately():
    # x * s + o with two float32 roundings, not one fused operation
    x = np.float32(1.0000001)
    s = np.float32(3.0000002)
    o = np.float32(-3.0000005)
    out = LN.apply_grid(np.full((1, 1), x), np.full((1, 1), s), np.full((1, 1), o), 1, 1)
    two_roundings = np.float32(np.float32(x * s) + o)
    fused = np.float32(float(x) * float(s) + float(o))
    assert out[0, 0] == two_roundings
    assert two_roundings != fused


def test_coefficients_offset_only_and_empty_tiles():
    src = np.full((4, 4), 5.0, dtype=np.float32)
    ref = np.arange(16, dtype=np.float32).reshape(4, 4)
    src[:2, 2:] = np.nan  # tile (0, 1) empty
    stats = LN.pair_grid_stats(src, ref, 2, 2)
    coeffs = LN.grid_coefficients(stats)
    assert coeffs["statuses"][0][1] == "empty"
    assert coeffs["statuses"][0][0] == "offset_only"  # constant source
    assert (coeffs["empty_tiles"], coeffs["offset_only_tiles"], coeffs["ok_tiles"]) == (1, 3, 0)
    assert coeffs["scales"][0, 1] == 1.0 and coeffs["offsets"][0, 1] == 0.0
    assert stats["source_mean"][0, 1] == 0.0 and stats["reference_std"][0, 1] == 0.0
    # offset-only maps the source mean onto the reference mean
    assert coeffs["scales"][0, 0] == 1.0
    assert coeffs["offsets"][0, 0] == np.float32(np.float32(2.5) - np.float32(5.0))
    # the summary means run over tiles with paired pixels only
    active = np.asarray(stats["valid_pixels"]) > 0
    assert coeffs["offset_mean"] == pytest.approx(float(np.mean(coeffs["offsets"][active])))


def test_global_coefficients_reference_fixture():
    # tests/test_cuda_resident_stack.py:1319-1338: moving = 2 * reference + 5
    reference = np.array([[10, 20], [30, 40]], dtype=np.float32)
    moving = reference * np.float32(2.0) + np.float32(5.0)
    status, scale, offset = LN.global_coefficients(frame_global_stats(reference), frame_global_stats(moving))
    assert status == "ok"
    assert scale == pytest.approx(0.5, rel=1e-12)
    assert offset == pytest.approx(25.0 - 55.0 * 0.5, rel=1e-12)
    np.testing.assert_allclose(LN.apply_global(moving, scale, offset), reference, rtol=1e-5, atol=1e-5)
    assert LN.global_coefficients(frame_global_stats(reference), {"valid_pixels": 0})[0] == "empty"
    flat = frame_global_stats(np.full((2, 2), 3.0, np.float32))
    assert LN.global_coefficients(frame_global_stats(reference), flat) == ("offset_only", 1.0, 25.0 - 3.0)


@pytest.mark.parametrize("mode", ["global_mean_std", "grid_mean_std"])
def test_driver_statuses_and_weights(mode):
    rng = np.random.default_rng(11)
    frames = rng.normal(100.0, 5.0, (5, 12, 10)).astype(np.float32)
    frames[3] = np.nan                      # no pixels at all
    frames[4, :6, :] = np.nan               # grid: some tiles empty
    weights = np.array([1.0, 1.0, 0.0, 1.0, 1.0], np.float32)
    out, w, recs = LN.resident_local_normalization(frames, 0, weights, mode, tile_size=6)
    assert [r["status"] for r in recs][:4] == ["reference", "ok", "skipped_zero_weight", "empty"]
    assert recs[4]["status"] == ("partial" if mode == "grid_mean_std" else "ok")
    assert w.tolist() == [1.0, 1.0, 0.0, 0.0, 1.0]
    assert np.array_equal(out[0], frames[0]) and np.array_equal(out[2], frames[2])
    assert np.array_equal(out[3], frames[3], equal_nan=True)
    # normalised frames take the reference's statistics
    ref = frame_global_stats(frames[0])
    if mode == "global_mean_std":
        got = frame_global_stats(out[1])
        assert got["mean"] == pytest.approx(ref["mean"], rel=1e-6)
        assert got["std"] == pytest.approx(ref["std"], rel=1e-5)
