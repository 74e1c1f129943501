"""GPU parity of the resident local normalisation (pair-grid statistics,
global / grid application, the frame-loop driver) through the C ABI against
oracle/local_norm.py.  The application is bit-exact for given coefficients;
the statistics agree to float32 rounding (FP64 sums in a different order)."""

from __future__ import annotations

import numpy as np
import pytest

from glass_amd import ResidentCalibratedStack
from glass_amd import local_norm as GLN
from glass_amd import pipeline as P
from oracle import local_norm as LN
from oracle.quality import frame_global_stats

pytestmark = pytest.mark.gpu

STAT_KEYS = ("source_mean", "source_std", "reference_mean", "reference_std")


def _stack_of(frames):
    frames = np.asarray(frames, np.float32)
    s = ResidentCalibratedStack(frames.shape[0], frames.shape[1], frames.shape[2])
    for i, f in enumerate(frames):
        s.upload_calibrated_frame(i, f)
    return s


def _frames(f, h, w, seed, nan_frac=0.05):
    rng = np.random.default_rng(seed)
    base = rng.normal(400.0, 30.0, (h, w)).astype(np.float32)
    out = np.empty((f, h, w), np.float32)
    for i in range(f):
        gy, gx = np.indices((h, w), dtype=np.float32)
        out[i] = base * np.float32(rng.uniform(0.7, 1.4)) + np.float32(rng.uniform(-50, 50)) \
            + (gx * np.float32(rng.uniform(-0.2, 0.2))) + rng.normal(0, 3.0, (h, w)).astype(np.float32)
        out[i][rng.random((h, w)) < nan_frac] = np.nan
    return out


def _assert_stats_close(got, want):
    assert np.array_equal(np.asarray(got["valid_pixels"]), np.asarray(want["valid_pixels"]))
    for k in STAT_KEYS:
        np.testing.assert_allclose(got[k], want[k], rtol=3e-7, atol=0, err_msg=k)


@pytest.mark.parametrize("shape,tile", [((70, 90), (16, 16)), ((64, 96), (512, 512)), ((33, 37), (5, 7)),
                                        ((48, 40), (1, 40)), ((40, 300), (13, 256))])
def test_pair_grid_stats_match_oracle(gpu, shape, tile):
    frames = _frames(6, *shape, seed=shape[0] * 31 + tile[1])
    frames[2, : shape[0] // 2, :] = np.inf  # empty tiles for source 2
    frames[0, -3:, -3:] = np.nan            # reference holes
    with _stack_of(frames) as s:
        batch = s.frame_pair_grid_stats_batch(0, np.array([1, 2, 3, 4, 5, 0], np.int32), *tile)
        single = s.frame_pair_grid_stats(0, 3, *tile)
    assert batch["source_count"] == 6 and len(batch["frames"]) == 6
    for item in batch["frames"]:
        want = LN.pair_grid_stats(frames[item["source_index"]], frames[0], *tile)
        assert (item["grid_rows"], item["grid_cols"]) == (want["grid_rows"], want["grid_cols"])
        _assert_stats_close(item, want)
        assert item["valid_pixel_total"] == want["valid_pixel_total"]
    # the batched and the single-source forms are the same computation
    for k in (*STAT_KEYS, "valid_pixels"):
        assert np.array_equal(batch["frames"][2][k], single[k])


def test_pair_grid_stats_many_sources_chunks(gpu):
    # more sources than one launch batch holds (64)
    frames = _frames(70, 24, 20, seed=5)
    sources = np.arange(1, 70, dtype=np.int32)[::-1].copy()
    with _stack_of(frames) as s:
        batch = s.frame_pair_grid_stats_batch(0, sources, 8, 8)
    for item in batch["frames"]:
        _assert_stats_close(item, LN.pair_grid_stats(frames[item["source_index"]], frames[0], 8, 8))


def test_reference_fixture_batch_stats_match_per_frame(gpu):
    # tests/test_cuda_resident_stack.py:1439-1478
    reference = (np.arange(36, dtype=np.float32).reshape(6, 6) + np.float32(25.0)) / np.float32(3.0)
    moving_a = reference * np.float32(1.5) + np.float32(4.0)
    moving_b = reference * np.float32(0.75) - np.float32(2.0)
    moving_b[0, 0] = np.nan
    moving_b[5, 5] = np.nan
    with _stack_of([reference, moving_a, moving_b]) as s:
        batch = s.frame_pair_grid_stats_batch(0, np.array([1, 2], dtype=np.int32), 3, 2)
        assert batch["model"] == "resident_grid_pair_mean_std_batch"
        assert batch["batch_model"] == "single_kernel_source_frame_tile_grid"
        assert (batch["source_count"], batch["grid_rows"], batch["grid_cols"], batch["grid_count"]) == (2, 2, 3, 6)
        assert batch["download_bytes"] > 0 and batch["index_bytes"] == 8
        assert [f["source_index"] for f in batch["frames"]] == [1, 2]
        for frame_stats in batch["frames"]:
            single = s.frame_pair_grid_stats(0, frame_stats["source_index"], 3, 2)
            for k in (*STAT_KEYS, "valid_pixels"):
                assert np.array_equal(frame_stats[k], single[k])
            assert frame_stats["valid_pixel_total"] == single["valid_pixel_total"]
            want = LN.pair_grid_stats([moving_a, moving_b][frame_stats["source_index"] - 1], reference, 3, 2)
            _assert_stats_close(frame_stats, want)


def test_reference_fixture_grid_stats_drive_in_vram_normalization(gpu):
    # tests/test_cuda_resident_stack.py:1394-1436
    reference = np.arange(16, dtype=np.float32).reshape(4, 4) + np.float32(10.0)
    scales_true = np.array([[2.0, 0.5], [1.5, 3.0]], dtype=np.float32)
    offsets_true = np.array([[5.0, -2.0], [8.0, 1.0]], dtype=np.float32)
    moving = np.empty_like(reference)
    for row in range(2):
        for col in range(2):
            sl = np.s_[row * 2:row * 2 + 2, col * 2:col * 2 + 2]
            moving[sl] = reference[sl] * scales_true[row, col] + offsets_true[row, col]
    with _stack_of([reference, moving]) as s:
        stats = s.frame_pair_grid_stats(0, 1, 2, 2)
        assert stats["model"] == "resident_grid_pair_mean_std"
        assert (stats["grid_rows"], stats["grid_cols"]) == (2, 2)
        assert np.all(stats["valid_pixels"] == 4)
        sc = stats["reference_std"] / stats["source_std"]
        of = stats["reference_mean"] - stats["source_mean"] * sc
        profile = s.apply_grid_normalization_frame(1, sc, of, 2, 2)
        normalized, weight_map = s.integrate_mean(np.array([0.0, 1.0], dtype=np.float32))
    assert profile["mode"] == "in_place_device_update" and profile["temporary_output_bytes"] == 0
    assert np.allclose(normalized, reference, rtol=1e-5, atol=1e-5)
    assert np.allclose(weight_map, 1.0)


def test_reference_fixture_global_normalization(gpu):
    # tests/test_cuda_resident_stack.py:1319-1338
    reference = np.array([[10, 20], [30, 40]], dtype=np.float32)
    moving = reference * np.float32(2.0) + np.float32(5.0)
    with _stack_of([reference, moving]) as s:
        ref_stats = s.frame_global_stats(0)
        mov_stats = s.frame_global_stats(1)
        scale = ref_stats["std"] / mov_stats["std"]
        offset = ref_stats["mean"] - mov_stats["mean"] * scale
        profile = s.apply_global_normalization_frame(1, scale, offset)
        normalized, weight_map = s.integrate_mean(np.array([0.0, 1.0], dtype=np.float32))
        got = s.download_frame(1)
    assert profile["mode"] == "in_place_device_update"
    assert profile["temporary_output_bytes"] == 0 and profile["device_to_device_copy_s"] == 0.0
    assert np.allclose(normalized, reference, rtol=1e-5, atol=1e-5)
    assert np.allclose(weight_map, 1.0)
    assert np.array_equal(got, LN.apply_global(moving, scale, offset))


def test_reference_fixture_batch_apply_matches_per_frame(gpu):
    # tests/test_cuda_resident_stack.py:1341-1391
    frame_a = np.arange(35, dtype=np.float32).reshape(5, 7)
    frame_b = frame_a + np.float32(100.0)
    scales = np.stack([np.array([[1.0, 2.0], [0.5, 1.25]]), np.array([[0.75, 1.5], [2.0, 0.25]])]).astype(np.float32)
    offsets = np.stack([np.array([[0.0, -3.0], [4.0, 10.0]]), np.array([[5.0, -8.0], [2.0, 1.0]])]).astype(np.float32)
    expected = [LN.apply_grid(f, scales[k], offsets[k], 3, 4) for k, f in enumerate((frame_a, frame_b))]
    with _stack_of([frame_a, frame_b]) as s:
        profile = s.apply_grid_normalization_frames(np.array([0, 1], dtype=np.int32), scales, offsets, 3, 4)
        got = s.download_frames_tile([0, 1], 0, 0, 7, 5)
    assert profile["mode"] == "in_place_device_update_batch"
    assert profile["model"] == "resident_grid_mean_std_batch_apply"
    assert profile["frame_count"] == 2 and profile["coefficient_count"] == 8
    assert profile["coefficient_bytes"] == (scales[0].nbytes + offsets[0].nbytes) * 2
    assert [item["frame_index"] for item in profile["frames"]] == [0, 1]
    for k in range(2):
        assert np.array_equal(got[k], expected[k])
    with _stack_of([frame_a]) as s:
        s.apply_grid_normalization_frame(0, scales[0], offsets[0], 3, 4)
        assert np.array_equal(s.download_frame(0), expected[0])


@pytest.mark.parametrize("shape,tile", [((64, 96), (16, 32)), ((37, 53), (5, 7)), ((32, 40), (32, 12)),
                                        ((20, 8), (3, 3)), ((31, 64), (100, 100))])
def test_apply_grid_bit_exact(gpu, shape, tile):
    rng = np.random.default_rng(shape[1] * 3 + tile[0])
    frames = _frames(5, *shape, seed=9)
    rows, cols = LN.grid_shape(*shape, *tile)
    scales = rng.uniform(0.5, 2.0, (5, rows, cols)).astype(np.float32)
    offsets = rng.uniform(-100, 100, (5, rows, cols)).astype(np.float32)
    with _stack_of(frames) as s:
        s.apply_grid_normalization_frames(np.array([4, 1, 2], np.int32), scales[:3], offsets[:3], *tile)
        got = s.download_frames_tile(list(range(5)), 0, 0, shape[1], shape[0])
    assert np.array_equal(got[0], frames[0], equal_nan=True)
    assert np.array_equal(got[3], frames[3], equal_nan=True)
    for k, f in enumerate((4, 1, 2)):
        assert np.array_equal(got[f], LN.apply_grid(frames[f], scales[k], offsets[k], *tile), equal_nan=True)


def test_apply_repeated_frame_applies_in_order_and_flags_overflow(gpu):
    frames = _frames(3, 16, 24, seed=3, nan_frac=0.0)
    scales = np.full((3, 2, 3), 2.0, np.float32)
    offsets = np.full((3, 2, 3), 1.0, np.float32)
    scales[2] = 3.0
    with _stack_of(frames) as s:
        assert not s.frame_has_nonfinite(1)
        s.apply_grid_normalization_frames(np.array([1, 2, 1], np.int32), scales, offsets, 8, 8)
        got = s.download_frames_tile([1, 2], 0, 0, 24, 16)
        once = LN.apply_grid(frames[1], scales[0], offsets[0], 8, 8)
        assert np.array_equal(got[0], LN.apply_grid(once, scales[2], offsets[2], 8, 8))
        assert np.array_equal(got[1], LN.apply_grid(frames[2], scales[1], offsets[1], 8, 8))
        # an overflow to inf marks the frame as holding a non-finite sample
        s.apply_global_normalization_frame(2, 1.0e38, 0.0)
        assert s.frame_has_nonfinite(2)
        assert not s.frame_has_nonfinite(0)


def test_apply_device_coefficients_on_caller_stream(gpu):
    import torch

    frames = _frames(4, 40, 64, seed=21)
    rows, cols = LN.grid_shape(40, 64, 16, 16)
    rng = np.random.default_rng(2)
    scales = rng.uniform(0.5, 2.0, (2, rows, cols)).astype(np.float32)
    offsets = rng.uniform(-10, 10, (2, rows, cols)).astype(np.float32)
    stream = torch.cuda.Stream()
    with _stack_of(frames) as s:
        d_s = torch.from_numpy(scales).cuda()
        d_o = torch.from_numpy(offsets).cuda()
        torch.cuda.synchronize()
        s.apply_grid_normalization_frames(np.array([3, 0], np.int32), d_s, d_o, 16, 16, stream=stream.cuda_stream)
        stream.synchronize()
        got = s.download_frames_tile([3, 0], 0, 0, 64, 40)
    assert np.array_equal(got[0], LN.apply_grid(frames[3], scales[0], offsets[0], 16, 16), equal_nan=True)
    assert np.array_equal(got[1], LN.apply_grid(frames[0], scales[1], offsets[1], 16, 16), equal_nan=True)


def test_local_norm_argument_errors(gpu):
    frames = _frames(2, 8, 8, seed=1)
    with _stack_of(frames) as s:
        with pytest.raises(ValueError, match="tile dimensions must be positive"):
            s.frame_pair_grid_stats(0, 1, 0, 4)
        with pytest.raises(IndexError):
            s.frame_pair_grid_stats_batch(0, np.array([5], np.int32), 4, 4)
        with pytest.raises(ValueError, match="one-dimensional"):
            s.frame_pair_grid_stats_batch(0, np.array([[1]], np.int32), 4, 4)
        with pytest.raises(ValueError, match="does not match"):
            s.apply_grid_normalization_frame(1, np.ones((3, 2), np.float32), np.zeros((3, 2), np.float32), 4, 4)
        with pytest.raises(ValueError, match="must match"):
            s.apply_grid_normalization_frames(np.array([1], np.int32), np.ones((2, 2, 2), np.float32),
                                              np.zeros((2, 2, 2), np.float32), 4, 4)
        empty = s.apply_grid_normalization_frames(np.zeros(0, np.int32), np.ones((0, 2, 2), np.float32),
                                                  np.zeros((0, 2, 2), np.float32), 4, 4)
        assert empty["frame_count"] == 0 and empty["frames"] == []
        assert s.frame_pair_grid_stats_batch(0, np.zeros(0, np.int32), 4, 4)["frames"] == []
    with ResidentCalibratedStack(2, 8, 8) as s:
        s.upload_calibrated_frame(0, frames[0])
        with pytest.raises(RuntimeError, match="must be loaded"):
            s.frame_pair_grid_stats(0, 1, 4, 4)
        with pytest.raises(RuntimeError, match="must be loaded"):
            s.apply_global_normalization_frame(1, 1.0, 0.0)


@pytest.mark.parametrize("mode", ["global_mean_std", "grid_mean_std"])
def test_driver_matches_oracle(gpu, mode):
    frames = _frames(9, 70, 90, seed=77)
    frames[5] = np.nan                 # no pixels: weight 0, status empty
    frames[6, :40, :] = np.inf         # grid: empty tiles -> partial
    weights = np.array([1, 1, 1, 0, 1, 1, 1, 2, 1], np.float32)
    exp_frames, exp_w, exp_recs = LN.resident_local_normalization(frames, 2, weights, mode, tile_size=32)
    with _stack_of(frames) as s:
        res = GLN.resident_local_normalization(s, 2, weights, mode, tile_size=32)
        got = s.download_frames_tile(list(range(9)), 0, 0, 90, 70)
    assert res["weights"].tolist() == exp_w.tolist()
    assert [r["status"] for r in res["frames"]] == [r["status"] for r in exp_recs]
    for i, (rec, exp) in enumerate(zip(res["frames"], exp_recs)):
        if rec["status"] in ("reference", "skipped_zero_weight", "empty"):
            assert np.array_equal(got[i], frames[i], equal_nan=True)
            continue
        if mode == "global_mean_std":
            assert rec["scale"] == pytest.approx(exp["scale"], rel=1e-12)
            assert rec["offset"] == pytest.approx(exp["offset"], rel=1e-10, abs=1e-9)
            # bit-exact application of the coefficients the driver used
            assert np.array_equal(got[i], LN.apply_global(frames[i], rec["scale"], rec["offset"]), equal_nan=True)
        else:
            gc = rec["grid_coefficients"]
            np.testing.assert_allclose(gc["scales"], exp["scales"], rtol=2e-6)
            np.testing.assert_allclose(gc["offsets"], exp["offsets"], rtol=2e-6, atol=1e-3)
            assert np.array_equal(gc["valid_pixels"], exp["valid_pixels"])
            assert np.array_equal(got[i], LN.apply_grid(frames[i], gc["scales"], gc["offsets"], 32, 32),
                                  equal_nan=True)
        np.testing.assert_allclose(got[i], exp_frames[i], rtol=1e-5, atol=1e-3, equal_nan=True)
    # normalised frames carry the reference's statistics
    if mode == "global_mean_std":
        ref = frame_global_stats(frames[2])
        assert frame_global_stats(got[1])["mean"] == pytest.approx(ref["mean"], rel=1e-6)


def test_pipeline_with_local_normalization(gpu):
    frames = _frames(6, 48, 64, seed=13, nan_frac=0.0)
    eye = [np.eye(3)] * 6
    masters = {"bias": None, "dark": None, "flat": None}
    res = P.run_resident_stack(frames, [1.0] * 6, masters, eye, local_normalization="grid_mean_std",
                               local_normalization_tile_size=16, reference_index=0)
    exp_frames, exp_w, _ = LN.resident_local_normalization(frames, 0, np.ones(6, np.float32), "grid_mean_std", 16)
    # integrating normalised frames lands on the reference's level
    assert np.nanmean(res.master) == pytest.approx(float(np.mean(frames[0])), rel=1e-3)
    assert res.weights.tolist() == exp_w.tolist()


def test_frames_global_stats_batch_equals_per_frame(gpu):
    frames = _frames(70, 20, 24, seed=8)
    frames[4] = np.nan
    with _stack_of(frames) as s:
        idx = list(range(69, -1, -1))
        batch = s.frames_global_stats(idx)
        for k, i in enumerate(idx):
            single = s.frame_global_stats(i)
            for key in ("mean", "std", "valid_pixels", "nonfinite_pixels", "min", "max"):
                assert batch[k][key] == single[key], (i, key)
            exp = frame_global_stats(frames[i])
            assert batch[k]["valid_pixels"] == exp["valid_pixels"]
            assert batch[k]["mean"] == pytest.approx(exp["mean"], rel=1e-12, abs=1e-12)
            assert batch[k]["std"] == pytest.approx(exp["std"], rel=1e-9, abs=1e-12)


def test_apply_global_batch_equals_per_frame(gpu):
    frames = _frames(5, 33, 35, seed=12)
    scales = [0.5, 1.75, 3.0]
    offsets = [10.0, -7.25, 0.125]
    with _stack_of(frames) as s:
        s.apply_global_normalization_frames([3, 0, 4], scales, offsets)
        got = s.download_frames_tile(list(range(5)), 0, 0, 35, 33)
    for k, i in enumerate((3, 0, 4)):
        assert np.array_equal(got[i], LN.apply_global(frames[i], scales[k], offsets[k]), equal_nan=True)
    assert np.array_equal(got[1], frames[1], equal_nan=True)
