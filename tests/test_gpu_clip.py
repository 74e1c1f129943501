"""GPU parity of the mean / sigma-clip integrations that sample the frames
themselves (resident stack as stored, or warped on the fly) against
oracle/sigma_clip.py: bit-exact for the stored and bilinear samplers, and for
Lanczos3 against the library's own resident warp followed by the oracle's clip
(same sampler), within the Lanczos3 tolerance against the FP64 oracle."""

from __future__ import annotations

import numpy as np
import pytest

from glass_amd import ResidentCalibratedStack
from oracle import sigma_clip as SC

pytestmark = pytest.mark.gpu

MAPS = ("master", "weight", "coverage", "low", "high")


def _stack_of(frames):
    frames = np.asarray(frames, np.float32)
    s = ResidentCalibratedStack(frames.shape[0], frames.shape[1], frames.shape[2])
    for i, f in enumerate(frames):
        s.upload_calibrated_frame(i, f)
    return s


def _frames(f, h, w, seed, nan_frac=0.03):
    rng = np.random.default_rng(seed)
    yy, xx = np.indices((h, w), dtype=np.float32)
    base = 100.0 + 5.0 * np.sin(xx * 0.3) + 0.2 * yy
    x = (base[None] + rng.normal(0, 2.0, (f, h, w))).astype(np.float32)
    x[rng.random((f, h, w)) < nan_frac] = np.nan
    x[rng.random((f, h, w)) < 0.01] = 500.0  # outliers
    return x


def _mats(f, seed, rot=0.01):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(f):
        a = rng.uniform(-rot, rot)
        out.append(np.array([[np.cos(a), -np.sin(a), rng.uniform(-3, 3)], [np.sin(a), np.cos(a), rng.uniform(-3, 3)],
                             [0, 0, 1]], np.float64))
    return np.stack(out)


def test_reference_fixture_sigma_clip_and_download_modes(gpu):
    frames = [np.array([[1, 5], [10, 2]], np.float32), np.array([[1, 6], [10, 2]], np.float32),
              np.array([[1, 7], [10, 2]], np.float32), np.array([[100, 8], [10, -50]], np.float32)]
    with _stack_of(frames) as s:
        master, weight, coverage, low, high = s.integrate_sigma_clip(None, 1.0, 1.0, False)
        only = s.integrate_sigma_clip(None, 1.0, 1.0, False, download_mode="master_only")
        mw = s.integrate_sigma_clip(None, 1.0, 1.0, False, download_mode="master_weight")
    assert master.tolist() == [[1.0, 6.5], [10.0, 2.0]]
    assert coverage.tolist() == [[3, 2], [4, 3]] and weight.tolist() == [[3, 2], [4, 3]]
    assert low.tolist() == [[0, 1], [0, 1]] and high.tolist() == [[1, 1], [0, 0]]
    assert np.array_equal(only[0], master) and only[1:] == (None, None, None, None)
    assert np.array_equal(mw[1], weight) and mw[2:] == (None, None, None)


@pytest.mark.parametrize("f,winsorize", [(7, True), (23, False), (70, True), (150, True)])
def test_sigma_clip_stack_bit_exact(gpu, f, winsorize):
    x = _frames(f, 33, 41, seed=f)
    rng = np.random.default_rng(f + 1)
    weights = rng.uniform(0.5, 2.0, f).astype(np.float32)
    weights[::9] = 0.0
    exp = SC.clip_integrate(x, weights, "sigma_clip", 2.5, 2.0, winsorize)
    with _stack_of(x) as s:
        got = s.integrate_sigma_clip(weights, 2.5, 2.0, winsorize)
        after = s.download_frames_tile(list(range(f)), 0, 0, 41, 33)
    for name, g in zip(MAPS, got):
        assert np.array_equal(g, exp[name]), name
    assert np.array_equal(after, x, equal_nan=True)


@pytest.mark.parametrize("f,rejection", [(5, "mean"), (9, "sigma_clip"), (140, "sigma_clip"), (260, "mean")])
def test_matrix_warped_bilinear_bit_exact(gpu, f, rejection):
    x = _frames(f, 30, 38, seed=3 * f)
    mats = _mats(f, f)
    weights = np.ones(f, np.float32)
    weights[1] = 0.0
    exp = SC.matrix_warped_integrate(x, mats, weights, "bilinear", rejection=rejection, low_sigma=2.0,
                                     high_sigma=2.0, winsorize=True)
    with _stack_of(x) as s:
        if rejection == "mean":
            master, weight, coverage, geometric, timing = s.integrate_matrix_warped_mean(mats, weights, "bilinear")
            got = {"master": master, "weight": weight, "coverage": coverage}
            assert timing["modifies_resident_stack"] is False and timing["rejection"] == "none"
        else:
            master, weight, coverage, low, high, geometric, timing = s.integrate_matrix_warped_sigma_clip(
                mats, weights, "bilinear", -1.0, 2.0, 2.0, True)
            got = {"master": master, "weight": weight, "coverage": coverage, "low": low, "high": high}
            assert timing["rejection"] == "winsorized_sigma"
        after = s.download_frames_tile(list(range(f)), 0, 0, 38, 30)
    for name, g in got.items():
        assert np.array_equal(g, exp[name]), name
    assert np.array_equal(geometric, exp["geometric"])
    assert np.array_equal(after, x, equal_nan=True)  # the stack is not modified


@pytest.mark.parametrize("f,clamp", [(6, 0.3), (6, -1.0), (140, 0.3)])
def test_matrix_warped_lanczos3(gpu, f, clamp):
    # 6 frames: 128-thread LDS store; 140: 64-thread LDS store
    h, w = 40, 44
    x = _frames(f, h, w, seed=77)
    x[2, 10:14, 12:15] = np.nan  # non-finite taps are skipped
    mats = _mats(f, 5)
    weights = np.tile(np.array([1, 0.5, 2, 1, 0, 1], np.float32), f // 6 + 1)[:f]
    active = [k for k in range(f) if weights[k] > 0]
    with _stack_of(x) as s:
        master, weight, coverage, low, high, geometric, _ = s.integrate_matrix_warped_sigma_clip(
            mats, weights, "lanczos3", clamp, 2.0, 2.0, True)
        mean = s.integrate_matrix_warped_mean(mats, weights, "lanczos3", clamp)
    # the library's own resident warp (per-pixel direct path: clamp or NaN
    # sources) followed by the oracle's clip: the same samples
    with _stack_of(x) as s:
        s.reset_warp_coverage()
        s.warp_frames(active, mats[active], "lanczos3", "resident", fill=np.nan, clamping_threshold=clamp)
        warped = s.download_frames_tile(list(range(f)), 0, 0, w, h)
        geo = s.warp_coverage_map()
    if clamp >= 0:
        exp = SC.clip_integrate(warped, weights, "sigma_clip", 2.0, 2.0, True)
        for name, g in zip(MAPS, (master, weight, coverage, low, high)):
            assert np.array_equal(g, exp[name]), name
        assert np.array_equal(geometric, geo)
        exp_mean = SC.clip_integrate(warped, weights, "mean")
        assert np.array_equal(mean[0], exp_mean["master"]) and np.array_equal(mean[2], exp_mean["coverage"])
    # the FP64-weight oracle, within the Lanczos3 tolerance
    ref = SC.matrix_warped_integrate(x, mats, weights, "lanczos3", clamp, "mean")
    assert np.array_equal(mean[2], ref["coverage"]) and np.array_equal(mean[3], ref["geometric"])
    np.testing.assert_allclose(mean[0], ref["master"], rtol=0, atol=2e-3)


@pytest.mark.parametrize("clamp", [-1.0, 0.3])
def test_matrix_warped_lanczos3_staged_and_unstaged_boxes(gpu, clamp):
    # Affine Lanczos3 maps take the tile kernel: each frame's source box of a
    # 128 x 16 output tile is staged into LDS when it fits (near-identity
    # maps), and the taps come from the frame when it does not (35-degree
    # rotations, a zoom, a shear), within one launch and one tile; partial
    # tiles at the right and bottom edges.  Samples and sums must equal the
    # library's resident warp followed by the oracle's clip bit for bit.
    h, w = 70, 300
    x = _frames(8, h, w, seed=79)
    x[3, 30:33, 100:110] = np.nan
    def rot(deg, tx, ty, s=1.0):
        a = np.deg2rad(deg)
        return np.array([[s * np.cos(a), -s * np.sin(a), tx], [s * np.sin(a), s * np.cos(a), ty], [0, 0, 1.0]])
    mats = np.stack([rot(0.3, 1.5, -2.0), rot(35.0, 60.0, -120.0), rot(-0.2, -2.5, 1.0), rot(-35.0, -20.0, 90.0),
                     rot(0.0, 0.5, 0.5, 1.8), np.array([[1.0, 0.6, -15.0], [0.0, 1.0, 3.0], [0, 0, 1.0]]),
                     rot(0.5, 3.0, 2.0), rot(0.0, -4.0, -1.5)])
    weights = np.array([1, 0.5, 2, 1, 1, 0, 1, 1.5], np.float32)
    active = [k for k in range(8) if weights[k] > 0]
    print(f"fused lanczos3 staged/unstaged boxes clamp={clamp}", flush=True)
    with _stack_of(x) as s:
        clip = s.integrate_matrix_warped_sigma_clip(mats, weights, "lanczos3", clamp, 2.0, 2.0, True)
        plain = s.integrate_matrix_warped_sigma_clip(mats, weights, "lanczos3", clamp, 2.0, 2.0, False)
        mean = s.integrate_matrix_warped_mean(mats, weights, "lanczos3", clamp)
    with _stack_of(x) as s:
        s.reset_warp_coverage()
        s.warp_frames(active, mats[active], "lanczos3", "resident", fill=np.nan, clamping_threshold=clamp)
        warped = s.download_frames_tile(list(range(8)), 0, 0, w, h)
        geo = s.warp_coverage_map()
    if clamp >= 0:
        # the clamped resident warp samples every pixel with the same direct
        # arithmetic (unclamped, finite frames take the table-weight warp)
        for got, wins in ((clip, True), (plain, False)):
            exp = SC.clip_integrate(warped, weights, "sigma_clip", 2.0, 2.0, wins)
            for name, g in zip(MAPS, got[:5]):
                assert np.array_equal(g, exp[name]), (name, wins)
            assert np.array_equal(got[5], geo)
        exp_mean = SC.clip_integrate(warped, weights, "mean")
        assert np.array_equal(mean[0], exp_mean["master"]) and np.array_equal(mean[1], exp_mean["weight"])
        assert np.array_equal(mean[2], exp_mean["coverage"])
    assert np.array_equal(mean[3], geo)
    # the FP64-weight oracle, within the Lanczos3 tolerance: float32 direct
    # weights and float32 tap sums, whose rounding grows with the 500-valued
    # outliers' ringing (relative 1e-5 beside the absolute 2e-3 at ~100)
    ref = SC.matrix_warped_integrate(x, mats, weights, "lanczos3", clamp, "mean")
    assert np.array_equal(mean[2], ref["coverage"]) and np.array_equal(mean[3], ref["geometric"])
    np.testing.assert_allclose(mean[0], ref["master"], rtol=1e-5, atol=2e-3)


def test_clip_argument_errors(gpu):
    x = _frames(3, 8, 8, seed=1)
    with _stack_of(x) as s:
        with pytest.raises(ValueError, match="sigma thresholds must be positive"):
            s.integrate_sigma_clip(None, 0.0, 1.0)
        with pytest.raises(ValueError, match="download_mode"):
            s.integrate_sigma_clip(None, 1.0, 1.0, True, download_mode="everything")
        with pytest.raises(ValueError, match="interpolation"):
            s.integrate_matrix_warped_mean(np.stack([np.eye(3)] * 3), None, "cubic")
        with pytest.raises(ValueError, match="frame_count, 3, 3"):
            s.integrate_matrix_warped_mean(np.stack([np.eye(3)] * 2))
        with pytest.raises(ValueError, match="weights must have shape"):
            s.integrate_sigma_clip(np.ones(2, np.float32))
    with ResidentCalibratedStack(2, 8, 8) as s:
        s.upload_calibrated_frame(0, x[0])
        with pytest.raises(RuntimeError, match="must be loaded"):
            s.integrate_sigma_clip()


def test_integrations_on_two_streams_keep_their_tables(gpu):
    # Two asynchronous integrations with different frame sets on two caller
    # streams, queued back to back: the second upload of the frame table must
    # not overwrite the one the first launch is still reading.
    import torch

    f, h, w = 40, 256, 320
    x = _frames(f, h, w, seed=17, nan_frac=0.0)
    wa = np.ones(f, np.float32)
    wb = np.ones(f, np.float32)
    wb[::2] = 0.0
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    with _stack_of(x) as s:
        exp_a = s.integrate_hardened_winsorized_sigma(wa)[0]
        exp_b = s.integrate_hardened_winsorized_sigma(wb)[0]
        outs = [{"master": torch.empty((h, w), dtype=torch.float32, device="cuda")} for _ in range(2)]
        torch.cuda.synchronize()
        for _ in range(3):
            s.integrate_to_device(outs[0], wa, dq_geometric=False, stream=sa.cuda_stream)
            s.integrate_to_device(outs[1], wb, dq_geometric=False, stream=sb.cuda_stream)
        torch.cuda.synchronize()
        got_a, got_b = outs[0]["master"].cpu().numpy(), outs[1]["master"].cpu().numpy()
    assert np.array_equal(got_a, exp_a)
    assert np.array_equal(got_b, exp_b)


def _tile_case(f, h, w, seed):
    rng = np.random.default_rng(seed)
    x = _frames(f, h, w, seed)
    weights = rng.uniform(0.25, 2.0, f).astype(np.float32)
    weights[::7] = 0.0
    weights[3] = np.nan
    target = (rng.random(f) < 0.5).astype(np.uint8)
    target[1] = 1
    # non-overlapping tiles over part of the frame, one of multiplier 0
    ext = np.array([[0, 0, w // 3, h // 2], [w // 3, 0, w, h // 3], [w // 4, h // 2, w - 1, h],
                    [0, h // 2, w // 4, h - 1]], np.int32)
    mult = np.array([2.0, 0.0, 0.375, 1e38], np.float32)
    return x, weights, target, ext, mult


@pytest.mark.filterwarnings("ignore:overflow encountered")  # multiplier 1e38: inf sums on both sides
@pytest.mark.parametrize("f", [5, 40, 100, 150])  # the register, 128-slot and re-read kernels
@pytest.mark.parametrize("sigma", [False, True])
def test_tile_local_matches_the_oracle_bit_exact(gpu, f, sigma):
    h, w = 37, 53
    x, weights, target, ext, mult = _tile_case(f, h, w, 100 + f)
    with _stack_of(x) as s:
        if sigma:
            *maps, timing = s.integrate_tile_local_sigma_clip(target, ext, mult, weights, 2.5, 2.0, True)
        else:
            *maps, timing = s.integrate_tile_local_mean(target, ext, mult, weights)
    want = SC.tile_local_integrate(x, weights, target, ext, mult, "sigma_clip" if sigma else "mean", 2.5, 2.0, True)
    names = MAPS if sigma else ("master", "weight")
    assert len(maps) == len(names) and timing["tile_count"] == 4
    assert timing["target_frame_count"] == int(np.count_nonzero(target))
    for name, got in zip(names, maps):
        assert np.array_equal(got, want[name], equal_nan=True), name


@pytest.mark.parametrize("sigma", [False, True])
def test_tile_local_reference_fixtures(gpu, sigma):
    yy, xx = np.indices((5, 6), dtype=np.float32)
    if sigma:
        frames = [10.0 + xx * 0.1, 12.0 + yy * 0.2, 50.0 + xx * 0.05]
    else:
        frames = [10.0 + xx, 20.0 + yy * 2.0, 30.0 + xx * 0.25 + yy * 0.5]
    frames = np.stack(frames).astype(np.float32)
    if sigma:
        frames[2][0, 0] = 11.0
    frames[1][2, 3] = np.nan
    weights = np.array([1.0, 1.0, 0.5], np.float32)
    target = np.array([0, 1, 0], np.uint8)
    ext = np.array([[1, 1, 5, 4]], np.int32)
    mult = np.array([2.0], np.float32)
    with _stack_of(frames) as s:
        if sigma:
            *maps, timing = s.integrate_tile_local_sigma_clip(target, ext, mult, weights, 3.0, 1.0, True)
            assert timing["timing_model"] == "native_resident_tile_local_sigma_clip_one_sync"
            assert timing["rejection"] == "winsorized_sigma"
        else:
            *maps, timing = s.integrate_tile_local_mean(target, ext, mult, weights)
            plain_master, plain_weight = s.integrate_mean(weights)
            assert timing["timing_model"] == "native_resident_tile_local_weighted_mean_one_sync"
            assert not np.allclose(maps[0][1:4, 1:5], plain_master[1:4, 1:5])
            assert maps[1][0, 0] == plain_weight[0, 0]
    assert timing["target_frame_count"] == 1 and timing["tile_count"] == 1
    assert timing["modifies_resident_stack"] is False
    want = SC.tile_local_integrate(frames, weights, target, ext, mult, "sigma_clip" if sigma else "mean", 3.0, 1.0,
                                   True)
    for name, got in zip(MAPS, maps):
        assert np.array_equal(got, want[name]), name


def test_tile_local_rejects_bad_policies(gpu):
    x = _frames(3, 6, 8, 1)
    w = np.ones(3, np.float32)
    with _stack_of(x) as s:
        t = np.array([0, 1, 0], np.uint8)
        one = np.array([1.0], np.float32)
        with pytest.raises(ValueError, match="select at least one"):
            s.integrate_tile_local_mean(np.zeros(3, np.uint8), [[0, 0, 2, 2]], one, w)
        with pytest.raises(ValueError, match=r"shape \(tile_count, 4\)"):
            s.integrate_tile_local_mean(t, [0, 0, 2, 2], one, w)
        with pytest.raises(ValueError, match="between 1 and INT_MAX"):
            s.integrate_tile_local_mean(t, np.zeros((0, 4), np.int32), np.zeros(0, np.float32), w)
        with pytest.raises(ValueError, match="tile_multipliers must have shape"):
            s.integrate_tile_local_mean(t, [[0, 0, 2, 2]], [1.0, 2.0], w)
        for bad in ([[0, 0, 9, 2]], [[2, 0, 2, 2]], [[-1, 0, 2, 2]], [[0, 0, 2, 7]]):
            with pytest.raises(ValueError, match="half-open rectangles"):
                s.integrate_tile_local_mean(t, bad, one, w)
        for bad in (-1.0, np.inf, np.nan):
            with pytest.raises(ValueError, match="finite non-negative"):
                s.integrate_tile_local_mean(t, [[0, 0, 2, 2]], [bad], w)
        with pytest.raises(ValueError, match="must not overlap"):
            s.integrate_tile_local_mean(t, [[0, 0, 3, 3], [2, 2, 4, 4]], [1.0, 1.0], w)
        with pytest.raises(ValueError, match="sigma thresholds"):
            s.integrate_tile_local_sigma_clip(t, [[0, 0, 2, 2]], one, w, 0.0, 3.0)
        # touching tiles do not overlap; the stack is untouched by either call
        before = s.download_frame(0)
        s.integrate_tile_local_sigma_clip(t, [[0, 0, 3, 3], [3, 0, 8, 6]], [1.0, 0.5], w)
        assert np.array_equal(s.download_frame(0), before, equal_nan=True)
