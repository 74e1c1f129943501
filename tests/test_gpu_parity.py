"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.  Bit-exact where the arithmetic is (calibration, bilinear,
rejection counts, rejection-free means), tolerance-based where the reference
itself only pins a tolerance (Lanczos3 weights in f32, FP64 sums in a
different order)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from glass_amd import ResidentCalibratedStack
from glass_amd import pipeline as P
from glass_amd.synthetic import encode_fits_u16be_bzero, make_set, to_adu

pytestmark = pytest.mark.gpu


def _stack_of(frames):
    frames = np.asarray(frames, np.float32)
    s = ResidentCalibratedStack(frames.shape[0], frames.shape[1], frames.shape[2])
    for i, f in enumerate(frames):
        s.upload_calibrated_frame(i, f)
    return s


# ------------------------------------------------------------- calibration
def test_calibrate_golden_314_bit_exact(gpu, golden_314):
    g = golden_314
    with ResidentCalibratedStack(3, 24, 24) as s:
        s.set_calibration_masters(g["master_bias"], g["master_dark"], g["master_flat"])
        for i in range(3):
            s.calibrate_frame(i, g["raw_light"][i], float(g["light_exposure_s"][i]), float(g["dark_exposure_s"][i]),
                              g["policy"])
        got = s.download_frames_tile([0, 1, 2], 0, 0, 24, 24)
    assert np.array_equal(got, g["calibrated_light"])


@pytest.mark.parametrize("includes_bias,scaling,pedestal", [(True, True, 0.0), (False, True, 0.0), (True, False, 5.0)])
def test_calibrate_formats_bit_exact(gpu, includes_bias, scaling, pedestal):
    rng = np.random.default_rng(42)
    h, w = 33, 47  # odd size exercises the scalar path
    bias = rng.uniform(90, 110, (h, w)).astype(np.float32)
    dark = rng.uniform(100, 140, (h, w)).astype(np.float32)
    flat = rng.uniform(0.5, 1.5, (h, w)).astype(np.float32)
    flat[3, 4] = 0.0  # below the floor
    adu = rng.integers(0, 65536, size=(3, h, w), dtype=np.uint16)
    policy = {"master_dark_includes_bias": includes_bias, "dark_scaling_enabled": scaling, "pedestal": pedestal,
              "flat_floor": 1e-6}
    with ResidentCalibratedStack(3, h, w) as s:
        s.set_calibration_masters(bias, dark, flat)
        s.calibrate_frame(0, adu[0].astype(np.float32), 37.0, 60.0, policy)
        raw = encode_fits_u16be_bzero(adu[1])
        s.calibrate_frames_fits_u16be_bzero_host_async_multistream_callback_release_timed(
            [1], [raw], [37.0], [60.0], 1, 1, None, policy)
        s.calibrate_frame(2, adu[2], 37.0, 60.0, policy)
        got = s.download_frames_tile([0, 1, 2], 0, 0, w, h)
    for i in range(3):
        exp = oracle.calibrate_light(adu[i], bias, dark, flat, 37.0, 60.0,
                                     master_dark_includes_bias=includes_bias, dark_scaling_enabled=scaling,
                                     pedestal=pedestal)
        assert np.array_equal(got[i], exp), i


def test_calibrate_device_raw_with_fused_stats(gpu):
    import torch

    rng = np.random.default_rng(3)
    h, w, f = 64, 96, 5
    adu = rng.integers(1000, 5000, size=(f, h, w), dtype=np.uint16)
    raw = torch.from_numpy(encode_fits_u16be_bzero(adu).reshape(f, -1).copy()).cuda()
    bias = rng.uniform(900, 1000, (h, w)).astype(np.float32)
    with ResidentCalibratedStack(f, h, w) as s:
        s.set_calibration_masters(bias=bias)
        s.calibrate_frames_device(list(range(f)), raw, "fits_u16be_bzero", [60.0] * f, None,
                                  accumulate_stats=True)
        got = s.download_frames_tile(list(range(f)), 0, 0, w, h)
        stats = [s.calibration_stats(i) for i in range(f)]
    for i in range(f):
        exp = oracle.calibrate_light(adu[i], bias, None, None, 60.0, None)
        assert np.array_equal(got[i], exp)
        ref = oracle.frame_global_stats(exp)
        assert stats[i]["valid_pixels"] == ref["valid_pixels"]
        assert stats[i]["mean"] == pytest.approx(ref["mean"], rel=1e-14)
        assert stats[i]["std"] == pytest.approx(ref["std"], rel=1e-9)


@pytest.mark.parametrize("stats", [False, True])
def test_calibrate_device_batches_flags_and_stats(gpu, stats):
    # 37 frames in shuffled order span three 16-frame batches (and 4-frame
    # statistics launches); a NaN flat pixel then makes every frame non-finite.
    import torch

    rng = np.random.default_rng(8)
    h, w, f = 40, 52, 37
    adu = rng.integers(500, 60000, size=(f, h, w), dtype=np.uint16)
    raw = torch.from_numpy(encode_fits_u16be_bzero(adu).reshape(f, -1).copy()).cuda()
    bias = rng.uniform(90, 110, (h, w)).astype(np.float32)
    dark = rng.uniform(100, 140, (h, w)).astype(np.float32)
    flat = rng.uniform(0.5, 1.5, (h, w)).astype(np.float32)
    nan_flat = flat.copy()
    nan_flat[7, 9] = np.nan
    exposures = list(rng.uniform(10, 90, f))
    policy = {"master_dark_includes_bias": False, "flat_floor": 1e-6}
    order = list(rng.permutation(f))
    with ResidentCalibratedStack(f, h, w) as s:
        s.set_calibration_masters(bias, dark, flat)
        s.calibrate_frames_device(order, raw[order], "fits_u16be_bzero", [exposures[i] for i in order], 60.0,
                                  policy, accumulate_stats=stats)
        got = s.download_frames_tile(list(range(f)), 0, 0, w, h)
        flags = [s.frame_has_nonfinite(i) for i in range(f)]
        st = [s.calibration_stats(i) for i in range(f)] if stats else None
        s.set_calibration_masters(bias, dark, nan_flat)
        s.calibrate_frames_device(list(range(f)), raw, "fits_u16be_bzero", exposures, 60.0, policy,
                                  accumulate_stats=stats)
        got_nan = s.download_frames_tile(list(range(f)), 0, 0, w, h)
        flags_nan = [s.frame_has_nonfinite(i) for i in range(f)]
    assert not any(flags) and all(flags_nan)
    for i in range(f):
        exp = oracle.calibrate_light(adu[i], bias, dark, flat, exposures[i], 60.0, master_dark_includes_bias=False)
        assert np.array_equal(got[i], exp), i
        exp_nan = oracle.calibrate_light(adu[i], bias, dark, nan_flat, exposures[i], 60.0,
                                         master_dark_includes_bias=False)
        assert np.array_equal(got_nan[i], exp_nan, equal_nan=True), i
        if stats:
            ref = oracle.frame_global_stats(exp)
            assert st[i]["valid_pixels"] == ref["valid_pixels"]
            assert st[i]["mean"] == pytest.approx(ref["mean"], rel=1e-13)
            assert st[i]["std"] == pytest.approx(ref["std"], rel=1e-9)


def test_calibrate_device_pixel_counts_off_the_batch_width(gpu):
    # 6 x 10 = 60 px (a multiple of 4, not of 8) and 7 x 9 = 63 px take the
    # per-frame kernels instead of the 8-pixel batched one; device raw frames
    # sit at a padded 128-byte stride.
    import torch

    rng = np.random.default_rng(9)
    for h, w in ((6, 10), (7, 9)):
        f = 3
        adu = rng.integers(500, 60000, size=(f, h, w), dtype=np.uint16)
        padded = np.zeros((f, 128), np.uint8)
        padded[:, :2 * h * w] = encode_fits_u16be_bzero(adu).reshape(f, -1)
        raw = torch.from_numpy(padded).cuda()
        bias = rng.uniform(90, 110, (h, w)).astype(np.float32)
        dark = rng.uniform(100, 140, (h, w)).astype(np.float32)
        flat = rng.uniform(0.5, 1.5, (h, w)).astype(np.float32)
        with ResidentCalibratedStack(f, h, w) as s:
            s.set_calibration_masters(bias, dark, flat)
            s.calibrate_frames_device(list(range(f)), raw, "fits_u16be_bzero", [30.0] * f, 60.0,
                                      {"master_dark_includes_bias": False}, frame_stride_bytes=128)
            got = s.download_frames_tile(list(range(f)), 0, 0, w, h)
        for i in range(f):
            exp = oracle.calibrate_light(adu[i], bias, dark, flat, 30.0, 60.0, master_dark_includes_bias=False)
            assert np.array_equal(got[i], exp), (h, w, i)


# ----------------------------------------------------------------- masters
def test_masters_golden_314_bit_exact(gpu, golden_314):
    g = golden_314
    pol = dict(g["policy"], tile_size=int(g["tile_size"]))
    mb = P.build_master_bias(g["raw_bias"], pol)
    md = P.build_master_dark(g["raw_dark"], mb, pol)
    mf, _ = P.build_master_flat(g["raw_flat"], mb, pol)
    assert np.array_equal(mb, g["master_bias"])
    assert np.array_equal(md, g["master_dark"])
    assert np.array_equal(mf, g["master_flat"])


def test_masters_with_rejection_match_oracle(gpu):
    rng = np.random.default_rng(17)
    frames = rng.normal(1000, 5, size=(20, 40, 36)).astype(np.float32)
    frames[3, 5, 5] = 5000.0
    frames[7, 10:12, 3] = -100.0
    frames[9, 20, 20] = np.nan  # one tile takes the per-pixel-count arithmetic
    pol = {"tile_size": 16, "master_dark_includes_bias": False}
    bias = rng.normal(100, 1, size=(40, 36)).astype(np.float32)
    got_b = P.build_master_bias(frames, pol)
    exp_b = oracle.build_master_bias(frames, pol)
    assert np.array_equal(got_b, exp_b)
    got_d = P.build_master_dark(frames, bias, pol)
    exp_d = oracle.build_master_dark(frames, bias, pol)
    assert np.array_equal(got_d, exp_d)
    flats = rng.normal(1.0, 0.01, size=(6, 40, 36)).astype(np.float32) * 20000
    got_f, sc = P.build_master_flat(flats, bias, pol)
    exp_f, esc = oracle.build_master_flat(flats, bias, pol)
    assert sc == esc
    assert np.array_equal(got_f, exp_f)


# -------------------------------------------------------------------- warp
def _rot(deg, tx, ty, scale=1.0):
    a = np.deg2rad(deg)
    return np.array([[np.cos(a) * scale, -np.sin(a) * scale, tx], [np.sin(a) * scale, np.cos(a) * scale, ty],
                     [0, 0, 1]], np.float32)


@pytest.mark.parametrize("interp", ["bilinear", "lanczos3"])
def test_warp_cpu_semantics_matches_oracle(gpu, interp):
    rng = np.random.default_rng(5)
    h, w = 57, 71
    frames = rng.normal(100, 10, size=(3, h, w)).astype(np.float32)
    mats = [_rot(0.7, 3.3, -2.6), _rot(-1.0, -0.4, 1.25, 1.0004), _rot(0.0, 2.0, 1.0)]
    with _stack_of(frames) as s:
        s.reset_warp_coverage()
        s.apply_matrix_frames_cpu_semantics([0, 1, 2], np.stack(mats), interp, fill=0.0)
        got = s.download_frames_tile([0, 1, 2], 0, 0, w, h)
        geo = s.warp_coverage_map()
    cov_sum = np.zeros((h, w), np.float32)
    for i in range(3):
        exp, cov = oracle.warp_matrix(frames[i], mats[i], fill=0.0, interpolation=interp)
        cov_sum += cov
        tol = 0.0 if interp == "bilinear" else 2e-4  # f32 Lanczos weights vs FP64 (values ~100)
        assert np.allclose(got[i], exp, rtol=0, atol=tol), (interp, i, np.max(np.abs(got[i] - exp)))
    assert np.array_equal(geo, cov_sum)


def test_warp_lanczos3_tiled_paths_match_oracle(gpu):
    # 150 x 300 spans 5 x 3 output tiles; each frame drives one branch of the
    # tiled kernel: LDS-staged taps (small rotation, zoom), global taps (40 deg
    # rotation: the source box exceeds the LDS tile), table taps with the
    # non-finite recheck (a frame holding NaN), and an integer shift (delta
    # weights: exact copy).
    rng = np.random.default_rng(11)
    h, w = 150, 300
    frames = rng.normal(100, 10, size=(5, h, w)).astype(np.float32)
    frames[2, 70, 150] = np.nan
    mats = [_rot(0.4, 2.7, -1.9), _rot(40.0, 60.0, -80.0), _rot(-0.6, 1.1, 0.3), _rot(0.0, 3.0, -2.0),
            _rot(0.2, -5.0, 4.0, 1.3)]
    with _stack_of(frames) as s:
        s.reset_warp_coverage()
        s.apply_matrix_frames_cpu_semantics(list(range(5)), np.stack(mats), "lanczos3", fill=0.0)
        got = s.download_frames_tile(list(range(5)), 0, 0, w, h)
        geo = s.warp_coverage_map()
    cov_sum = np.zeros((h, w), np.float32)
    for i in range(5):
        exp, cov = oracle.warp_matrix(frames[i], mats[i], fill=0.0, interpolation="lanczos3")
        cov_sum += cov
        assert np.array_equal(np.isnan(got[i]), np.isnan(exp)), i
        ok = ~np.isnan(exp)
        assert np.allclose(got[i][ok], exp[ok], rtol=0, atol=2e-4), (i, np.max(np.abs(got[i][ok] - exp[ok])))
    assert np.array_equal(geo, cov_sum)
    # integer shift: output (y, x) takes source (y + 2, x - 3) exactly
    assert np.array_equal(got[3][2:h - 5, 5:w - 3], frames[3][4:h - 3, 2:w - 6])


def test_warp_lanczos3_cpu_semantics_nonfinite_sources(gpu):
    # CPU semantics over sources holding NaN / +-inf: the table paths keep
    # finite pixels and recompute the non-finite ones per tap.  The reference
    # takes every tap with a non-zero FP64 weight, which at an integer shift
    # is every tap but the last of each axis: a NaN spreads to the pixels
    # whose 5 x 5 taps hold it, exactly as in the oracle.
    rng = np.random.default_rng(13)
    h, w = 140, 290
    frames = rng.normal(100, 10, size=(4, h, w)).astype(np.float32)
    frames[0, 60:63, :] = np.nan
    frames[0, 20, 40] = np.inf
    frames[0, 100, 200] = -np.inf
    for f in (1, 2, 3):
        ys, xs = rng.integers(5, h - 5, 25), rng.integers(5, w - 5, 25)
        frames[f, ys, xs] = np.nan
    frames[3, 30, 10:280] = np.nan
    mats = [_rot(0.5, 1.7, -2.3), _rot(0.0, 3.0, -2.0), _rot(40.0, 60.0, -80.0), _rot(0.3, -4.0, 2.5, 1.3)]
    with _stack_of(frames) as s:
        s.reset_warp_coverage()
        s.apply_matrix_frames_cpu_semantics(list(range(4)), np.stack(mats), "lanczos3", fill=np.nan)
        got = s.download_frames_tile(list(range(4)), 0, 0, w, h)
        geo = s.warp_coverage_map()
    cov_sum = np.zeros((h, w), np.float32)
    for i in range(4):
        exp, cov = oracle.warp_matrix(frames[i], mats[i], fill=np.nan, interpolation="lanczos3")
        cov_sum += cov
        assert np.array_equal(np.isfinite(got[i]), np.isfinite(exp)), i
        nf = ~np.isfinite(exp)
        assert np.array_equal(got[i][nf], exp[nf].astype(np.float32), equal_nan=True), i
        ok = ~nf
        assert np.allclose(got[i][ok], exp[ok], rtol=0, atol=2e-4), (i, np.max(np.abs(got[i][ok] - exp[ok])))
        assert np.isnan(got[i]).any() and ok.sum() > h * w // 2, i
    assert np.array_equal(geo, cov_sum)
    # integer shift: every finite output pixel is its source pixel exactly
    shifted, src = got[1][2:h - 5, 5:w - 3], frames[1][4:h - 3, 2:w - 6]
    fin = np.isfinite(shifted)
    assert np.array_equal(shifted[fin], src[fin]) and (~fin).sum() > 25 * 20


@pytest.mark.parametrize("interp", ["lanczos3", "bilinear"])
def test_warp_edge_geometries(gpu, interp):
    # Frames smaller than the tap window and than one output tile, a 90 degree
    # rotation, a map that leaves every output pixel outside the footprint, and
    # a projective map (direct kernel), against the oracle with NaN fill.
    rng = np.random.default_rng(14)
    proj = np.array([[1.0, 0.002, 1.5], [-0.003, 1.0, -2.0], [2e-5, -1e-5, 1.0]])
    cases = [((5, 7), _rot(0.3, 0.4, -0.2)), ((6, 6), _rot(0.0, 1.0, 0.0)), ((20, 300), _rot(90.0, 160.0, -140.0)),
             ((150, 300), _rot(0.5, 1000.0, 0.0)), ((70, 90), proj), ((37, 250), _rot(-1.0, 3.3, 0.7, 0.9))]
    for (h, w), m in cases:
        frame = rng.normal(100, 10, size=(h, w)).astype(np.float32)
        with _stack_of(frame[None]) as s:
            s.reset_warp_coverage()
            s.apply_matrix_frames_cpu_semantics([0], np.asarray(m)[None], interp, fill=np.nan)
            got = s.download_frame(0)
            geo = s.warp_coverage_map()
        exp, cov = oracle.warp_matrix(frame, m, fill=np.nan, interpolation=interp)
        fin = np.isfinite(exp)
        assert np.array_equal(np.isfinite(got), fin), (h, w)
        assert np.array_equal(geo, cov), (h, w)
        tol = 0.0 if interp == "bilinear" else 2e-4
        assert np.allclose(got[fin], exp[fin], rtol=0, atol=tol), (h, w, np.max(np.abs(got[fin] - exp[fin]), initial=0))


def test_warp_lanczos3_resident_tiled_matches_oracle(gpu):
    rng = np.random.default_rng(12)
    h, w = 100, 260
    frames = rng.normal(500, 20, size=(3, h, w)).astype(np.float32)
    frames[1, 40:43, 100] = np.nan
    mats = [_rot(0.9, -1.4, 2.2), _rot(-0.3, 0.6, -0.8), _rot(25.0, 30.0, -40.0)]
    with _stack_of(frames) as s:
        s.reset_warp_coverage()
        s.apply_matrix_lanczos3_frames([0, 1, 2], np.stack(mats), np.nan)
        got = s.download_frames_tile([0, 1, 2], 0, 0, w, h)
        geo = s.warp_coverage_map()
    cov_sum = np.zeros((h, w), np.float32)
    for i in range(3):
        exp, cov = oracle.warp_matrix_resident(frames[i], mats[i], np.nan, "lanczos3")
        cov_sum += cov
        assert np.array_equal(np.isfinite(got[i]), cov > 0), i
        ok = cov > 0
        assert np.allclose(got[i][ok], exp[ok], rtol=0, atol=1e-3), (i, np.max(np.abs(got[i][ok] - exp[ok])))
    assert np.array_equal(geo, cov_sum)


def test_warp_resident_semantics_nan_taps_and_clamp(gpu):
    yy, xx = np.indices((12, 13), dtype=np.float32)
    data = 0.2 * xx + 0.4 * yy
    data[:, 7:] += 5.0
    data[6, 6] = np.nan
    m = _rot(1.0, 0.35, -0.25)
    with _stack_of(data[None]) as s:
        s.apply_matrix_lanczos3_frame(0, m, np.nan, 0.30)
        got = s.download_frame(0)
    exp, cov = oracle.warp_matrix_resident(data, m, np.nan, "lanczos3", 0.30)
    assert np.array_equal(np.isfinite(got), cov > 0)
    assert np.allclose(got[cov > 0], exp[cov > 0], atol=3e-5)


@pytest.mark.parametrize("t", [0.0, 0.01])
def test_warp_resident_clamp_step_edge_closed_form(gpu, t):
    # A half-pixel shift of a 0/100 step rings ~11% past both plateaus: the
    # clamp must put those pixels exactly on 100 + 100 t / -100 t (float32),
    # and leave the others at the unclamped Lanczos3 value.
    yy, xx = np.indices((16, 20), dtype=np.float32)
    img = np.where(xx >= 10, 100.0, 0.0).astype(np.float32)
    m = np.array([[1, 0, 0.5], [0, 1, 0.0], [0, 0, 1.0]])
    with _stack_of(img[None]) as s:
        s.apply_matrix_lanczos3_frame(0, m, np.nan, t)
        got = s.download_frame(0)
    raw, cov = oracle.warp_matrix_resident(img, m, np.nan, "lanczos3", -1.0)
    ok = cov > 0
    hi, lo = np.float32(100.0 + 100.0 * t), np.float32(-100.0 * t)
    over, under = ok & (raw > hi + 1.0), ok & (raw < lo - 1.0)
    assert over.sum() >= 10 and under.sum() >= 10
    assert np.all(got[over] == hi) and np.all(got[under] == lo)
    inside = ok & (raw < hi - 1.0) & (raw > lo + 1.0)
    assert np.allclose(got[inside], raw[inside], rtol=0, atol=2e-4)
    assert np.array_equal(np.isfinite(got), ok)


def test_warp_reference_case_bilinear(gpu):
    # tests/test_gpu_warp_vs_cpu.py:154-173 on the resident class
    data = np.arange(49, dtype=np.float32).reshape(7, 7)
    m = _rot(3.0, 1.2, -0.4)
    with _stack_of(data[None]) as s:
        s.apply_matrix_bilinear_frame(0, m, -1.0)
        got = s.download_frame(0)
    exp, cov = oracle.warp_matrix_resident(data, m, -1.0, "bilinear")
    assert np.array_equal(got, exp)


# --------------------------------------------------------------- integrate
def _random_stack(seed, f, h, w, nan_frac=0.02, outliers=True):
    rng = np.random.default_rng(seed)
    stack = rng.normal(1000.0, 4.0, size=(f, h, w)).astype(np.float32)
    if outliers:
        idx = rng.random(stack.shape) < 0.01
        stack[idx] += rng.uniform(50, 500, size=int(idx.sum())).astype(np.float32)
        idx = rng.random(stack.shape) < 0.005
        stack[idx] -= rng.uniform(50, 500, size=int(idx.sum())).astype(np.float32)
    if nan_frac:
        stack[rng.random(stack.shape) < nan_frac] = np.nan
    return stack


def _compare(got, exp, exact_counts=True):
    gm, gw, gc, gl, gh = got
    em, ew, ec, el, eh = exp
    assert np.array_equal(gc, ec)
    assert np.array_equal(gl, el)
    assert np.array_equal(gh, eh)
    assert np.array_equal(gw, ew) or np.allclose(gw, ew, rtol=1e-7)
    assert np.allclose(gm, em, rtol=2e-7, atol=0), np.max(np.abs(gm - em) / np.maximum(np.abs(em), 1))


@pytest.mark.parametrize("frames", [1, 2, 3, 5, 16, 17, 33, 64, 65, 100, 129, 193, 208, 209, 256])
def test_winsorized_unit_weights_matches_oracle(gpu, frames):
    stack = _random_stack(frames, frames, 9, 31)
    print(f"winsorized unit weights frames={frames}", flush=True)
    with _stack_of(stack) as s:
        got = s.integrate_hardened_winsorized_sigma(None, 3.0, 3.0, 3, 0.5)
    exp = oracle.weighted_integrate_stack(stack, rejection="winsorized_sigma")
    _compare(got, exp)


@pytest.mark.parametrize("frames,weighted", [(24, False), (200, False), (256, True), (300, False), (300, True)])
def test_winsorized_sums_straddling_zero_and_wide_magnitudes(gpu, frames, weighted):
    # Samples of both signs spanning ~2^40 in magnitude, with NaNs: FP64 sums
    # of such samples are not exact, so the kernel must add them in the
    # oracle's frame order for means, thresholds and counts to match.
    rng = np.random.default_rng(700 + frames + weighted)
    h, w = 7, 29
    mag = 10.0 ** rng.uniform(-6, 6, size=(frames, h, w))
    stack = (np.where(rng.random((frames, h, w)) < 0.5, -1.0, 1.0) * mag).astype(np.float32)
    stack[:, :, :10] = rng.normal(0.0, 1e-3, (frames, h, 10)).astype(np.float32) + np.float32(1e-4)
    stack[rng.random(stack.shape) < 0.02] = np.nan
    weights = rng.uniform(0.5, 2.0, frames).astype(np.float32) if weighted else None
    print(f"winsorized wide-magnitude sums frames={frames} weighted={weighted}", flush=True)
    with _stack_of(stack) as s:
        got = s.integrate_hardened_winsorized_sigma(weights, 3.0, 3.0, 3, 0.5)
    exp = oracle.weighted_integrate_stack(stack, weights=weights, rejection="winsorized_sigma")
    _compare(got, exp)


@pytest.mark.parametrize("frames,weighted", [(16, False), (48, True), (64, False), (128, True), (160, False),
                                             (193, False), (200, False), (200, True), (256, False)])
def test_winsorized_pruned_network_waves(gpu, frames, weighted):
    # Waves whose valid counts all lie within 8 of the sort capacity take a
    # pruned network that orders only the quartile ranks and the ends; each
    # 64-pixel row below is one wave.  Rows: clean; 20-40 low and 25-35 high
    # outliers per pixel (the clip scans walk past the ordered ends into the
    # unordered middle); the same with one NaN in the row (full network);
    # integer-valued samples (ties); three NaNs per pixel (+inf keys inside
    # the window when the capacity allows); outliers in every other pixel.
    rng = np.random.default_rng(900 + frames + weighted)
    h, w = 6, 64
    stack = rng.normal(1000.0, 4.0, size=(frames, h, w)).astype(np.float32)
    n_lo, n_hi = min(40, frames // 5), min(35, frames // 6)
    for row in (1, 2, 5):
        for x in range(0 if row != 5 else 1, w, 1 if row != 5 else 2):
            k_lo = int(rng.integers(n_lo // 2, n_lo + 1))
            k_hi = int(rng.integers(n_hi // 2, n_hi + 1))
            idx = rng.permutation(frames)
            stack[idx[:k_lo], row, x] -= rng.uniform(50, 300, k_lo).astype(np.float32)
            stack[idx[k_lo:k_lo + k_hi], row, x] += rng.uniform(50, 300, k_hi).astype(np.float32)
    stack[int(rng.integers(frames)), 2, 17] = np.nan
    stack[:, 3, :] = np.round(stack[:, 3, :] / 3.0)
    for x in range(w):
        stack[rng.permutation(frames)[:3], 4, x] = np.nan
    weights = rng.uniform(0.5, 2.0, frames).astype(np.float32) if weighted else None
    print(f"winsorized pruned-network waves frames={frames} weighted={weighted}", flush=True)
    with _stack_of(stack) as s:
        got = s.integrate_hardened_winsorized_sigma(weights, 3.0, 3.0, 3, 0.5)
    exp = oracle.weighted_integrate_stack(stack, weights=weights, rejection="winsorized_sigma")
    _compare(got, exp)
    # every outlier was rejected: from 200 frames on, more than the 16 / 24
    # ordered end slots
    assert np.min(exp[3][1]) >= n_lo // 2 and np.min(exp[4][1]) >= n_hi // 2


@pytest.mark.parametrize("frames", [4, 20, 40, 120])
def test_winsorized_weights_and_zero_weight_frames(gpu, frames):
    rng = np.random.default_rng(100 + frames)
    stack = _random_stack(frames, frames, 8, 24)
    weights = rng.uniform(0.5, 2.0, frames).astype(np.float32)
    weights[rng.random(frames) < 0.15] = 0.0
    print(f"winsorized weighted frames={frames}", flush=True)
    with _stack_of(stack) as s:
        got = s.integrate_hardened_winsorized_sigma(weights, 2.5, 3.0, 2, 0.4)
    exp = oracle.weighted_integrate_stack(stack, weights=weights, rejection="winsorized_sigma", low_sigma=2.5,
                                          high_sigma=3.0, min_samples=2, max_reject_fraction=0.4)
    _compare(got, exp)


def test_winsorized_all_valid_variant(gpu):
    stack = _random_stack(7, 25, 10, 12, nan_frac=0.0)
    with _stack_of(stack) as s:
        got = s.integrate_hardened_winsorized_sigma()
    exp = oracle.weighted_integrate_stack(stack, rejection="winsorized_sigma")
    _compare(got, exp)


def test_winsorized_reference_literals(gpu):
    frames = np.stack([
        np.array([[1, 1], [10, 4]], np.float32), np.array([[1, 2], [10, 4]], np.float32),
        np.array([[1, 3], [10, 4]], np.float32), np.array([[12, 4], [10, 40]], np.float32),
    ])
    weights = np.array([1.0, 2.0, 1.0, 1.0], np.float32)
    with _stack_of(frames) as s:
        master, weight, cov, low, high, prof = s.integrate_hardened_winsorized_sigma_timed(weights, 2.4, 2.4)
    exp = oracle.weighted_integrate_stack(frames, weights=weights, rejection="winsorized_sigma", low_sigma=2.4,
                                          high_sigma=2.4)
    _compare((master, weight, cov, low, high), exp)
    assert high[0, 0] == 1 and high[1, 1] == 1
    frames = np.stack([np.full((2, 3), 10, np.float32)] * 4 + [np.array([[100, 10, 10], [10, -50, 10]], np.float32)])
    with _stack_of(frames) as s:
        got = s.integrate_hardened_winsorized_sigma(np.ones(5, np.float32), 2.0, 2.0, min_samples=1,
                                                    max_reject_fraction=1.0, count_map_dtype="uint16")
    assert got[2].dtype == np.uint16
    assert int(got[4][0, 0]) == 1 and int(got[3][1, 1]) == 1
    exp = oracle.weighted_integrate_stack(frames, weights=np.ones(5, np.float32), rejection="winsorized_sigma",
                                          low_sigma=2.0, high_sigma=2.0, min_samples=1, max_reject_fraction=1.0)
    _compare(tuple(np.asarray(x, np.float32) for x in got), exp)


def test_winsorized_empty_and_guard_edge_cases(gpu):
    stack = np.full((6, 3, 4), np.nan, np.float32)
    stack[:, 0, 0] = [1, 1, 1, 1, 1, 50]
    stack[:2, 1, 1] = [3, 4]
    with _stack_of(stack) as s:
        got = s.integrate_hardened_winsorized_sigma(None, 3.0, 3.0, 3, 0.5)
        mean = s.integrate_mean()
    exp = oracle.weighted_integrate_stack(stack, rejection="winsorized_sigma")
    _compare(got, exp)
    em, ew = oracle.mean_integrate([np.nan_to_num(f, nan=0.0) for f in stack])  # only for shape
    assert mean[0].shape == em.shape


@pytest.mark.parametrize("frames", [257, 300, 512, 513, 1100])
def test_winsorized_many_frames_lds_path(gpu, frames):
    # Beyond the register kernels' 256 samples: the LDS-sort path (512, 1024
    # and 2048-key segments; 32, 16 and 8 pixels per block), NaN samples.
    stack = _random_stack(frames, frames, 5, 13)
    print(f"winsorized lds frames={frames}", flush=True)
    with _stack_of(stack) as s:
        got = s.integrate_hardened_winsorized_sigma(None, 3.0, 3.0, 3, 0.5)
    exp = oracle.weighted_integrate_stack(stack, rejection="winsorized_sigma")
    _compare(got[:5], exp)


def test_winsorized_many_frames_weighted_all_valid_and_degenerate(gpu):
    rng = np.random.default_rng(41)
    f = 400
    # weighted with zero-weight frames, no NaN (all-valid arithmetic)
    stack = _random_stack(42, f, 6, 11, nan_frac=0.0)
    weights = rng.uniform(0.2, 3.0, f).astype(np.float32)
    weights[::37] = 0.0
    with _stack_of(stack) as s:
        got = s.integrate_hardened_winsorized_sigma(weights, 3.0, 2.5, 3, 0.5)
    _compare(got[:5], oracle.weighted_integrate_stack(stack, weights=weights, rejection="winsorized_sigma",
                                                      high_sigma=2.5))
    # unit weights, all valid; a degenerate IQR (constant pixels) takes the std fallback
    stack = _random_stack(43, f, 4, 9, nan_frac=0.0)
    stack[:, 1, :] = 7.0
    stack[:3, 1, 2] = 500.0
    with _stack_of(stack) as s:
        got = s.integrate_hardened_winsorized_sigma(None, 3.0, 3.0, 3, 0.5)
    _compare(got[:5], oracle.weighted_integrate_stack(stack, rejection="winsorized_sigma"))


def test_winsorized_frame_limit(gpu):
    with _stack_of(np.ones((16385, 1, 2), np.float32)) as s:
        with pytest.raises(ValueError):
            s.integrate_hardened_winsorized_sigma(None, 3.0, 3.0, 3, 0.5)


def test_integrate_mean_matches_oracle(gpu):
    stack = _random_stack(9, 11, 7, 9)
    weights = np.linspace(0.5, 1.5, 11).astype(np.float32)
    with _stack_of(stack) as s:
        master, weight = s.integrate_mean(weights)
    em, ew, *_ = oracle.weighted_integrate_stack(stack, weights=weights)
    assert np.array_equal(master, em)
    assert np.array_equal(weight, ew)


def test_integrate_rejects_bad_arguments(gpu):
    with _stack_of(np.ones((3, 2, 2), np.float32)) as s:
        with pytest.raises(ValueError):
            s.integrate_hardened_winsorized_sigma(np.ones(2, np.float32))
        with pytest.raises(ValueError):
            s.integrate_hardened_winsorized_sigma(None, -1.0)
        with pytest.raises(ValueError):
            s.integrate_hardened_winsorized_sigma(None, 3.0, 3.0, 0)
        with pytest.raises(ValueError):
            s.integrate_hardened_winsorized_sigma(None, 3.0, 3.0, 3, 1.5)
        with pytest.raises(IndexError):
            s.download_frame_tile(0, 0, 0, 3, 2)
        with pytest.raises(IndexError):
            s.upload_calibrated_frame(5, np.ones((2, 2), np.float32))


# ------------------------------------------------------------ stats and DQ
def test_frame_stats_exact_median_mad(gpu):
    rng = np.random.default_rng(8)
    for n_even in (0, 1):
        frame = rng.normal(10, 3, size=(37, 41 + n_even)).astype(np.float32)
        frame[0, :5] = np.nan
        with _stack_of(frame[None]) as s:
            st = s.frame_global_stats(0, exact_median=True)
        ref = oracle.scan_quality_stats(frame)
        assert st["median"] == ref["median"]
        assert st["mad"] == ref["mad"]
        assert st["valid_pixels"] == ref["pixel_count"]
        assert st["mean"] == pytest.approx(ref["mean"], rel=1e-13)
        assert st["std"] == pytest.approx(ref["rms"], rel=1e-9)


def test_end_to_end_config0_and_dq(gpu):
    data = make_set(seed=0)
    pol = {"tile_size": 512}
    mb = P.build_master_bias(data.bias, pol)
    md = P.build_master_dark(data.dark, mb, pol)
    mf, _ = P.build_master_flat(data.flat, mb, pol)
    assert np.array_equal(mb, oracle.build_master_bias(data.bias, pol))
    res = P.run_resident_stack(data.light, [data.light_exposure_s] * len(data.light),
                               {"bias": mb, "dark": md, "flat": mf}, data.matrices,
                               dark_exposure_s=data.dark_exposure_s, interpolation="bilinear")
    cal = [oracle.calibrate_light(l, mb, md, mf, data.light_exposure_s, data.dark_exposure_s) for l in data.light]
    warped, covs = [], []
    for i, c in enumerate(cal):
        if np.array_equal(data.matrices[i], np.eye(3)):
            warped.append(c)
            covs.append(np.ones_like(c))
        else:
            wimg, cov = oracle.warp_matrix(c, data.matrices[i], fill=0.0, interpolation="bilinear")
            warped.append(wimg)
            covs.append(cov)
    exp = oracle.weighted_integrate_stack(np.stack(warped), coverage=np.stack(covs), rejection="winsorized_sigma")
    _compare((res.master, res.weight, res.coverage, res.low, res.high), exp)
    geo = np.sum(covs, axis=0).astype(np.float32)
    assert np.array_equal(res.geometric, geo)
    edq = oracle.resident_dq_map(exp[0], exp[1], exp[2], exp[3], exp[4], geo, active_frame_count=len(cal))
    assert np.array_equal(res.dq, edq)


def test_end_to_end_simple_snr_weights_and_inactive_frame(gpu):
    # The resident simple_snr rule (resident_cuda.py:1948-1955, :12762-12779)
    # from frame_global_stats of the calibrated frames, one frame masked, a
    # bilinear warp; the five maps and DQ come from one integration launch.
    data = make_set(seed=5, n_lights=7, height=40, width=56, n_cal=3, n_stars=10)
    pol = {"tile_size": 512}
    mb = P.build_master_bias(data.bias, pol)
    md = P.build_master_dark(data.dark, mb, pol)
    mf, _ = P.build_master_flat(data.flat, mb, pol)
    active = [True, True, True, False, True, True, True]
    res = P.run_resident_stack(data.light, [data.light_exposure_s] * 7, {"bias": mb, "dark": md, "flat": mf},
                               data.matrices, dark_exposure_s=data.dark_exposure_s, interpolation="bilinear",
                               weighting="simple_snr", active=active)
    cal = [oracle.calibrate_light(l, mb, md, mf, data.light_exposure_s, data.dark_exposure_s) for l in data.light]
    ew = np.asarray([oracle.simple_snr_weight(oracle.frame_global_stats(c)) if a else 0.0
                     for c, a in zip(cal, active)], np.float32)
    assert np.allclose(res.weights, ew, rtol=1e-6, atol=0) and res.weights[3] == 0.0
    warped, covs = [], []
    for i, c in enumerate(cal):
        if np.array_equal(data.matrices[i], np.eye(3)):
            warped.append(c)
            covs.append(np.ones_like(c))
        else:
            wimg, cov = oracle.warp_matrix(c, data.matrices[i], fill=0.0, interpolation="bilinear")
            warped.append(wimg)
            covs.append(cov)
    exp = oracle.weighted_integrate_stack(np.stack(warped), weights=res.weights, coverage=np.stack(covs),
                                          rejection="winsorized_sigma")
    _compare((res.master, res.weight, res.coverage, res.low, res.high), exp)
    edq = oracle.resident_dq_map(exp[0], exp[1], exp[2], exp[3], exp[4], np.sum(covs, axis=0).astype(np.float32),
                                 active_frame_count=6)
    assert np.array_equal(res.dq, edq)


def test_end_to_end_config1_stagewise(gpu):
    # Config 1's recipe at a 256 x 256 crop (32 lights, 16 + 16 + 16 calibration
    # frames, raw u16 ADU, drift +-20 px / +-0.5 deg / scale 1 +- 5e-4, cosmic
    # rays, 2 satellite trails, Lanczos3 with NaN fill, variance-aware weights).
    # Each stage is compared with the oracle on that stage's own inputs, so a
    # last-bit warp difference cannot flip a rejection decision downstream.
    data = make_set(seed=1, n_lights=32, height=256, width=256, n_cal=16, n_stars=25, fwhm=3.0, max_shift=20.0,
                    max_rot_deg=0.5, max_scale=5e-4, cosmic_density=1e-4, trails=2)
    f, h, w = data.light.shape
    adu = to_adu(data.light)
    bias, dark, flat = (to_adu(x).astype(np.float32) for x in (data.bias, data.dark, data.flat))
    pol = {"tile_size": 128}
    mb = P.build_master_bias(bias, pol)
    md = P.build_master_dark(dark, mb, pol)
    mf, _ = P.build_master_flat(flat, mb, pol)
    assert np.array_equal(mb, oracle.build_master_bias(bias, pol))
    assert np.array_equal(md, oracle.build_master_dark(dark, mb, pol))
    assert np.array_equal(mf, oracle.build_master_flat(flat, mb, pol)[0])
    moving = [i for i in range(f) if not np.array_equal(data.matrices[i], np.eye(3))]
    with ResidentCalibratedStack(f, h, w) as s:
        s.set_calibration_masters(mb, md, mf)
        for i in range(f):
            s.calibrate_frame(i, adu[i], data.light_exposure_s, data.dark_exposure_s, None)
        cal = s.download_frames_tile(list(range(f)), 0, 0, w, h)
        rms = [s.frame_global_stats(i)["std"] for i in range(f)]
        s.reset_warp_coverage()
        for _ in range(f - len(moving)):
            s.accumulate_full_warp_coverage_frame()
        s.apply_matrix_frames_cpu_semantics(moving, np.stack([data.matrices[i] for i in moving]), "lanczos3",
                                            fill=np.nan)
        warped = s.download_frames_tile(list(range(f)), 0, 0, w, h)
        weights = P.variance_aware_weights(rms)
        got = s.integrate_hardened_winsorized_sigma(weights, 3.0, 3.0, 3, 0.5)
    # calibration: bit-exact
    for i in range(f):
        exp = oracle.calibrate_light(adu[i], mb, md, mf, data.light_exposure_s, data.dark_exposure_s)
        assert np.array_equal(cal[i], exp), i
    # weights: variance_aware over the frames' FP64 rms
    ref_rms = [oracle.frame_global_stats(c)["std"] for c in cal]
    assert np.allclose(rms, ref_rms, rtol=1e-9)
    exp_w = oracle.quality_weights([{"background_rms": r} for r in ref_rms], "variance_aware")
    assert np.allclose(weights, exp_w, rtol=1e-6)
    # warp: tolerance (f32 table weights vs FP64), non-finite pattern exact
    for i in moving:
        exp, _cov = oracle.warp_matrix(cal[i], data.matrices[i], fill=np.nan, interpolation="lanczos3")
        fin = np.isfinite(exp)
        assert np.array_equal(np.isfinite(warped[i]), fin), i
        err = np.abs(warped[i][fin] - exp[fin])
        scale = np.max(np.abs(exp[fin]))
        assert err.max() <= 3e-6 * scale and np.median(err) <= 1e-6 * np.median(np.abs(exp[fin])), \
            (i, err.max(), scale)
    # integration of the GPU's warped stack
    exp = oracle.weighted_integrate_stack(warped, weights=weights, rejection="winsorized_sigma")
    _compare(got[:5], exp)
    assert got[3].sum() + got[4].sum() > 0  # cosmic rays / trails are rejected


# ------------------------------------------------------------ translations
def test_translation_warps_bit_exact(gpu):
    from oracle.warp import translate_bilinear_f32, translate_integer

    rng = np.random.default_rng(6)
    h, w = 37, 53
    frames = rng.normal(100, 10, (3, h, w)).astype(np.float32)
    frames[1, 5, 7] = np.nan
    with _stack_of(frames) as s:
        s.reset_warp_coverage()
        s.apply_translation_frame(0, 3, -2)
        s.apply_translation_bilinear_frame(1, 1.3, 0.45, fill=-1.0)
        s.apply_translation_bilinear_frame(2, -2.0, 1.0)
        got = s.download_frames_tile([0, 1, 2], 0, 0, w, h)
        cov = s.warp_coverage_map()
        assert s.warp_coverage_frame_count == 3
        flags = [s.frame_has_nonfinite(i) for i in range(3)]
    e0, c0 = translate_integer(frames[0], 3, -2)
    e1, c1 = translate_bilinear_f32(frames[1], 1.3, 0.45, fill=-1.0)
    e2, c2 = translate_bilinear_f32(frames[2], -2.0, 1.0)
    assert np.array_equal(got[0], e0, equal_nan=True)
    assert np.array_equal(got[1], e1, equal_nan=True)
    assert np.array_equal(got[2], e2, equal_nan=True)
    assert np.array_equal(cov, c0 + c1 + c2)
    assert flags == [True, True, True]  # NaN fill / NaN source reach the output


def test_reference_fixture_translation_uses_nan_coverage(gpu):
    # tests/test_cuda_resident_stack.py:1481-1494
    from oracle.warp import translate_integer

    frame = np.arange(20, dtype=np.float32).reshape(4, 5)
    with _stack_of([frame]) as s:
        s.apply_translation_frame(0, 1, -1, np.nan)
        master, weight_map = s.integrate_mean()
        with pytest.raises(TypeError):
            s.apply_translation_frame(0, 0.5, 0)
    expected, _ = translate_integer(frame, 1, -1)
    valid = np.isfinite(expected)
    assert np.allclose(master[valid], expected[valid], rtol=1e-5, atol=1e-5)
    assert np.all(master[~valid] == 0.0)
    assert np.allclose(weight_map, valid.astype(np.float32))
