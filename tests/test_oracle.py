"""Pins the CPU oracle (oracle/) against the reference's own golden data and
known answers, per-pixel brute-force restatements of the reference's test
oracles, and closed forms.  CPU only."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle.integration import stack_engine_combine
from oracle.rejection import _percentile_valid, center_and_scale, uses_all_valid_variant


# --------------------------------------------------------------- golden 314
def test_golden_masters_bit_exact(golden_314):
    g = golden_314
    pol = dict(g["policy"], tile_size=int(g["tile_size"]))
    mb = oracle.build_master_bias(g["raw_bias"], pol)
    md = oracle.build_master_dark(g["raw_dark"], mb, pol)
    mf, scalars = oracle.build_master_flat(g["raw_flat"], mb, pol)
    assert np.array_equal(mb, g["master_bias"])
    assert np.array_equal(md, g["master_dark"])
    assert np.array_equal(mf, g["master_flat"])
    assert len(scalars) == 3


def test_golden_calibrated_lights_bit_exact(golden_314):
    g = golden_314
    for i in range(3):
        out = oracle.calibrate_light(
            g["raw_light"][i], g["master_bias"], g["master_dark"], g["master_flat"],
            g["light_exposure_s"][i], g["dark_exposure_s"][i],
        )
        assert np.array_equal(out, g["calibrated_light"][i])


def test_golden_identity_warp_and_single_frame_integration(golden_314):
    g = golden_314
    warped, cov = oracle.warp_matrix(g["calibrated_light"][0], np.eye(3), interpolation="bilinear")
    assert np.array_equal(warped, g["registered_F000010"])
    assert np.array_equal(cov, g["coverage_F000010"])
    master, weight, coverage, low, high = oracle.weighted_integrate_stack(warped[None], coverage=cov[None])
    assert np.array_equal(master, g["integration_master_H"])
    assert np.array_equal(weight, g["integration_weight_map_H"])
    assert np.array_equal(coverage, g["integration_coverage_map_H"])
    assert np.array_equal(low, g["integration_low_rejection_H"])
    assert np.array_equal(high, g["integration_high_rejection_H"])


# ------------------------------------------------------- calibration literals
def test_calibration_known_answers():
    # tests/test_cpu_calibration.py:11-38
    light = np.full((4, 4), 1200.0, np.float32)
    bias = np.full((4, 4), 100.0, np.float32)
    flat = np.full((4, 4), 2.0, np.float32)
    out = oracle.calibrate_light(light, bias, np.full((4, 4), 120.0, np.float32), flat, 60.0, 60.0,
                                 master_dark_includes_bias=True, dark_scaling_enabled=False)
    assert np.allclose(out, 540.0)
    out = oracle.calibrate_light(light, bias, np.full((4, 4), 20.0, np.float32), flat, 60.0, 60.0,
                                 master_dark_includes_bias=False, dark_scaling_enabled=False)
    assert np.allclose(out, 540.0)
    out = oracle.calibrate_light(np.full((2, 2), 20.0, np.float32), None, None, np.full((2, 2), 2.0, np.float32),
                                 1.0, None, pedestal=5.0)
    assert np.allclose(out, 15.0)


def test_fits_u16be_decode_round_trip():
    from oracle.calibration import encode_fits_u16be_bzero

    adu = np.array([0, 1, 255, 256, 32767, 32768, 65534, 65535], dtype=np.uint16)
    raw = encode_fits_u16be_bzero(adu)
    # stored value is the signed int16 ADU - 32768, big-endian
    assert raw[:2].tolist() == [0x80, 0x00]
    assert np.array_equal(oracle.decode_fits_u16be_bzero(raw), adu.astype(np.float32))


# ------------------------------------------------------- integration literals
def test_weighted_integrate_with_coverage_literal():
    # tests/test_cpu_integration.py:26-42
    stack = np.stack([np.ones((3, 3), np.float32), np.ones((3, 3), np.float32) * 3])
    coverage = np.ones_like(stack)
    coverage[1, 0, 0] = 0
    master, weight, cov, low, high = oracle.weighted_integrate_stack(stack, coverage=coverage, weights=np.array([1, 3]))
    assert np.allclose(master[1:, 1:], 2.5)
    assert master[0, 0] == 1.0 and weight[0, 0] == 1.0 and cov[0, 0] == 1.0
    assert low.sum() == 0 and high.sum() == 0


def test_winsorized_rejects_outlier_literal():
    # tests/test_cpu_integration.py:45-52
    stack = np.stack([np.ones((4, 4), np.float32)] * 4 + [np.ones((4, 4), np.float32) * 9])
    master, weight, cov, low, high = oracle.weighted_integrate_stack(stack, rejection="winsorized_sigma", high_sigma=1.0)
    assert np.allclose(master, 1) and np.all(weight == 4) and np.all(cov == 4)
    assert low.sum() == 0 and high.sum() > 0


def test_winsorized_low_sample_literal_and_guard():
    # tests/test_cpu_integration.py:55-96
    stack = np.stack([np.ones((2, 2), np.float32)] * 3 + [np.ones((2, 2), np.float32) * 12])
    master, weight, cov, low, high = oracle.weighted_integrate_stack(
        stack, rejection="winsorized_sigma", low_sigma=2.4, high_sigma=2.4)
    assert np.allclose(master, 1.0) and np.all(weight == 3) and np.all(cov == 3)
    assert low.sum() == 0 and high.sum() == 4
    master, weight, cov, low, high = oracle.weighted_integrate_stack(
        stack, rejection="winsorized_sigma", low_sigma=2.4, high_sigma=2.4, max_reject_fraction=0.05)
    assert np.allclose(master, 3.75) and np.all(weight == 4) and np.all(cov == 4)
    assert low.sum() == 0 and high.sum() == 0


def test_resident_hardened_literals():
    # tests/test_cuda_resident_stack.py:2878-2932 and :2934-2991
    frames = np.stack([
        np.array([[1, 1], [10, 4]], np.float32),
        np.array([[1, 2], [10, 4]], np.float32),
        np.array([[1, 3], [10, 4]], np.float32),
        np.array([[12, 4], [10, 40]], np.float32),
    ])
    _, _, _, low, high = oracle.weighted_integrate_stack(
        frames, weights=np.array([1.0, 2.0, 1.0, 1.0], np.float32), rejection="winsorized_sigma",
        low_sigma=2.4, high_sigma=2.4)
    assert high[0, 0] == 1 and high[1, 1] == 1
    frames = np.stack([np.full((2, 3), 10, np.float32)] * 4 + [np.array([[100, 10, 10], [10, -50, 10]], np.float32)])
    _, _, _, low, high = oracle.weighted_integrate_stack(
        frames, weights=np.ones(5, np.float32), rejection="winsorized_sigma", low_sigma=2.0, high_sigma=2.0,
        min_samples=1, max_reject_fraction=1.0)
    assert high[0, 0] == 1 and low[1, 1] == 1


def test_stack_engine_outlier_literal():
    # tests/test_stack_engine.py:206-227 (tile size 2)
    frames = np.stack([np.full((4, 4), 10.0, np.float32)] * 19 + [np.full((4, 4), 500.0, np.float32)])
    master, weight, cov, low, high = stack_engine_combine(frames, tile_size=2)
    assert np.allclose(master, 10.0) and np.all(cov == 19) and high.sum() == 16


# ------------------------------------------- brute-force per-pixel restatement
def _linear_percentile(values, fraction):
    """tests/test_stack_engine.py:124-135, per pixel."""
    if values.size == 1:
        return np.float32(values[0])
    position = np.float32(values.size - 1) * np.float32(fraction)
    lower = int(np.floor(position))
    upper = min(lower + 1, values.size - 1)
    t = np.float32(position - np.float32(lower))
    lo = np.float32(values[lower])
    if upper == lower or t <= 0:
        return lo
    return np.float32(lo * (np.float32(1.0) - t) + np.float32(values[upper]) * t)


def _winsorized_brute_force(stack, valid, low_sigma, high_sigma):
    """tests/test_stack_engine.py:71-121: one pixel at a time."""
    data = np.asarray(stack, np.float32)
    ok = np.asarray(valid, bool) & np.isfinite(data)
    center = np.zeros(data.shape[1:], np.float32)
    scale = np.zeros(data.shape[1:], np.float32)
    for y in range(data.shape[1]):
        for x in range(data.shape[2]):
            values = np.sort(data[:, y, x][ok[:, y, x]]).astype(np.float32)
            if values.size == 0:
                continue
            c0 = _linear_percentile(values, 0.5)
            q25 = _linear_percentile(values, 0.25)
            q75 = _linear_percentile(values, 0.75)
            mean = np.float32(np.sum(values.astype(np.float64)) / values.size)
            fallback = np.float32(np.sqrt(np.sum((values.astype(np.float64) - np.float64(mean)) ** 2) / values.size))
            s0 = np.float32((q75 - q25) / np.float32(1.349))
            if not (s0 > 0 and np.isfinite(s0)):
                s0 = fallback
            lo = np.float32(c0 - np.float32(low_sigma) * s0)
            hi = np.float32(c0 + np.float32(high_sigma) * s0)
            w = np.clip(values, lo, hi).astype(np.float32)
            wm = np.float32(np.sum(w.astype(np.float64)) / values.size)
            ws = np.float32(np.sqrt(np.sum((w.astype(np.float64) - np.float64(wm)) ** 2) / values.size))
            center[y, x] = wm
            scale[y, x] = ws if ws > 0 and np.isfinite(ws) else s0
    return center, scale


def test_center_and_scale_all_valid_matches_brute_force_seed_594():
    rng = np.random.default_rng(594)
    stack = rng.normal(loc=100.0, scale=3.0, size=(20, 6, 7)).astype(np.float32)
    stack[-1, 2, 3] = np.float32(500.0)
    valid = np.ones_like(stack, bool)
    assert uses_all_valid_variant(stack, valid)
    c, s = center_and_scale(stack, valid, method="winsorized_sigma", low_sigma=3.0, high_sigma=3.0)
    ec, es = _winsorized_brute_force(stack, valid, 3.0, 3.0)
    assert np.allclose(c, ec, rtol=1e-6, atol=1e-5)
    assert np.allclose(s, es, rtol=1e-6, atol=1e-5)


def test_center_and_scale_mixed_valid_matches_brute_force_seed_626():
    rng = np.random.default_rng(626)
    stack = rng.normal(loc=1000.0, scale=4.0, size=(65, 5, 6)).astype(np.float32)
    stack[0, 2, :] -= np.float32(90.0)
    stack[1, :, 3] += np.float32(110.0)
    stack[8, 2, 3] = np.nan
    valid = np.ones_like(stack, bool)
    valid[4, 1, 2] = False
    valid[12, 4, 5] = False
    c, s = center_and_scale(stack, valid, method="winsorized_sigma", low_sigma=2.4, high_sigma=2.4)
    ec, es = _winsorized_brute_force(stack, valid, 2.4, 2.4)
    # the valid-count arithmetic is the same as the brute force: bit-exact
    assert np.array_equal(c, ec)
    assert np.array_equal(s, es)


def test_valid_percentile_matches_numpy_linear_percentile():
    rng = np.random.default_rng(7)
    for count in (1, 2, 3, 4, 5, 8, 13, 64, 193):
        values = np.sort(rng.normal(size=count).astype(np.float32))
        padded = np.concatenate([values, np.full(7, np.inf, np.float32)])[:, None, None]
        for fraction in (0.25, 0.5, 0.75):
            got = _percentile_valid(padded, np.array([[count]]), fraction)[0, 0]
            expected = np.percentile(values.astype(np.float64), fraction * 100.0, method="linear")
            assert abs(float(got) - expected) <= 4 * np.finfo(np.float32).eps * max(1.0, abs(expected))


def test_winsorized_closed_forms():
    # constant stack: nothing rejected, master = the constant
    stack = np.full((9, 3, 3), 7.25, np.float32)
    m, w, c, lo, hi = oracle.weighted_integrate_stack(stack, rejection="winsorized_sigma")
    assert np.all(m == 7.25) and np.all(c == 9) and lo.sum() == 0 and hi.sum() == 0
    # uniform samples on [99, 101] never leave mean +- 3 sigma (sigma = 1/sqrt(3)),
    # so a single far outlier is the only rejection
    rng = np.random.default_rng(11)
    stack = rng.uniform(99.0, 101.0, size=(30, 4, 4)).astype(np.float32)
    stack[5, 1, 2] = 1.0e4
    m, w, c, lo, hi = oracle.weighted_integrate_stack(stack, rejection="winsorized_sigma")
    assert hi[1, 2] == 1 and hi.sum() == 1 and lo.sum() == 0
    keep = np.delete(stack[:, 1, 2], 5).astype(np.float64)
    assert m[1, 2] == np.float32(keep.sum() / keep.size)


# ----------------------------------------------------------------------- warp
def _lanczos_brute_cpu(image, matrix):
    """engine/warp.py:121-152 one pixel at a time (Python scalars, FP64)."""
    img = np.asarray(image, np.float32)
    h, w = img.shape
    inv = np.linalg.inv(np.asarray(matrix, np.float64))
    out = np.zeros((h, w), np.float32)
    cov = np.zeros((h, w), np.float32)

    def sinc(v):
        if abs(v) < 1.0e-8:
            return 1.0
        return float(np.sin(np.pi * v) / (np.pi * v))

    def lw(v):
        v = abs(float(v))
        return 0.0 if v >= 3.0 else sinc(v) * sinc(v / 3.0)

    for y in range(h):
        for x in range(w):
            sx = float((inv[0, 0] * x + inv[0, 1] * y) + inv[0, 2])
            sy = float((inv[1, 0] * x + inv[1, 1] * y) + inv[1, 2])
            if not (2.0 <= sx < w - 3 and 2.0 <= sy < h - 3):
                continue
            bx, by = int(np.floor(sx)), int(np.floor(sy))
            acc = wsum = 0.0
            for yy in range(by - 2, by + 4):
                wy = lw(sy - yy)
                if wy == 0.0:
                    continue
                for xx in range(bx - 2, bx + 4):
                    wx = lw(sx - xx)
                    if wx == 0.0:
                        continue
                    acc += float(img[yy, xx]) * (wx * wy)
                    wsum += wx * wy
            out[y, x] = np.float32(acc / wsum) if abs(wsum) > 1e-12 else 0.0
            cov[y, x] = 1.0
    return out, cov


def _bilinear_brute(image, matrix, fill):
    """tests/test_gpu_warp_vs_cpu.py:35-64."""
    img = np.asarray(image, np.float32)
    h, w = img.shape
    out = np.full((h, w), fill, np.float32)
    cov = np.zeros((h, w), np.float32)
    inv = np.linalg.inv(np.asarray(matrix, np.float64))
    for y in range(h):
        for x in range(w):
            s = inv @ np.asarray([x, y, 1.0])
            sx, sy = float(s[0] / s[2]), float(s[1] / s[2])
            if sx < 0 or sx > w - 1 or sy < 0 or sy > h - 1:
                continue
            x0, y0 = int(np.floor(sx)), int(np.floor(sy))
            x1, y1 = min(x0 + 1, w - 1), min(y0 + 1, h - 1)
            tx, ty = np.float32(sx - x0), np.float32(sy - y0)
            top = img[y0, x0] * (1.0 - tx) + img[y0, x1] * tx
            bot = img[y1, x0] * (1.0 - tx) + img[y1, x1] * tx
            out[y, x] = top * (1.0 - ty) + bot * ty
            cov[y, x] = 1.0
    return out, cov


def _rot(deg, tx, ty):
    a = np.deg2rad(deg)
    return np.array([[np.cos(a), -np.sin(a), tx], [np.sin(a), np.cos(a), ty], [0, 0, 1]], np.float32)


def test_warp_lanczos3_matches_per_pixel_brute_force():
    yy, xx = np.indices((12, 13), dtype=np.float32)
    data = 0.2 * xx + 0.4 * yy
    data[:, 7:] += 5.0
    m = _rot(1.0, 0.35, -0.25)
    got, cov = oracle.warp_matrix(data, m, interpolation="lanczos3")
    exp, ecov = _lanczos_brute_cpu(data, m)
    assert np.array_equal(cov, ecov)
    assert np.allclose(got, exp, rtol=0, atol=2e-6)


def test_warp_bilinear_matches_per_pixel_brute_force():
    data = np.arange(49, dtype=np.float32).reshape(7, 7)
    m = _rot(3.0, 1.2, -0.4)
    got, cov = oracle.warp_matrix_resident(data, m, fill=-1.0, interpolation="bilinear")
    exp, ecov = _bilinear_brute(data, m, -1.0)
    assert np.array_equal(cov, ecov)
    assert np.allclose(got, exp, atol=1e-6)


def test_warp_closed_forms():
    rng = np.random.default_rng(3)
    img = rng.normal(size=(20, 24)).astype(np.float32)
    shift = np.array([[1, 0, 3], [0, 1, -2], [0, 0, 1]], np.float64)
    for interp in ("bilinear", "lanczos3"):
        out, cov = oracle.warp_matrix(img, shift, interpolation=interp)
        ys, xs = np.nonzero(cov)
        # integer translation reproduces the source (Lanczos taps at integers are delta)
        assert np.allclose(out[ys, xs], img[ys + 2, xs - 3], atol=1e-6)
    const = np.full((16, 16), 3.5, np.float32)
    out, cov = oracle.warp_matrix(const, _rot(2.0, 0.3, 0.7), interpolation="lanczos3")
    assert np.allclose(out[cov > 0], 3.5, atol=1e-6)
    yy, xx = np.indices((16, 16), dtype=np.float32)
    ramp = 0.5 * xx - 0.25 * yy + 1.0
    out, cov = oracle.warp_matrix(ramp, _rot(0.0, 0.3, 0.6), interpolation="bilinear")
    ys, xs = np.nonzero(cov)
    expected = 0.5 * (xs - 0.3) - 0.25 * (ys - 0.6) + 1.0
    assert np.allclose(out[ys, xs], expected, atol=1e-5)


# -------------------------------------------------------- quality / weights
def test_quality_stats_closed_forms():
    rng = np.random.default_rng(5)
    for n in (1001, 1000):
        frame = rng.normal(3.0, 2.0, size=n).astype(np.float32)
        frame[7] = np.nan
        st = oracle.scan_quality_stats(frame)
        finite = frame[np.isfinite(frame)]
        assert st["median"] == pytest.approx(float(np.median(finite.astype(np.float64))), abs=0)
        dev = np.abs(finite - np.float32(st["median"]))
        assert st["mad"] == pytest.approx(float(np.median(dev.astype(np.float64))), abs=0)
        assert st["pixel_count"] == finite.size


def test_quality_weights_normalised():
    # engine/integration.py:80-106 on the tests/test_cpu_integration.py frame set
    q = [
        {"snr": 20.0, "quality_score": 10.0, "weight": 10.0},
        {"snr": 10.0, "quality_score": 5.0, "weight": 5.0},
        {"snr": 5.0, "quality_score": 1.0, "noise_sigma": 4.0, "background_rms": 4.0},
        {"snr": 5.0, "quality_score": 1.0, "noise_sigma": 2.0, "background_rms": 2.0},
    ]
    w = oracle.quality_weights(q, "simple_snr")
    assert np.allclose(w, [20 / 7.5, 10 / 7.5, 5 / 7.5, 5 / 7.5])
    w = oracle.quality_weights(q, "variance_aware")
    raw = np.array([10.0, 5.0, 1 / 16, 1 / 4])
    assert np.allclose(w, raw / np.median(raw.astype(np.float32)))


# ----------------------------------------------------------------------- DQ
def test_resident_dq_rules():
    master = np.array([[1.0, np.nan, 1.0, 1.0, 1.0]], np.float32)
    weight = np.array([[3.0, 3.0, 0.0, 3.0, 3.0]], np.float32)
    coverage = np.array([[3.0, 3.0, 3.0, 0.0, 3.0]], np.float32)
    low = np.array([[1.0, 0, 0, 0, 0]], np.float32)
    high = np.array([[0, 0, 0, 0, 2.0]], np.float32)
    geo = np.array([[3.0, 3.0, 3.0, 3.0, 2.0]], np.float32)
    dq = oracle.resident_dq_map(master, weight, coverage, low, high, geo, active_frame_count=3)
    assert dq[0, 0] == oracle.DQ_LOW_REJECTED
    assert dq[0, 1] == oracle.DQ_NO_DATA
    assert dq[0, 2] == oracle.DQ_NO_DATA
    assert dq[0, 3] == oracle.DQ_NO_DATA | oracle.DQ_WARP_EDGE
    assert dq[0, 4] == oracle.DQ_WARP_EDGE | oracle.DQ_HIGH_REJECTED


def test_translation_oracle_closed_forms():
    from oracle.warp import translate_bilinear_f32, translate_integer

    img = np.arange(20, dtype=np.float32).reshape(4, 5)
    out, cov = translate_integer(img, 1, -1)
    # out[y, x] = img[y + 1, x - 1]: column 0 and the last row are outside
    assert np.array_equal(out[:3, 1:], img[1:, :4])
    assert np.all(np.isnan(out[:, 0])) and np.all(np.isnan(out[3, :]))
    assert cov.sum() == 12
    # whole-pixel bilinear equals the integer shift on finite data
    b, bcov = translate_bilinear_f32(img, 1.0, -1.0)
    assert np.array_equal(b, out, equal_nan=True) and np.array_equal(bcov, cov)
    # a linear ramp is reproduced exactly at half-pixel offsets (exact in f32)
    yy, xx = np.indices((6, 9), dtype=np.float32)
    ramp = 4 * xx + 2 * yy + 1
    r, rcov = translate_bilinear_f32(ramp, 0.5, -1.5)
    ok = rcov > 0
    exp = 4 * (xx - 0.5) + 2 * (yy + 1.5) + 1
    assert np.array_equal(r[ok], exp[ok])
    # the sample position is computed in float32: far from the origin the
    # fraction is the float32 one, not the exact one
    wide = np.zeros((1, 20000), np.float32)
    wide[0, 16385] = 1.0
    w32, _ = translate_bilinear_f32(wide, 0.3, 0.0)
    sx = np.float32(16385) - np.float32(0.3)
    # output 16385 samples between 16384 (0) and 16385 (1): the value is the
    # float32 fraction, 0.69921875 there, not 0.7
    assert w32[0, 16385] == sx - np.float32(16384) == np.float32(0.69921875)


# ------------------------------------------- resident Lanczos3 clamp branch
# Closed forms of the clamp (tests/test_gpu_warp_vs_cpu.py:116-124): the value
# is limited to [min - t*range, max + t*range] of the finite taps.  A step
# edge shifted by half a pixel rings past both plateaus (about 11%), so the
# overshooting pixels must land exactly on the bounds; values inside the
# bounds, and every value of a smooth ramp, must be left unchanged.
def _step_edge():
    yy, xx = np.indices((16, 20), dtype=np.float32)
    return np.where(xx >= 10, 100.0, 0.0).astype(np.float32), np.array([[1, 0, 0.5], [0, 1, 0.0], [0, 0, 1.0]])


@pytest.mark.parametrize("t", [0.0, 0.01, 0.05])
def test_resident_clamp_step_edge_closed_form(t):
    img, m = _step_edge()
    raw, cov = oracle.warp_matrix_resident(img, m, np.nan, "lanczos3", -1.0)
    got, cov2 = oracle.warp_matrix_resident(img, m, np.nan, "lanczos3", t)
    assert np.array_equal(cov, cov2)
    ok = cov > 0
    hi, lo = np.float32(100.0 + 100.0 * t), np.float32(-100.0 * t)
    over, under = ok & (raw > hi), ok & (raw < lo)
    assert over.sum() >= 10 and under.sum() >= 10           # the edge rings past both bounds
    assert np.all(got[over] == hi) and np.all(got[under] == lo)
    inside = ok & ~over & ~under
    assert np.array_equal(got[inside], raw[inside])
    assert np.nanmax(got) == hi and np.nanmin(got) == lo


def test_resident_clamp_leaves_smooth_ramp_and_loose_bounds_unchanged():
    yy, xx = np.indices((14, 17), dtype=np.float32)
    ramp = (3.0 * xx + 0.5 * yy).astype(np.float32)
    m = np.array([[np.cos(0.01), -np.sin(0.01), 0.37], [np.sin(0.01), np.cos(0.01), -0.21], [0, 0, 1.0]])
    raw, cov = oracle.warp_matrix_resident(ramp, m, np.nan, "lanczos3", -1.0)
    for t in (0.0, 0.2):
        got, _ = oracle.warp_matrix_resident(ramp, m, np.nan, "lanczos3", t)
        assert np.array_equal(got, raw, equal_nan=True)
    img, m = _step_edge()
    raw, _ = oracle.warp_matrix_resident(img, m, np.nan, "lanczos3", -1.0)
    got, _ = oracle.warp_matrix_resident(img, m, np.nan, "lanczos3", 1.0e6)
    assert np.array_equal(got, raw, equal_nan=True)


def test_resident_clamp_bounds_ignore_nonfinite_taps():
    img, m = _step_edge()
    img = img.copy()
    img[5:11, 3] = np.nan     # NaN taps inside the low plateau's windows
    img[5:11, 15] = -np.inf   # and -inf inside the high plateau's
    got, cov = oracle.warp_matrix_resident(img, m, np.nan, "lanczos3", 0.0)
    ok = cov > 0
    assert np.all(got[ok] >= 0.0) and np.all(got[ok] <= 100.0)
    assert np.all(np.isfinite(got[ok]))
