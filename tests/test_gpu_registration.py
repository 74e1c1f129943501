"""GPU parity of star detection (GPU candidate pass + library FP64
measurement) against the oracle restatement of
src/glass/cpu/star_detect.py:121-168: star lists must match field by field,
bit for bit."""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle.registration as O
from glass_amd import ResidentCalibratedStack
from glass_amd import registration as R

pytestmark = pytest.mark.gpu


def _field(seed, h, w, n_stars, noise=2.0, background=100.0):
    rng = np.random.default_rng(seed)
    yy, xx = np.indices((h, w), dtype=np.float32)
    img = rng.normal(background, noise, size=(h, w)).astype(np.float32)
    for _ in range(n_stars):
        x, y = rng.uniform(3, w - 3), rng.uniform(3, h - 3)
        f, s = rng.uniform(50, 2000), rng.uniform(1.0, 2.2)
        r2 = (xx - x) ** 2 + (yy - y) ** 2
        img += (f * np.exp(-r2 / (2.0 * s * s))).astype(np.float32)
    return img.astype(np.float32)


def _assert_same(got, exp):
    assert len(got) == len(exp)
    for g, e in zip(got, exp):
        for name in ("x", "y", "flux", "peak", "fwhm_px", "eccentricity", "snr", "sharpness", "npix"):
            assert getattr(g, name) == getattr(e, name), (name, getattr(g, name), getattr(e, name))


@pytest.mark.parametrize("seed,h,w,n", [(1, 160, 200, 40), (2, 257, 311, 120), (3, 64, 64, 3)])
def test_detect_stars_bit_exact(gpu, seed, h, w, n):
    img = _field(seed, h, w, n)
    img[17, 23] = np.nan
    img[h // 2, w // 2 + 1] = np.inf
    exp = O.detect_stars(img)
    got = R.detect_stars(img)
    _assert_same(got, exp)


def test_detect_stars_options_and_resident_frame(gpu):
    img = _field(7, 120, 180, 30)
    kw = dict(threshold_sigma=3.0, max_stars=10, window_radius=3, min_separation_px=4.0, origin_x=5, origin_y=-2)
    exp = O.detect_stars(img, **kw)
    with ResidentCalibratedStack(2, 120, 180) as s:
        s.upload_calibrated_frame(1, img)
        got = s.detect_stars(1, **kw)
    _assert_same(got, exp)
    exp = O.detect_stars(img, background=101.0, noise=3.0)
    _assert_same(R.detect_stars(img, background=101.0, noise=3.0), exp)


def test_detect_stars_flat_background_std_fallback(gpu):
    # Constant background: MAD = 0, the noise falls back to np.std in float32
    # (tests/test_cpu_registration.py:23-37 star field).
    h, w = 96, 112
    yy, xx = np.indices((h, w), dtype=np.float32)
    img = np.full((h, w), 10.0, dtype=np.float32)
    for x, y, f in [(12, 17, 100.0), (30, 42, 220.0), (71, 15, 160.0), (88, 63, 180.0), (45, 79, 250.0),
                    (19, 86, 130.0), (101, 33, 145.0), (53, 55, 190.0)]:
        img += f * np.exp(-(((xx - x) ** 2 + (yy - y) ** 2) / (2.0 * 1.4 ** 2)))
    img = img.astype(np.float32)
    exp = O.detect_stars(img)
    got = R.detect_stars(img)
    assert len(exp) == 8
    _assert_same(got, exp)


def test_detect_then_register_recovers_shift(gpu):
    img = _field(11, 200, 240, 60, noise=1.0)
    shifted = np.full_like(img, 100.0)
    shifted[3:, :-5] = img[:-3, 5:]  # moving = reference shifted by (-5, +3)
    ref = R.detect_stars(img)
    mov = R.detect_stars(shifted)
    r = R.estimate_star_transform(ref, mov, "translation", min_inliers=6, tolerance_px=1.0)
    o = O.estimate_star_transform(O.detect_stars(img), O.detect_stars(shifted), "translation", min_inliers=6,
                                  tolerance_px=1.0)
    assert r.status == "ok" and r.inliers >= 20
    assert r.matrix == o.matrix and sorted(r.pairs) == sorted(o.pairs)
    assert math.isclose(r.matrix[0][2], 5.0, abs_tol=0.05) and math.isclose(r.matrix[1][2], -3.0, abs_tol=0.05)


def test_reference_fixture_ncc_translation_then_bilinear_shift(gpu):
    # tests/test_cuda_resident_stack.py:1684-1737
    from glass_amd import ResidentCalibratedStack
    from oracle.registration import ncc_translation_search
    from oracle.warp import translate_bilinear_f32

    image = np.zeros((64, 64), dtype=np.float32)
    yy, xx = np.indices(image.shape, dtype=np.float32)
    for x, y, flux in [(12.25, 14.50, 150.0), (24.00, 20.75, 230.0), (41.50, 18.25, 180.0), (16.75, 42.00, 130.0),
                       (49.25, 45.50, 210.0), (35.50, 33.25, 170.0)]:
        image += flux * np.exp(-(((xx - x) ** 2 + (yy - y) ** 2) / (2.0 * 1.2**2)))
    reference = (image + 10.0 + 0.01 * xx + 0.02 * yy).astype(np.float32)
    moving, _ = translate_bilinear_f32(reference, -2.25, 1.5, fill=0.0)
    with ResidentCalibratedStack(2, 64, 64) as s:
        s.upload_calibrated_frame(0, reference)
        s.upload_calibrated_frame(1, moving)
        coarse = s.estimate_translation_to_reference(0, 1, 4, 4)
        refined = s.estimate_translation_subpixel_to_reference(0, 1, float(coarse["dx"]), float(coarse["dy"]), 3, 0.25)
        s.apply_translation_bilinear_frame(1, refined["dx"], refined["dy"], np.nan)
        aligned, weight_map = s.integrate_mean(np.array([0.0, 1.0], dtype=np.float32))
    valid = weight_map > 0.0
    valid &= (xx >= 8) & (xx < 56) & (yy >= 8) & (yy < 56)
    rms = float(np.sqrt(np.mean((aligned[valid] - reference[valid]) ** 2)))
    assert coarse["model"] == "resident_translation_integer_ncc" and coarse["search_count"] == 81
    assert refined["model"] == "resident_translation_subpixel_ncc" and refined["candidate_count"] == 49
    assert abs(refined["dx"] - 2.25) <= 0.25 and abs(refined["dy"] + 1.5) <= 0.25
    assert refined["score"] > 0.98 and rms < 4.0
    # the same candidates and the same winner as the oracle
    odx, ody, oscore, _ = ncc_translation_search(reference, moving, 4, 4)
    assert (coarse["dx"], coarse["dy"]) == (odx, ody) and coarse["score"] == pytest.approx(oscore, abs=1e-6)
    sdx, sdy, sscore, _ = ncc_translation_search(reference, moving, 3, 3, True, coarse["dx"], coarse["dy"], 0.25)
    assert (refined["dx"], refined["dy"]) == (sdx, sdy) and refined["score"] == pytest.approx(sscore, abs=1e-6)


@pytest.mark.parametrize("stride", [1, 3])
def test_ncc_translation_matches_oracle_with_nan_and_stride(gpu, stride):
    from glass_amd import ResidentCalibratedStack
    from oracle.registration import ncc_translation_search
    from oracle.warp import translate_integer

    rng = np.random.default_rng(stride)
    ref = rng.normal(50, 8, (70, 90)).astype(np.float32)
    ref[10:14, 20:25] = np.nan
    mov, _ = translate_integer(ref, 5, -4, fill=np.nan)
    mov += rng.normal(0, 0.5, mov.shape).astype(np.float32)
    with ResidentCalibratedStack(2, 70, 90) as s:
        s.upload_calibrated_frame(0, ref)
        s.upload_calibrated_frame(1, mov)
        got = s.estimate_translation_to_reference(0, 1, 7, 6, stride)
        sub = s.estimate_translation_subpixel_to_reference(0, 1, -5.0, 4.0, 2, 0.5, stride)
    dx, dy, score, _ = ncc_translation_search(ref, mov, 7, 6, sample_stride=stride)
    assert (got["dx"], got["dy"]) == (dx, dy) == (-5.0, 4.0)
    assert got["score"] == pytest.approx(score, abs=1e-6)
    sdx, sdy, sscore, _ = ncc_translation_search(ref, mov, 2, 2, True, -5.0, 4.0, 0.5, stride)
    assert (sub["dx"], sub["dy"]) == (sdx, sdy) and sub["score"] == pytest.approx(sscore, abs=1e-6)


def _stack_of(*frames):
    s = ResidentCalibratedStack(len(frames), frames[0].shape[0], frames[0].shape[1])
    for i, f in enumerate(frames):
        s.upload_calibrated_frame(i, f)
    return s


def _smooth_pair(h=180, w=230, seed=5):
    rng = np.random.default_rng(seed)
    yy, xx = np.indices((h, w), dtype=np.float64)
    ref = (100 + 20 * np.sin(xx / 9.0) * np.cos(yy / 11.0) + 0.05 * xx + rng.normal(0, 0.5, (h, w)))
    ref = ref.astype(np.float32)
    ref[rng.random((h, w)) < 0.003] = np.nan
    mov = np.full_like(ref, np.nan)
    mov[2:, :-3] = ref[:-2, 3:]      # moving(x - 3, y + 2) = reference(x, y)
    return ref, mov


def test_matrix_translation_candidates_match_the_oracle(gpu):
    ref, mov = _smooth_pair()
    a = 0.002
    seeds = np.stack([np.eye(3), [[np.cos(a), -np.sin(a), 3.5], [np.sin(a), np.cos(a), -1.25], [0, 0, 1]]])
    with _stack_of(ref, mov) as s:
        res = s.refine_matrix_translation_candidates_to_reference(0, 1, seeds, 4.0, 0.5, 0.5, 0.125, 3, 1)
        single = s.refine_matrix_translation_candidates_to_reference(0, 1, np.eye(3), 4.0, 0.5, 0.0, 0.125, 3, 2)
    want, sel = O.refine_matrix_translation_candidates(ref, mov, seeds, 4.0, 0.5, 0.5, 0.125, 3, 1)
    assert res["selected_index"] == sel and res["seed_count"] == 2
    assert res["model"] == "resident_cuda_matrix_metric_translation_multi_seed_refine_grid"
    for got, exp in zip(res["seed_results"], want):
        assert (got["dx_correction"], got["dy_correction"]) == (exp["dx"], exp["dy"])
        assert np.array_equal(np.asarray(got["matrix"], np.float32), exp["matrix"])
        m = got["metrics"]
        assert m["valid_pixels"] == exp["valid_pixels"] and m["sample_stride"] == 1
        for key in ("rms", "mean_abs_diff", "ncc"):
            assert m[key] == pytest.approx(exp[key], rel=1e-9, abs=1e-12), key
        assert got["coarse_candidates"] == 17 * 17 and got["fine_candidates"] == 9 * 9
    # the identity seed recovers the shift exactly: correction (+3, -2)
    best = res["seed_results"][0]
    assert (best["dx_correction"], best["dy_correction"]) == (3.0, -2.0) and best["metrics"]["rms"] == 0.0
    with _stack_of(ref, mov, mov) as s3:
        batch = s3.refine_matrix_translation_candidates_batch_to_reference(0, [1, 2], seeds, 4.0, 0.5, 0.5, 0.125,
                                                                           3, 1)
        # no frames: [] when a seed is given; an empty seed stack is an error
        # (parse_matrix_stack, native_bindings.cpp:1377-1379), one (3, 3) is a stack of one
        assert s3.refine_matrix_translation_candidates_batch_to_reference(0, [], seeds[:1]) == []
        with pytest.raises(ValueError, match="at least one matrix"):
            s3.refine_matrix_translation_candidates_batch_to_reference(0, [], seeds[:0])
        one = s3.refine_matrix_translation_candidates_batch_to_reference(0, [1], seeds[0])
        assert len(one) == 1 and one[0]["moving_index"] == 1
        with pytest.raises(ValueError, match="one seed matrix per moving frame"):
            s3.refine_matrix_translation_candidates_batch_to_reference(0, [1], seeds)
    assert [b["moving_index"] for b in batch] == [1, 2]
    for b, exp in zip(batch, want):
        assert (b["dx_correction"], b["dy_correction"]) == (exp["dx"], exp["dy"])
        assert b["metrics"]["valid_pixels"] == exp["valid_pixels"]
        assert b["metrics"]["rms"] == pytest.approx(exp["rms"], rel=1e-9, abs=1e-12)
        assert b["fine_candidates"] == 81 and b["coarse_candidates"] == 289
    # no fine grid: the coarse best rescored alone at the final stride
    ws, _ = O.refine_matrix_translation_candidates(ref, mov, [np.eye(3)], 4.0, 0.5, 0.0, 0.125, 3, 2)
    assert single["seed_results"][0]["fine_candidates"] == 1
    assert single["metrics"]["valid_pixels"] == ws[0]["valid_pixels"]
    assert single["metrics"]["rms"] == pytest.approx(ws[0]["rms"], rel=1e-9, abs=1e-12)


def test_matrix_translation_candidates_errors(gpu):
    ref, mov = _smooth_pair(20, 24)
    with _stack_of(ref, mov) as s:
        with pytest.raises(ValueError, match="radii must be non-negative"):
            s.refine_matrix_translation_candidates_to_reference(0, 1, np.eye(3), -1.0)
        with pytest.raises(ValueError, match="steps must be positive"):
            s.refine_matrix_translation_candidates_to_reference(0, 1, np.eye(3), 1.0, 0.0)
        with pytest.raises(ValueError, match="strides must be positive"):
            s.refine_matrix_translation_candidates_to_reference(0, 1, np.eye(3), coarse_sample_stride=0)
        with pytest.raises(ValueError, match=r"shape \(N, 3, 3\)"):
            s.refine_matrix_translation_candidates_to_reference(0, 1, np.eye(2))
        with pytest.raises(ValueError, match="invertible"):
            s.refine_matrix_translation_candidates_to_reference(0, 1, np.zeros((3, 3)))


@pytest.mark.parametrize("name", ["plain", "nan_pairs_plateau", "given_background_truncated"])
def test_detect_stars_matches_reference_golden(gpu, name):
    # inputs and outputs of the reference's own glass.cpu.star_detect.detect_stars
    # (tests/golden/make_registration_golden.py): the GPU path must give the
    # same list, field by field, bit for bit
    from test_oracle_registration import GOLDEN, STAR_FIELDS, golden_detection_kwargs

    print(f"golden detection case {name}", flush=True)
    stars = R.detect_stars(GOLDEN[f"det/{name}/image"], **golden_detection_kwargs(name))
    assert len(stars) == len(GOLDEN[f"det/{name}/x"])
    for f in STAR_FIELDS:
        got = np.asarray([np.nan if getattr(s, f) is None else getattr(s, f) for s in stars], np.float64)
        assert np.array_equal(got, GOLDEN[f"det/{name}/{f}"].astype(np.float64), equal_nan=True), f
