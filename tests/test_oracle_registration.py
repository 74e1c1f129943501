"""Pins the star detection / registration oracle (CPU) against the reference
tests' literal cases (tests/test_cpu_registration.py:64-112,
tests/test_cpu_star_detect.py:30-40) and closed forms, and checks the
library's host-side transform search against it (no GPU needed)."""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle.registration as O

POINTS = np.array([(10, 10), (20, 10), (10, 25), (30, 30), (5, 35), (35, 5), (25, 18), (14, 32)], dtype=np.float64)


def _gaussian_field(h, w, stars, sigma=1.4, background=10.0):
    yy, xx = np.indices((h, w), dtype=np.float32)
    img = np.full((h, w), background, dtype=np.float32)
    for x, y, f in stars:
        img += f * np.exp(-(((xx - x) ** 2 + (yy - y) ** 2) / (2.0 * sigma ** 2)))
    return img.astype(np.float32)


def test_translation_with_outlier_reference_case():
    # tests/test_cpu_registration.py:64-84
    ref = [O.Star(float(x), float(y), 100.0 - 10 * i) for i, (x, y) in enumerate(POINTS)]
    mov = [O.Star(s.x - 3, s.y + 2, s.flux) for s in ref] + [O.Star(100.0, 100.0, 10.0)]
    r = O.estimate_star_transform(ref, mov, "translation", min_inliers=6, tolerance_px=0.5)
    assert r.status == "ok" and r.inliers == 8 and r.rms_px == 0.0
    assert abs(r.matrix[0][2] - 3.0) < 1e-6 and abs(r.matrix[1][2] + 2.0) < 1e-6


def _similarity_case(angle_deg, scale, t):
    a = math.radians(angle_deg)
    lin = scale * np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
    mov = (POINTS - np.asarray(t)) @ np.linalg.inv(lin).T
    ref = [O.Star(float(x), float(y), 1000.0 - i) for i, (x, y) in enumerate(POINTS)]
    movs = [O.Star(float(x), float(y), 1000.0 - i) for i, (x, y) in enumerate(mov)]
    return ref, movs, lin


def test_similarity_reference_case():
    # tests/test_cpu_registration.py:87-111
    ref, mov, lin = _similarity_case(5.0, 1.02, (3.0, -2.0))
    r = O.estimate_star_transform(ref, mov, "similarity", min_inliers=6, tolerance_px=0.5)
    assert r.status == "ok" and r.inliers == 8 and r.rms_px < 1e-5
    m = np.asarray(r.matrix)
    assert np.allclose(m[:2, :2], lin, atol=1e-6) and np.allclose(m[:2, 2], [3.0, -2.0], atol=1e-6)


def test_affine_recovers_exact_map():
    lin = np.array([[0.99, -0.05], [0.04, 1.01]])
    t = np.array([4.5, 1.25])
    mov = (POINTS - t) @ np.linalg.inv(lin).T
    ref = [O.Star(float(x), float(y), 1000.0 - i) for i, (x, y) in enumerate(POINTS)]
    movs = [O.Star(float(x), float(y), 1000.0 - i) for i, (x, y) in enumerate(mov)]
    r = O.estimate_star_transform(ref, movs, "affine", min_inliers=6, tolerance_px=0.5)
    assert r.status == "ok" and r.inliers == 8
    m = np.asarray(r.matrix)
    assert np.allclose(m[:2, :2], lin, atol=1e-9) and np.allclose(m[:2, 2], t, atol=1e-9)


def test_detect_symmetric_star_centroid_and_order():
    img = _gaussian_field(48, 64, [(20, 16, 300.0), (44, 30, 500.0)])
    stars = O.detect_stars(img, threshold_sigma=3.0)
    assert len(stars) == 2
    # brightest first; a star centred on a pixel has its centroid there by symmetry
    assert abs(stars[0].x - 44.0) < 1e-9 and abs(stars[0].y - 30.0) < 1e-9
    assert abs(stars[1].x - 20.0) < 1e-9 and abs(stars[1].y - 16.0) < 1e-9
    # second moments of a sampled Gaussian: fwhm ~ 2.3548 sigma, round star
    assert 0.9 * 2.3548 * 1.4 < stars[0].fwhm_px < 1.05 * 2.3548 * 1.4
    assert stars[0].eccentricity < 1e-6


def test_detect_elongated_reference_case():
    # tests/test_cpu_star_detect.py:30-40
    yy, xx = np.mgrid[0:64, 0:64]
    img = (1000.0 * np.exp(-(((xx - 32.0) / 2.3) ** 2 + ((yy - 32.0) / 1.0) ** 2) / 2.0)).astype(np.float32)
    stars = O.detect_stars(img, threshold_sigma=3.0)
    assert len(stars) == 1 and stars[0].eccentricity > 0.7


def test_background_uses_float32_median_pair_average():
    # np.median of float32 averages the middle pair in float32: f32(a + b) / 2
    a, b = np.float32(1000.1), np.float32(1000.2)
    img = np.array([[a, b, 5.0, 2000.0]], dtype=np.float32)
    med, _rms, _mad, _noise = O.robust_background(img)
    assert med == float(np.float32(a + b) / np.float32(2.0))
    assert med != (float(a) + float(b)) / 2.0


def test_dedupe_keeps_brighter_of_close_pair():
    s = [O.Star(10.0, 10.0, 5.0), O.Star(11.0, 10.5, 9.0), O.Star(30.0, 30.0, 1.0)]
    kept = O.dedupe_stars(s, 2.0)
    assert [k.flux for k in kept] == [9.0, 1.0]


# -------------------------------------------- library host-side transforms
@pytest.mark.parametrize("model", ["translation", "similarity", "affine"])
def test_library_transform_matches_oracle(model):
    from glass_amd import registration as R

    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 400, size=(40, 2))
    a = math.radians(0.7)
    lin = np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]]) * (1.0 if model != "affine" else 1.001)
    if model == "translation":
        lin = np.eye(2)
    t = np.array([12.3, -7.6])
    mov = (pts - t) @ np.linalg.inv(lin).T + rng.normal(0, 0.05, size=pts.shape)
    ref = [O.Star(float(x), float(y), 1000.0 - i) for i, (x, y) in enumerate(pts)]
    movs = [O.Star(float(x), float(y), 1000.0 - i) for i, (x, y) in enumerate(mov)]
    movs += [O.Star(1000.0, 1000.0, 1.0)]
    o = O.estimate_star_transform(ref, movs, model, min_inliers=6, tolerance_px=1.0)
    g = R.estimate_star_transform(ref, movs, model, min_inliers=6, tolerance_px=1.0)
    assert g.status == o.status == "ok"
    assert sorted(g.pairs) == sorted(o.pairs)
    tol = 0.0 if model == "translation" else 1e-9  # lstsq / svd last bits are not pinned
    assert np.allclose(np.asarray(g.matrix), np.asarray(o.matrix), rtol=0, atol=tol)
    assert g.rms_px == pytest.approx(o.rms_px, rel=1e-9)


def test_library_transform_rejects_unknown_model():
    from glass_amd import registration as R

    with pytest.raises(ValueError):
        R.estimate_star_transform([O.Star(1.0, 1.0, 1.0)], [O.Star(1.0, 1.0, 1.0)], "homography")


def test_ncc_translation_search_closed_forms():
    from oracle.registration import ncc_translation_search
    from oracle.warp import translate_bilinear_f32, translate_integer

    rng = np.random.default_rng(5)
    ref = rng.normal(10, 3, (40, 48)).astype(np.float32)
    mov, _ = translate_integer(ref, -3, 2)  # moving = reference shifted by (-3, 2)
    dx, dy, score, scores = ncc_translation_search(ref, mov, 4, 4)
    # the estimate maps moving onto reference: the applied shift undone
    assert (dx, dy) == (3.0, -2.0) and score == pytest.approx(1.0, abs=1e-6)
    # a score is the Pearson coefficient of the overlapping finite pairs
    ys, xs = np.indices(ref.shape)
    ok = (xs - 1 >= 0) & (ys + 2 < 40)
    r = ref[ok]
    m = mov[ys[ok] + 2, xs[ok] - 1]
    fin = np.isfinite(m)
    idx = (-2 + 4) * 9 + (1 + 4)  # candidate (ox=1, oy=-2)
    assert scores[idx] == pytest.approx(np.corrcoef(r[fin], m[fin])[0, 1], rel=1e-9)
    # subpixel grid around the integer estimate, the reference fixture's way
    yy, xx = np.indices((64, 64), dtype=np.float32)
    image = np.zeros((64, 64), np.float32)
    for x, y, flux in [(12.25, 14.5, 150.0), (24.0, 20.75, 230.0), (41.5, 18.25, 180.0), (16.75, 42.0, 130.0),
                       (49.25, 45.5, 210.0), (35.5, 33.25, 170.0)]:
        image += flux * np.exp(-(((xx - x) ** 2 + (yy - y) ** 2) / (2.0 * 1.2 ** 2)))
    reference = (image + 10.0 + 0.01 * xx + 0.02 * yy).astype(np.float32)
    moving, _ = translate_bilinear_f32(reference, -2.25, 1.5, fill=0.0)
    cdx, cdy, _, _ = ncc_translation_search(reference, moving, 4, 4)
    sdx, sdy, sscore, sc = ncc_translation_search(reference, moving, 3, 3, True, cdx, cdy, 0.25)
    assert len(sc) == 49 and sscore > 0.98
    assert abs(sdx - 2.25) <= 0.25 and abs(sdy - (-1.5)) <= 0.25


def test_matrix_candidate_refinement_closed_forms():
    """An integer shift of a smooth frame: the correction that undoes it scores
    rms 0 and ncc 1 and wins the coarse and the fine grid; the float32 offset
    grid of radius 1, step 0.25 has 81 points including the exact shift."""
    from oracle import registration as R

    yy, xx = np.indices((40, 52), dtype=np.float64)
    ref = (100 + 20 * np.sin(xx / 5.0) * np.cos(yy / 7.0) + 0.3 * xx).astype(np.float32)
    mov = np.full_like(ref, np.nan)
    mov[:, 1:] = ref[:, :-1]         # moving(x + 1, y) = reference(x, y)
    mov[1:, :] = mov[:-1, :].copy()  # and one row down: moving(x + 1, y + 1) = reference(x, y)
    offs = R.translation_offsets(0.0, 0.0, 1.0, 0.25)
    assert len(offs) == 81 and (np.float32(-1), np.float32(-1)) in offs
    results, sel = R.refine_matrix_translation_candidates(ref, mov, [np.eye(3)], 1.0, 0.25, 0.25, 0.0625, 2, 1)
    best = results[sel]
    assert (best["dx"], best["dy"]) == (-1.0, -1.0)
    assert best["rms"] == 0.0 and best["mean_abs_diff"] == 0.0 and abs(best["ncc"] - 1.0) < 1e-12
    # the inverse of the winning matrix is the +1 / +1 translation, exactly
    assert np.array_equal(R.invert_matrix_f32(best["matrix"]), np.array([1, 0, 1, 0, 1, 1, 0, 0, 1], np.float32))
    # the sums: every valid pixel pairs reference(x, y) with itself
    s = R.alignment_sums(ref, mov, R.invert_matrix_f32(best["matrix"]), 1)
    assert s[0] == s[1] and s[5] == 0.0 and s[7] == (40 - 1) * (52 - 1)


# ---------------------------- golden vectors of the reference's own functions
# tests/golden/registration_golden.npz holds inputs and outputs of the
# reference's glass.cpu.star_detect.detect_stars and
# glass.cpu.registration.estimate_star_transform /
# estimate_triangle_asterism_transform (written by
# tests/golden/make_registration_golden.py).
from pathlib import Path  # noqa: E402

GOLDEN = np.load(Path(__file__).resolve().parent / "golden" / "registration_golden.npz", allow_pickle=False)
STAR_FIELDS = ("x", "y", "flux", "peak", "fwhm_px", "eccentricity", "snr", "sharpness", "npix")


def golden_detection_kwargs(name):
    p = lambda k: float(GOLDEN[f"det/{name}/{k}"])  # noqa: E731
    kw = {"threshold_sigma": p("threshold_sigma"), "max_stars": int(p("max_stars"))}
    for k in ("background", "noise", "min_separation_px"):
        if math.isfinite(p(k)):
            kw[k] = p(k)
    for k in ("window_radius", "origin_x", "origin_y"):
        if math.isfinite(p(k)):
            kw[k] = int(p(k))
    return kw


def golden_star_lists(name):
    ref = GOLDEN[f"reg/{name}/reference_xy"]
    mov = GOLDEN[f"reg/{name}/moving_xy"]
    rs = [O.Star(float(x), float(y), float(f)) for (x, y), f in zip(ref, np.linspace(1000, 10, len(ref)))]
    ms = [O.Star(float(x), float(y), float(f)) for (x, y), f in zip(mov, np.linspace(900, 9, len(mov)))]
    return rs, ms


def star_arrays(stars):
    return {f: np.asarray([np.nan if getattr(s, f) is None else getattr(s, f) for s in stars], np.float64)
            for f in STAR_FIELDS}


@pytest.mark.parametrize("name", [str(n) for n in GOLDEN["det_names"]])
def test_oracle_detection_matches_reference_golden(name):
    stars = O.detect_stars(GOLDEN[f"det/{name}/image"], **golden_detection_kwargs(name))
    got = star_arrays(stars)
    assert len(stars) == len(GOLDEN[f"det/{name}/x"]) > 0
    for f in STAR_FIELDS:
        assert np.array_equal(got[f], GOLDEN[f"det/{name}/{f}"].astype(np.float64), equal_nan=True), f


@pytest.mark.parametrize("model", ["translation", "similarity", "affine"])
@pytest.mark.parametrize("name", [str(n) for n in GOLDEN["reg_names"]])
def test_oracle_transform_matches_reference_golden(name, model):
    rs, ms = golden_star_lists(name)
    key = f"reg/{name}/star_{model}"
    r = O.estimate_star_transform(rs, ms, model)
    assert r.status == str(GOLDEN[f"{key}/status"])
    assert np.array_equal(np.asarray(r.pairs, np.int64).reshape(-1, 2), GOLDEN[f"{key}/pairs"])
    # bit-exact, lstsq / svd included (same numpy build as the reference run)
    assert np.array_equal(np.asarray(r.matrix), GOLDEN[f"{key}/matrix"])
    assert r.rms_px == float(GOLDEN[f"{key}/rms_px"]) or (math.isnan(r.rms_px) and
                                                          math.isnan(float(GOLDEN[f"{key}/rms_px"])))


@pytest.mark.parametrize("model", ["translation", "similarity", "affine"])
@pytest.mark.parametrize("name", [str(n) for n in GOLDEN["reg_names"]])
def test_library_transform_matches_reference_golden(name, model):
    from glass_amd import registration as R

    rs, ms = golden_star_lists(name)
    key = f"reg/{name}/star_{model}"
    g = R.estimate_star_transform(rs, ms, model)
    assert g.status == str(GOLDEN[f"{key}/status"])
    assert np.array_equal(np.asarray(g.pairs, np.int64).reshape(-1, 2), GOLDEN[f"{key}/pairs"])
    # the library solves the fits with its own FP64 QR / closed forms: the
    # matrices agree with LAPACK's to 1e-12 absolute (pixel units; 1e-8 px
    # over a 10^4 px frame), the translation model exactly
    tol = 0.0 if model == "translation" else 1e-12
    assert np.max(np.abs(np.asarray(g.matrix) - GOLDEN[f"{key}/matrix"])) <= tol
    assert g.rms_px == pytest.approx(float(GOLDEN[f"{key}/rms_px"]), rel=1e-12, abs=1e-12)
