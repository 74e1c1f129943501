"""GPU star candidate catalogues (gh_star_candidates through the C ABI)
against oracle/star_candidates.py: the reference's literal cases and random
star fields, bit-exact positions, fluxes and centroids."""

from __future__ import annotations

import numpy as np
import pytest

from glass_amd import ResidentCalibratedStack
from oracle import star_candidates as SCand

pytestmark = pytest.mark.gpu


def _stack(*frames):
    h, w = frames[0].shape
    s = ResidentCalibratedStack(len(frames), h, w)
    for i, f in enumerate(frames):
        s.upload_calibrated_frame(i, np.asarray(f, np.float32))
    return s


def _xyf(res):
    return [(float(x), float(y), float(f)) for x, y, f in zip(res["x"], res["y"], res["flux"])]


def _field(seed, h=150, w=170, stars=60, noise=1.0):
    rng = np.random.default_rng(seed)
    yy, xx = np.indices((h, w), dtype=np.float32)
    img = 20.0 + rng.normal(0, noise, (h, w)).astype(np.float32)
    for _ in range(stars):
        x, y, f = rng.uniform(3, w - 3), rng.uniform(3, h - 3), rng.uniform(30, 300)
        img += (f * np.exp(-((xx - x) ** 2 + (yy - y) ** 2) / (2 * 1.5 ** 2))).astype(np.float32)
    img[rng.random((h, w)) < 0.002] = np.nan
    img[40:44, 50:54] = 400.0   # saturated plateau: ties broken by y, then x
    img[10, 10] = -0.0          # signed zero below the threshold
    return img.astype(np.float32)


def test_reference_literal_cases(gpu):
    close = np.zeros((48, 48), np.float32)
    close[10, 10], close[11, 11], close[30, 30], close[40, 7] = 100, 95, 90, 85
    plateau = np.zeros((48, 48), np.float32)
    plateau[4:6, 4:6] = 100.0
    with _stack(close, plateau) as s:
        r = s.star_top_nms_candidates(0, 10.0, 8, 4, 4.0)
        assert r["count"] == 3 and r["stored_count"] == 3
        assert [(x, y) for x, y, _ in _xyf(r)] == [(10, 10), (30, 30), (7, 40)]
        top = s.star_top_candidates(1, 10.0, 2)
        assert [(x, y) for x, y, _ in _xyf(top)] == [(4, 4), (5, 4)] and top["count"] == 4
        g = s.star_grid_top_nms_candidates(1, 10.0, 1, 1, 4, 4, 0.0)
        assert [(x, y) for x, y, _ in _xyf(g)] == [(4, 4), (5, 4), (4, 5), (5, 5)]
    grid = np.zeros((31, 37), np.float32)
    for x, y, f in [(4, 4, 100), (6, 4, 99), (12, 8, 180), (25, 7, 150), (30, 14, 170), (7, 20, 160),
                    (18, 21, 130), (31, 25, 140), (34, 27, 135)]:
        grid[y, x] = f
    with _stack(grid) as s:
        g = s.star_grid_top_nms_candidates(0, 10.0, 5, 3, 3, 8, 3.0)
        assert g["grid_capacity"] == 45 and g["count"] == 9
        assert _xyf(g) == [(12, 8, 180), (30, 14, 170), (7, 20, 160), (25, 7, 150), (31, 25, 140), (34, 27, 135),
                           (18, 21, 130), (4, 4, 100)]


@pytest.mark.parametrize("seed", [1, 2])
def test_catalogues_match_the_oracle(gpu, seed):
    img = _field(seed)
    with _stack(img) as s:
        allc = s.star_candidates(0, 25.0, 100000)
        top = s.star_top_candidates(0, 25.0, 40)
        nms = s.star_top_nms_candidates(0, 25.0, 80, 30, 5.0)
        grid = s.star_grid_top_nms_candidates(0, 25.0, 4, 3, 2, 20, 3.0)
        det = s.star_grid_top_nms_candidates_deterministic(0, 25.0, 4, 3, 2, 20, 3.0)
        short = s.star_candidates(0, 25.0, 5)
    cands = SCand.local_maxima(img, 25.0)
    assert allc["count"] == len(cands) and _xyf(allc) == cands
    assert short["count"] == len(cands) and _xyf(short) == cands[:5]
    count, ranked = SCand.top_candidates(img, 25.0, 80)
    assert top["count"] == count and _xyf(top) == ranked[:40]
    assert _xyf(nms) == SCand.separation_filter(ranked, 30, 5.0)
    gcount, merged = SCand.grid_top(img, 25.0, 4, 3, 2)
    assert grid["count"] == gcount and grid["cell_selected_count"] == len(merged)
    assert _xyf(grid) == SCand.separation_filter(merged, 20, 3.0)
    assert _xyf(det) == _xyf(grid)


@pytest.mark.parametrize("global_mean", [False, True])
def test_centroids_match_the_oracle(gpu, global_mean):
    img = _field(7)
    with _stack(img) as s:
        r = s.star_top_nms_candidates_centroid(0, 25.0, 60, 25, 4.0, 3, global_mean)
        g = s.star_grid_top_nms_candidates_centroid(0, 25.0, 3, 3, 3, 12, 4.0, 9, global_mean)
        plain = s.star_top_nms_candidates(0, 25.0, 60, 25, 4.0)
    bg = SCand.global_mean_background(img) if global_mean else np.nan
    info = r["centroid_refine"]
    assert info["enabled"] and info["background_mode"] == ("global_mean" if global_mean else "local_median")
    if global_mean:
        assert info["background"] == pytest.approx(float(bg), rel=1e-6)
        bg = np.float32(info["background"])  # FP64 sum order may move the last bit
    for res, radius, src in ((r, 3, _xyf(plain)), (g, 9, None)):
        if src is None:
            _, merged = SCand.grid_top(img, 25.0, 3, 3, 3)
            src = SCand.separation_filter(merged, 12, 4.0)
        want = [SCand.refine_centroid(img, x, y, f, radius, bg) for x, y, f in src]
        assert [(x, y) for x, y, _ in want] == [(x, y) for x, y, _ in _xyf(res)]
        assert [m for _, _, m in want] == [bool(v) for v in res["centroid_refine"]["status"]]
        assert res["centroid_refine"]["refined_count"] == sum(m for _, _, m in want)
        assert [f for _, _, f in src] == list(res["flux"])


def test_candidate_errors_and_empty_frames(gpu):
    flat = np.full((9, 9), 3.0, np.float32)
    with _stack(flat) as s:
        assert s.star_top_candidates(0, 5.0, 4)["count"] == 0
        r = s.star_grid_top_nms_candidates_centroid(0, 5.0, 2, 2, 2, 4, 1.0, 2)
        assert r["count"] == 0 and r["stored_count"] == 0 and r["x"].size == 0
        # a flat frame above the threshold: every interior pixel is a maximum
        r = s.star_top_candidates(0, 1.0, 3)
        assert r["count"] == 49 and [(x, y) for x, y, _ in _xyf(r)] == [(1, 1), (2, 1), (3, 1)]
        with pytest.raises(ValueError, match="max_candidates must be positive"):
            s.star_top_candidates(0, 1.0, 0)
        with pytest.raises(ValueError, match="greater than or equal"):
            s.star_top_nms_candidates(0, 1.0, 2, 3, 1.0)
        with pytest.raises(ValueError, match="non-negative"):
            s.star_top_nms_candidates(0, 1.0, 3, 3, -1.0)
        with pytest.raises(ValueError, match="grid dimensions"):
            s.star_grid_top_nms_candidates(0, 1.0, 0, 2, 2, 4, 1.0)
    with ResidentCalibratedStack(2, 9, 9) as s:
        with pytest.raises(RuntimeError, match="loaded before star detection"):
            s.star_top_candidates(1, 1.0, 3)


def _star_field():
    image = np.zeros((96, 112), np.float32)
    yy, xx = np.indices(image.shape, dtype=np.float32)
    for x, y, flux in [(12, 17, 100.0), (30, 42, 220.0), (71, 15, 160.0), (88, 63, 180.0), (45, 79, 250.0),
                       (19, 86, 130.0), (101, 33, 145.0)]:
        image += flux * np.exp(-(((xx - x) ** 2 + (yy - y) ** 2) / (2.0 * 1.4 ** 2)))
    image += 10.0 + 0.01 * xx + 0.02 * yy
    return image.astype(np.float32)


def _shift(data, dx, dy):
    out = np.zeros_like(data)
    h, w = data.shape
    out[max(0, dy):min(h, h + dy), max(0, dx):min(w, w + dx)] = \
        data[max(0, -dy):min(h, h - dy), max(0, -dx):min(w, w - dx)]
    return out


KEYS = ("dx", "dy", "inliers", "refined_dx", "refined_dy", "mutual_inliers", "rms_px")


@pytest.mark.parametrize("grid", [False, True])
def test_star_translation_reference_case(gpu, grid):
    ref = _star_field()
    mov = _shift(ref, 4, -3)
    kw = {"grid_cols": 4, "grid_rows": 4} if grid else {}
    with _stack(ref, mov) as s:
        r = s.estimate_translation_from_stars_to_reference(0, 1, 30.0, 16, 0.25, 8.0, 8.0, **kw)
    assert r["model"] == "resident_star_catalog_pair_offset_translation"
    assert r["candidate_selection"] == ("grid_brightest_per_cell" if grid else "top_flux")
    assert r["catalog_capacity"] == 16 and (r["dx"], r["dy"]) == (-4.0, 3.0) and r["mutual_inliers"] >= 6
    assert abs(r["refined_dx"] + 4.0) < 1e-5 and abs(r["refined_dy"] - 3.0) < 1e-5
    if grid:
        _, rc = SCand.grid_brightest_catalogue(ref, 30.0, 4, 4)
        _, mc = SCand.grid_brightest_catalogue(mov, 30.0, 4, 4)
    else:
        rc = [(x, y) for x, y, _ in SCand.top_candidates(ref, 30.0, 16)[1]]
        mc = [(x, y) for x, y, _ in SCand.top_candidates(mov, 30.0, 16)[1]]
    want = SCand.catalogue_translation(rc, mc, 0.25, 8.0, 8.0)
    assert tuple(r[k] for k in KEYS) == want


@pytest.mark.parametrize("seed", [3, 4])
def test_star_translation_random_fields(gpu, seed):
    ref = _field(seed, stars=40, noise=0.5)
    mov = _shift(ref, -6, 2)
    with _stack(ref, mov) as s:
        a = s.estimate_translation_from_stars_to_reference(0, 1, 40.0, 48, 0.5)
        b = s.estimate_translation_from_stars_to_reference(0, 1, 40.0, 48, 0.5, 10.0, None, 5.0, -2.0, 3.0)
        c = s.estimate_translation_from_stars_to_reference(0, 1, 40.0, 48, 0.5, grid_cols=5, grid_rows=4)
        d = s.estimate_translation_from_stars_to_reference(0, 1, 40.0, 48, 0.5, 0.5, 0.5)
        with pytest.raises(ValueError, match="both be positive"):
            s.estimate_translation_from_stars_to_reference(0, 1, 40.0, 48, 0.5, grid_cols=3)
        with pytest.raises(RuntimeError, match="found no stars"):
            s.estimate_translation_from_stars_to_reference(0, 1, 1e9, 48, 0.5)
    rc = [(x, y) for x, y, _ in SCand.top_candidates(ref, 40.0, 48)[1]]
    mc = [(x, y) for x, y, _ in SCand.top_candidates(mov, 40.0, 48)[1]]
    assert tuple(a[k] for k in KEYS) == SCand.catalogue_translation(rc, mc, 0.5)
    assert (a["dx"], a["dy"]) == (6.0, -2.0)
    assert tuple(b[k] for k in KEYS) == SCand.catalogue_translation(rc, mc, 0.5, 10.0, 10.0, 5.0, -2.0, 3.0)
    assert b["max_abs_dy"] == 10.0
    _, rg = SCand.grid_brightest_catalogue(ref, 40.0, 5, 4)
    _, mg = SCand.grid_brightest_catalogue(mov, 40.0, 5, 4)
    assert tuple(c[k] for k in KEYS) == SCand.catalogue_translation(rg, mg, 0.5)
    got_d, want_d = tuple(d[k] for k in KEYS), SCand.catalogue_translation(rc, mc, 0.5, 0.5, 0.5)
    assert np.array_equal(np.array(got_d, float), np.array(want_d, float), equal_nan=True)
