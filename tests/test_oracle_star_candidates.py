"""oracle/star_candidates.py against the literal cases of the reference's
tests (tests/test_gpu_registration_search.py:1339-1560) and closed forms of
the window centroid."""

from __future__ import annotations

import numpy as np

from oracle import star_candidates as SCand


def test_close_peaks_are_suppressed():
    image = np.zeros((48, 48), np.float32)
    image[10, 10], image[11, 11], image[30, 30], image[40, 7] = 100, 95, 90, 85
    count, ranked = SCand.top_candidates(image, 10.0, 8)
    assert count == 3  # (11, 11) is below its neighbour (10, 10)
    kept = SCand.separation_filter(ranked, 4, 4.0)
    assert [(x, y) for x, y, _ in kept] == [(10, 10), (30, 30), (7, 40)]


def test_plateau_ties_break_by_y_then_x():
    image = np.zeros((12, 12), np.float32)
    image[4:6, 4:6] = 100.0
    assert [(x, y) for x, y, _ in SCand.top_candidates(image, 10.0, 2)[1]] == [(4, 4), (5, 4)]
    count, merged = SCand.grid_top(image, 10.0, 1, 1, 4)
    assert count == 4
    assert [(x, y) for x, y, _ in SCand.separation_filter(merged, 4, 0.0)] == [(4, 4), (5, 4), (4, 5), (5, 5)]


def test_grid_keeps_spatial_candidates():
    image = np.zeros((64, 64), np.float32)
    image[8, 8], image[10, 10], image[8, 40], image[40, 8], image[40, 40] = 100, 95, 90, 85, 80
    count, merged = SCand.grid_top(image, 10.0, 2, 2, 2)
    kept = SCand.separation_filter(merged, 8, 4.0)
    assert count == 5 and len(kept) == 4
    assert {(x, y) for x, y, _ in kept} == {(8, 8), (40, 8), (8, 40), (40, 40)}


def test_non_power_of_two_grid_by_hand():
    image = np.zeros((31, 37), np.float32)
    for x, y, f in [(4, 4, 100), (6, 4, 99), (12, 8, 180), (25, 7, 150), (30, 14, 170), (7, 20, 160),
                    (18, 21, 130), (31, 25, 140), (34, 27, 135)]:
        image[y, x] = f
    count, merged = SCand.grid_top(image, 10.0, 5, 3, 3)
    kept = SCand.separation_filter(merged, 8, 3.0)
    # no cell holds more than 2 of the 9 peaks; (6, 4) is 2 px from (4, 4);
    # (34, 27) is sqrt(13) px from (31, 25); the 8th kept ends the scan
    assert count == 9
    assert [(x, y, f) for x, y, f in kept] == [(12, 8, 180), (30, 14, 170), (7, 20, 160), (25, 7, 150),
                                               (31, 25, 140), (34, 27, 135), (18, 21, 130), (4, 4, 100)]


def test_per_cell_cut_and_nan_neighbours():
    image = np.zeros((10, 10), np.float32)
    image[2, 2], image[2, 4], image[2, 6] = 50, 60, 70   # one cell of a 1x1 grid
    image[7, 7] = 10.0
    image[6, 6] = np.nan                                  # NaN neighbour: (7, 7) still counts
    image[0, 5] = 500.0                                   # frame edge: never a candidate
    count, merged = SCand.grid_top(image, 5.0, 1, 1, 2)
    assert count == 4 and [(x, y) for x, y, _ in merged] == [(6, 2), (4, 2)]
    assert SCand.local_maxima(image, 10.0) == [(2.0, 2.0, 50.0), (4.0, 2.0, 60.0), (6.0, 2.0, 70.0)]


def test_centroid_closed_forms():
    yy, xx = np.indices((21, 23), dtype=np.float64)
    img = (5.0 + 80.0 * np.exp(-((xx - 11.0) ** 2 + (yy - 9.0) ** 2) / 4.0)).astype(np.float32)
    # a star symmetric about its peak centres exactly at the peak
    x, y, moved = SCand.refine_centroid(img, 11.0, 9.0, float(img[9, 11]), 3)
    assert moved and (x, y) == (11.0, 9.0)
    # two pixels above a zero background: weights 3 and 1 -> x = 10.25
    two = np.zeros((9, 21), np.float32)
    two[4, 10], two[4, 11] = 3.0, 1.0
    x, y, moved = SCand.refine_centroid(two, 10.0, 4.0, 3.0, 2)
    assert moved and (x, y) == (10.25, 4.0)
    # the window median: 25 samples, 23 zeros -> background 0 (same result)
    assert SCand.window_median_f32(two[2:7, 8:13].ravel()) == 0.0
    # an even count averages the middle pair in float32
    assert SCand.window_median_f32(np.array([1, 4, 2, 3], np.float32)) == 2.5
    # a finite override replaces the median; one at the peak leaves no weight
    assert SCand.refine_centroid(two, 10.0, 4.0, 3.0, 2, background=3.0)[2] is False
    assert SCand.refine_centroid(two, 10.0, 4.0, -1e31, 2)[2] is False


def _star_field():
    """tests/test_cuda_resident_stack.py:45-59 (the reference's star field)."""
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


def test_catalogue_translation_reference_star_field():
    ref = _star_field()
    mov = _shift(ref, 4, -3)
    _, rc = SCand.top_candidates(ref, 30.0, 16)
    _, mc = SCand.top_candidates(mov, 30.0, 16)
    dx, dy, inl, rdx, rdy, mutual, rms = SCand.catalogue_translation(
        [(x, y) for x, y, _ in rc], [(x, y) for x, y, _ in mc], 0.25, 8.0, 8.0)
    # the 7 mutual matches all offset by exactly (-4, 3); the kernel's mean is
    # sum * (1 / 7) in float32: -28 * f32(1/7) rounds to -4, 21 * f32(1/7) to
    # the float above 3, and the rms is that last-bit residual
    assert (dx, dy, mutual, inl) == (-4.0, 3.0, 7, 7)
    assert rdx == -4.0 and rdy == float(np.float32(21) * (np.float32(1) / np.float32(7))) and rms < 1e-6
    _, rg = SCand.grid_brightest_catalogue(ref, 30.0, 4, 4)
    _, mg = SCand.grid_brightest_catalogue(mov, 30.0, 4, 4)
    assert len(rg) == len(mg) == 7
    out = SCand.catalogue_translation(rg, mg, 0.25, 8.0, 8.0)
    assert out[:3] == (-4.0, 3.0, 7) and out[5] == 7 and abs(out[4] - 3.0) < 1e-5


def test_catalogue_translation_by_hand():
    ref = [(10.0, 10.0), (30.0, 12.0), (50.0, 40.0), (5.0, 60.0)]
    mov = [(8.0, 11.0), (28.0, 13.0), (48.0, 41.0), (70.0, 70.0)]   # shift (2, -1); one stray star
    dx, dy, inl, rdx, rdy, mutual, rms = SCand.catalogue_translation(ref, mov, 0.5)
    assert (dx, dy, inl, rdx, rdy, mutual, rms) == (2.0, -1.0, 3, 2.0, -1.0, 3, 0.0)
    # |dy| <= 0.5 admits no pair: no offset, no refinement
    dx, dy, inl, rdx, rdy, mutual, rms = SCand.catalogue_translation(ref, mov, 0.5, max_abs_dx=100.0, max_abs_dy=0.5)
    assert np.isnan(dx) and np.isnan(rdx) and np.isnan(rms) and (inl, mutual) == (0, 0)
    # a prior of radius 0.5 around (2, -1) keeps exactly the 3 true pairs
    assert SCand.catalogue_translation(ref, mov, 0.5, prior_dx=2.0, prior_dy=-1.0, prior_radius_px=0.5)[:3] == \
        (2.0, -1.0, 3)
    # a prior far from every offset leaves none
    assert np.isnan(SCand.catalogue_translation(ref, mov, 0.5, prior_dx=-90.0, prior_dy=0.0, prior_radius_px=3.0)[0])
    # a jitter of 0.25 px on one star: the mean of the three mutual offsets
    jit = [(8.25, 11.0)] + mov[1:]
    out = SCand.catalogue_translation(ref, jit, 0.5)
    assert out[5] == 3 and out[3] == np.float32((np.float32(1.75) + np.float32(2.0) + np.float32(2.0))
                                                * (np.float32(1) / np.float32(3)))
