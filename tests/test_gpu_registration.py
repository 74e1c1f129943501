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
