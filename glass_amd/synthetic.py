"""Seeded synthetic deep-sky calibration/light sets for tests and benchmarks.

Frames are built in normalised units (bias 0.02, sky 0.05 ...), as
raw = bias + dark_current * exposure + flat * (sky + stars) + noise, and can be
quantised to uint16 ADU (x 65535) and encoded as FITS BITPIX=16/BZERO=32768
big-endian bytes, the raw format the resident path reads.

Each light has a known moving->reference affine matrix M: a star at reference
position p is rendered at M^-1 p in the light, so warping the light through M
puts it back on the reference grid.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class SyntheticSet:
    bias: np.ndarray
    dark: np.ndarray
    flat: np.ndarray
    light: np.ndarray
    matrices: np.ndarray
    light_exposure_s: float
    dark_exposure_s: float
    meta: dict = field(default_factory=dict)


def radial_vignetting(height, width):
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float64)
    cy, cx = (height - 1) / 2.0, (width - 1) / 2.0
    r2 = ((yy - cy) / max(cy, 1.0)) ** 2 + ((xx - cx) / max(cx, 1.0)) ** 2
    return (1.0 - 0.5 * np.clip(r2, 0.0, 2.0)) .astype(np.float64)


def render_stars(height, width, xs, ys, peaks, fwhm, radius=None):
    """Sum of circular Gaussians (FWHM in px) rendered in local windows."""
    sigma = np.broadcast_to(np.asarray(fwhm, dtype=np.float64) / 2.3548200450309493, np.shape(xs))
    out = np.zeros((height, width), dtype=np.float64)
    for x, y, a, s in zip(xs, ys, peaks, sigma):
        r = int(radius if radius is not None else np.ceil(4.0 * s))
        x0, x1 = max(int(np.floor(x)) - r, 0), min(int(np.floor(x)) + r + 2, width)
        y0, y1 = max(int(np.floor(y)) - r, 0), min(int(np.floor(y)) + r + 2, height)
        if x0 >= x1 or y0 >= y1:
            continue
        gy, gx = np.mgrid[y0:y1, x0:x1]
        out[y0:y1, x0:x1] += a * np.exp(-((gx - x) ** 2 + (gy - y) ** 2) / (2.0 * s * s))
    return out


def affine_matrix(tx, ty, rot_deg=0.0, scale=1.0):
    a = np.deg2rad(rot_deg)
    c, s = np.cos(a) * scale, np.sin(a) * scale
    return np.array([[c, -s, tx], [s, c, ty], [0.0, 0.0, 1.0]], dtype=np.float64)


def make_set(
    seed=0,
    n_lights=8,
    height=128,
    width=128,
    n_cal=4,
    n_stars=40,
    fwhm=2.5,
    max_shift=5.0,
    max_rot_deg=0.0,
    max_scale=0.0,
    cosmic_density=0.0,
    trails=0,
    sky=0.05,
    noise=0.01,
    bias_level=0.02,
    read_noise=0.003,
    dark_current=0.001,
    light_exposure_s=60.0,
    dark_exposure_s=60.0,
    identity_first=True,
) -> SyntheticSet:
    """Config-0-style set (defaults) or larger ones by argument."""
    rng = np.random.default_rng(seed)
    h, w = height, width
    bias = (bias_level + rng.normal(0.0, read_noise, size=(n_cal, h, w))).astype(np.float32)
    dark = (bias_level + dark_current * dark_exposure_s + rng.normal(0.0, read_noise, size=(n_cal, h, w)))
    dark = dark.astype(np.float32)
    vig = 0.8 + 0.2 * radial_vignetting(h, w)
    flat = (bias_level + vig[None] * (1.0 + rng.normal(0.0, 0.01, size=(n_cal, h, w)))).astype(np.float32)
    star_x = rng.uniform(0, w - 1, n_stars)
    star_y = rng.uniform(0, h - 1, n_stars)
    star_peak = np.exp(rng.uniform(np.log(0.05), np.log(0.9), n_stars))
    lights = np.empty((n_lights, h, w), dtype=np.float32)
    mats = np.empty((n_lights, 3, 3), dtype=np.float64)
    for i in range(n_lights):
        if i == 0 and identity_first:
            m = np.eye(3)
        else:
            m = affine_matrix(
                rng.uniform(-max_shift, max_shift), rng.uniform(-max_shift, max_shift),
                rng.uniform(-max_rot_deg, max_rot_deg), 1.0 + rng.uniform(-max_scale, max_scale),
            )
        mats[i] = m
        inv = np.linalg.inv(m)
        px = inv[0, 0] * star_x + inv[0, 1] * star_y + inv[0, 2]
        py = inv[1, 0] * star_x + inv[1, 1] * star_y + inv[1, 2]
        signal = sky + render_stars(h, w, px, py, star_peak, fwhm)
        if trails and i < trails:
            yy, xx = np.mgrid[0:h, 0:w]
            ang = rng.uniform(0, np.pi)
            d = np.abs((xx - w / 2) * np.sin(ang) - (yy - h / 2) * np.cos(ang) - rng.uniform(-h / 4, h / 4))
            signal = signal + 0.3 * (d <= 1.5)
        raw = bias_level + dark_current * light_exposure_s + vig * signal + rng.normal(0.0, noise, size=(h, w))
        if cosmic_density > 0:
            hits = rng.random((h, w)) < cosmic_density
            raw = np.where(hits, raw + rng.uniform(0.5, 1.0, size=(h, w)), raw)
        lights[i] = raw.astype(np.float32)
    return SyntheticSet(bias, dark, flat, lights, mats, light_exposure_s, dark_exposure_s,
                        {"seed": seed, "fwhm": fwhm, "stars": n_stars})


def to_adu(frames):
    """Normalised units -> uint16 ADU (x 65535, rounded, clipped)."""
    return np.clip(np.rint(np.asarray(frames, dtype=np.float64) * 65535.0), 0, 65535).astype(np.uint16)


def encode_fits_u16be_bzero(adu):
    """uint16 ADU -> FITS BITPIX=16 / BZERO=32768 big-endian bytes."""
    stored = (np.asarray(adu, dtype=np.uint16).astype(np.uint32) ^ 0x8000).astype(">u2")
    return np.frombuffer(stored.tobytes(), dtype=np.uint8).reshape(np.shape(adu) + (2,))
