"""Seeded synthetic raw stacks generated directly in HBM (torch on the GPU).

Same model as glass_amd.synthetic (normalised units, raw = bias + dark current
* t + vignetting * (sky + stars) + noise), quantised to uint16 ADU (x 65535)
and stored as FITS BITPIX=16 / BZERO=32768 big-endian samples, frame-major,
so a 200 x 9600 x 6422 light set (24.7 GB) never touches the host.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from .synthetic import affine_matrix


def _encode_be(raw: torch.Tensor) -> torch.Tensor:
    """Normalised float frame -> int16 tensor holding big-endian FITS bytes."""
    adu = torch.clamp(torch.round(raw * 65535.0), 0, 65535).to(torch.int32)
    stored = adu ^ 0x8000
    swapped = ((stored & 0xFF) << 8) | ((stored >> 8) & 0xFF)
    return swapped.to(torch.int16)


def vignetting(height, width, device):
    yy = torch.arange(height, device=device, dtype=torch.float32)[:, None]
    xx = torch.arange(width, device=device, dtype=torch.float32)[None, :]
    cy, cx = (height - 1) / 2.0, (width - 1) / 2.0
    r2 = ((yy - cy) / max(cy, 1.0)) ** 2 + ((xx - cx) / max(cx, 1.0)) ** 2
    return 0.8 + 0.2 * (1.0 - 0.5 * torch.clamp(r2, 0.0, 2.0))


def render_stars_(frame, xs, ys, peaks, sigmas, radius):
    """Add circular Gaussians to `frame` (H, W) in place via scatter-add."""
    h, w = frame.shape
    off = torch.arange(-radius, radius + 1, device=frame.device)
    cx = torch.floor(xs).long()
    cy = torch.floor(ys).long()
    gx = cx[:, None, None] + off[None, None, :]
    gy = cy[:, None, None] + off[None, :, None]
    dx = gx.float() - xs[:, None, None]
    dy = gy.float() - ys[:, None, None]
    val = peaks[:, None, None] * torch.exp(-(dx * dx + dy * dy) / (2.0 * sigmas[:, None, None] ** 2))
    ok = (gx >= 0) & (gx < w) & (gy >= 0) & (gy < h)
    idx = (gy * w + gx)[ok]
    frame.view(-1).index_add_(0, idx, val[ok])


def make_calibration_raw(kind, count, height, width, seed, device="cuda", dark_exposure_s=60.0):
    g = torch.Generator(device=device).manual_seed(seed)
    out = torch.empty((count, height * width), dtype=torch.int16, device=device)
    vig = vignetting(height, width, device) if kind == "flat" else None
    for i in range(count):
        noise = torch.randn((height, width), generator=g, device=device)
        if kind == "bias":
            frame = 0.02 + 0.003 * noise
        elif kind == "dark":
            frame = 0.02 + 0.001 * dark_exposure_s + 0.003 * noise
        else:
            frame = 0.02 + 0.4 * vig * (1.0 + 0.01 * noise)
        out[i] = _encode_be(frame).view(-1)
    return out


def make_light_raw(count, height, width, seed, *, n_stars=3000, fwhm=(2.0, 4.0), sky=0.02, noise=0.01,
                   max_shift=60.0, max_rot_deg=1.0, blurred=(), light_exposure_s=60.0, device="cuda",
                   cosmic_density=0.0):
    """Returns (raw int16 [count, H*W], moving->reference matrices [count,3,3])."""
    rng = np.random.default_rng(seed)
    g = torch.Generator(device=device).manual_seed(seed + 1)
    star_x = torch.tensor(rng.uniform(0, width - 1, n_stars), device=device, dtype=torch.float32)
    star_y = torch.tensor(rng.uniform(0, height - 1, n_stars), device=device, dtype=torch.float32)
    peaks = torch.tensor(np.exp(rng.uniform(math.log(0.05), math.log(0.9), n_stars)), device=device,
                         dtype=torch.float32)
    fw = torch.tensor(rng.uniform(fwhm[0], fwhm[1], n_stars), device=device, dtype=torch.float32)
    vig = vignetting(height, width, device)
    mats = np.empty((count, 3, 3), dtype=np.float64)
    out = torch.empty((count, height * width), dtype=torch.int16, device=device)
    blurred = set(blurred)
    for i in range(count):
        m = np.eye(3) if i == 0 else affine_matrix(rng.uniform(-max_shift, max_shift),
                                                   rng.uniform(-max_shift, max_shift),
                                                   rng.uniform(-max_rot_deg, max_rot_deg))
        mats[i] = m
        inv = np.linalg.inv(m)
        px = float(inv[0, 0]) * star_x + float(inv[0, 1]) * star_y + float(inv[0, 2])
        py = float(inv[1, 0]) * star_x + float(inv[1, 1]) * star_y + float(inv[1, 2])
        scale = 2.0 if i in blurred else 1.0
        sig = fw * scale / 2.3548200450309493
        signal = torch.full((height, width), sky, device=device, dtype=torch.float32)
        render_stars_(signal, px, py, peaks / (scale * scale), sig, int(math.ceil(4.0 * 2.0 * scale)))
        raw = 0.02 + 0.001 * light_exposure_s + vig * signal + noise * torch.randn(
            (height, width), generator=g, device=device)
        if cosmic_density > 0:
            hits = torch.rand((height, width), generator=g, device=device) < cosmic_density
            raw = torch.where(hits, raw + 0.75, raw)
        out[i] = _encode_be(raw).view(-1)
    return out, mats
