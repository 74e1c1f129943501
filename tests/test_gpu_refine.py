"""GPU checks of the affine Lanczos3 adjoint (config-4 registration
refinement).  There is no reference implementation of the derivatives: the
forward is compared with the FP64 CPU-path warp, the parameter gradient with
FP64 central differences of that warp (h = 1e-4 px), the pixel adjoint with
the dot-product identity <A u, v> = <u, A^T v>, and Gauss-Newton with a
known transform."""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle.warp import warp_lanczos3_source_map

pytestmark = pytest.mark.gpu


def _scene_at(seed, h, w, xx, yy, n_stars=40, min_sigma=1.2):
    """Analytic scene (smooth background + Gaussian stars) at coordinates."""
    rng = np.random.default_rng(seed)
    img = 100.0 + 5.0 * np.sin(xx / 17.0) * np.cos(yy / 23.0)
    for _ in range(n_stars):
        x, y, f, s = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(50, 400), rng.uniform(min_sigma, 2.5)
        img = img + f * np.exp(-((xx - x) ** 2 + (yy - y) ** 2) / (2 * s * s))
    return img.astype(np.float32)


def _scene(seed, h, w, n_stars=40):
    yy, xx = np.indices((h, w), dtype=np.float64)
    return _scene_at(seed, h, w, xx, yy, n_stars)


def _map(angle_deg, scale, tx, ty, h, w):
    a = math.radians(angle_deg)
    c, s = scale * math.cos(a), scale * math.sin(a)
    cx, cy = (w - 1) / 2, (h - 1) / 2
    # rotate/scale about the centre, then shift: reference pixel -> source pixel
    return np.array([c, -s, cx - c * cx + s * cy + tx, s, c, cy - s * cx - c * cy + ty], dtype=np.float64)


def test_forward_matches_fp64_cpu_warp(gpu):
    import torch
    from glass_amd.refine import affine_lanczos3

    h, w = 90, 120
    img = _scene(1, h, w)
    b = _map(0.8, 1.002, 1.3, -0.7, h, w)
    got = affine_lanczos3(torch.from_numpy(img).cuda(), torch.from_numpy(b)).cpu().numpy()
    exp = warp_lanczos3_source_map(img, b, 0.0, np.float32)
    assert np.allclose(got, exp, rtol=0, atol=2e-4 * 400), np.max(np.abs(got - exp))
    assert np.array_equal(got == 0.0, exp == 0.0)


def test_pixel_adjoint_dot_product_identity(gpu):
    import torch
    from glass_amd.refine import affine_lanczos3

    h, w = 70, 96
    rng = np.random.default_rng(2)
    u = torch.from_numpy(rng.normal(size=(h, w)).astype(np.float32)).cuda().requires_grad_(True)
    v = torch.from_numpy(rng.normal(size=(h, w)).astype(np.float32)).cuda()
    b = torch.from_numpy(_map(-1.5, 0.99, 2.2, 0.4, h, w))
    out = affine_lanczos3(u, b)
    (out * v).sum().backward()
    lhs = float((out.detach().double() * v.double()).sum())
    rhs = float((u.detach().double() * u.grad.double()).sum())
    assert lhs == pytest.approx(rhs, rel=1e-5)


def test_parameter_gradient_matches_fp64_finite_differences(gpu):
    import torch
    from glass_amd.refine import affine_lanczos3

    h, w = 64, 80
    img = _scene(3, h, w, 25)
    rng = np.random.default_rng(4)
    g = rng.normal(size=(h, w))
    b = _map(0.6, 1.001, 0.37, -0.81, h, w)
    bt = torch.from_numpy(b.copy()).requires_grad_(True)
    out = affine_lanczos3(torch.from_numpy(img).cuda(), bt)
    (out.double() * torch.from_numpy(g).cuda()).sum().backward()
    analytic = bt.grad.numpy()
    # FP64 central differences, each parameter stepped so the source
    # coordinates move by about 1e-4 px.
    scale = np.array([w, h, 1.0, w, h, 1.0])
    numeric = np.zeros(6)
    for k in range(6):
        step = 1e-4 / scale[k]
        bp, bm = b.copy(), b.copy()
        bp[k] += step
        bm[k] -= step
        lp = float(np.sum(g * warp_lanczos3_source_map(img, bp)))
        lm = float(np.sum(g * warp_lanczos3_source_map(img, bm)))
        numeric[k] = (lp - lm) / (2 * step)
    rel = np.abs(analytic - numeric) / np.maximum(np.abs(numeric), 1e-6 * np.max(np.abs(numeric)))
    assert np.all(rel < 2e-3), (analytic, numeric, rel)


def test_gauss_newton_recovers_affine(gpu):
    from glass_amd import ResidentCalibratedStack
    from glass_amd.refine import source_map_of

    h, w = 192, 256
    # The same analytic scene rendered on the reference grid and, through a
    # known moving -> reference map, on the moving grid (no resampling, no
    # fill band), with stars wide enough for Lanczos3 to interpolate well.
    true_m = np.array([[math.cos(math.radians(0.3)), -math.sin(math.radians(0.3)), 3.4],
                       [math.sin(math.radians(0.3)), math.cos(math.radians(0.3)), -2.1], [0, 0, 1.0]])
    yy, xx = np.indices((h, w), dtype=np.float64)
    ref = _scene_at(5, h, w, xx, yy, 60, min_sigma=2.0)
    mx = true_m[0, 0] * xx + true_m[0, 1] * yy + true_m[0, 2]
    my = true_m[1, 0] * xx + true_m[1, 1] * yy + true_m[1, 2]
    mov = _scene_at(5, h, w, mx, my, 60, min_sigma=2.0)
    # start 0.3 px / 0.02 deg away
    a = math.radians(0.3 + 0.02)
    start = np.array([[math.cos(a), -math.sin(a), 3.4 + 0.3], [math.sin(a), math.cos(a), -2.1 - 0.3], [0, 0, 1.0]])
    with ResidentCalibratedStack(2, h, w) as s:
        s.upload_calibrated_frame(0, ref)
        s.upload_calibrated_frame(1, mov)
        refined, rep = s.refine_affine(0, 1, start, max_iterations=30, tolerance_px=1e-5)
    corners = np.array([[0, 0, 1], [w - 1, 0, 1], [0, h - 1, 1], [w - 1, h - 1, 1]], dtype=np.float64).T
    err = np.max(np.abs((refined @ corners)[:2] - (true_m @ corners)[:2]))
    assert err < 0.02, (err, rep)
    assert rep["iterations"] <= 30 and rep["rms"] < 1.0
