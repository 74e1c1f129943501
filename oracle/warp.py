"""Affine warp oracle (test infrastructure; see oracle/__init__.py).

``warp_matrix`` restates the CPU path src/glass/engine/warp.py:155-228
(``_warp_tile_matrix``) with ``_sample_bilinear`` (:99-118) and
``_sample_windowed`` (:121-152, Lanczos3 weights :73-84), vectorised over
output pixels but with the same per-pixel arithmetic and tap order:
* source coordinates in FP64 from ``np.linalg.inv`` of the moving->reference
  matrix: ``(i00*x + i01*y) + i02`` divided by the projective denominator;
* bilinear: valid 0 <= s <= dim-1, f32 lerps, x1 = min(x0+1, W-1);
* Lanczos3: valid 2 <= s < dim-3, taps floor(s)-2 .. floor(s)+3, FP64
  weights sinc(x)*sinc(x/3) with sinc(|x|<1e-8)=1, taps with weight exactly 0
  skipped, FP64 sums, result 0 when |sum w| <= 1e-12; non-finite taps are NOT
  skipped on this path.

``warp_matrix_resident`` restates the resident-backend semantics that the
reference GPU tests pin with their own brute force
(tests/test_gpu_warp_vs_cpu.py:35-128): non-finite taps skipped, sinc
threshold 1e-6, optional clamp to [min - t*range, max + t*range] of the taps.
"""

from __future__ import annotations

import numpy as np


def invert_matrix(matrix):
    """FP64 inverse as engine/warp.py:169-176 (np.linalg.inv)."""
    m = np.asarray(matrix, dtype=np.float64)
    if m.shape != (3, 3):
        raise ValueError(f"warp matrix must be 3x3, got {m.shape}")
    try:
        return np.linalg.inv(m)
    except np.linalg.LinAlgError as exc:  # pragma: no cover - same message as reference
        raise ValueError("warp matrix is singular") from exc


def source_coordinates(inverse, height, width, x0=0, y0=0):
    """engine/warp.py:177-195: FP64 source coordinates of output pixels
    [y0, y0 + height) x [x0, x0 + width), NaN where the projective
    denominator is <= 1e-12 in magnitude."""
    ys, xs = np.mgrid[y0:y0 + height, x0:x0 + width].astype(np.float64)
    inv = np.asarray(inverse, dtype=np.float64)
    denom = inv[2, 0] * xs + inv[2, 1] * ys + inv[2, 2]
    ok = np.abs(denom) > 1.0e-12
    sx = np.divide(inv[0, 0] * xs + inv[0, 1] * ys + inv[0, 2], denom, out=np.full_like(xs, np.nan), where=ok)
    sy = np.divide(inv[1, 0] * xs + inv[1, 1] * ys + inv[1, 2], denom, out=np.full_like(ys, np.nan), where=ok)
    return sx, sy


def _sinc(x, threshold):
    x = np.asarray(x, dtype=np.float64)
    pix = np.pi * x
    with np.errstate(divide="ignore", invalid="ignore"):
        value = np.sin(pix) / pix
    return np.where(np.abs(x) < threshold, 1.0, value)


def lanczos3_weight(x, threshold=1.0e-8):
    """engine/warp.py:73-84 (threshold 1e-8) / test oracle :67-77 (1e-6)."""
    ax = np.abs(np.asarray(x, dtype=np.float64))
    w = _sinc(ax, threshold) * _sinc(ax / 3.0, threshold)
    return np.where(ax >= 3.0, 0.0, w)


def _bilinear(image, sx, sy, origin=(0, 0), full_shape=None):
    """`image` holds the source rows/columns from `origin` (x, y) of a frame
    of `full_shape` (the whole frame by default)."""
    h, w = full_shape if full_shape is not None else image.shape
    ox, oy = origin
    x0 = np.floor(sx).astype(np.int64)
    y0 = np.floor(sy).astype(np.int64)
    x1 = np.minimum(x0 + 1, w - 1)
    y1 = np.minimum(y0 + 1, h - 1)
    tx = (sx - x0).astype(np.float32)
    ty = (sy - y0).astype(np.float32)
    one = np.float32(1.0)
    x0, x1, y0, y1 = x0 - ox, x1 - ox, y0 - oy, y1 - oy
    top = image[y0, x0] * (one - tx) + image[y0, x1] * tx
    bottom = image[y1, x0] * (one - tx) + image[y1, x1] * tx
    return (top * (one - ty) + bottom * ty).astype(np.float32)


def _lanczos_cpu(image, sx, sy, origin=(0, 0)):
    """engine/warp.py:121-152 vectorised: yy outer, xx inner, FP64 sums.
    `image` holds the source from `origin` (x, y) on."""
    ox, oy = origin
    bx = np.floor(sx).astype(np.int64)
    by = np.floor(sy).astype(np.int64)
    acc = np.zeros(sx.shape, dtype=np.float64)
    wsum = np.zeros(sx.shape, dtype=np.float64)
    for dy in range(-2, 4):
        yy = by + dy
        wy = lanczos3_weight(sy - yy.astype(np.float64), 1.0e-8)
        for dx in range(-2, 4):
            xx = bx + dx
            wx = lanczos3_weight(sx - xx.astype(np.float64), 1.0e-8)
            take = (wy != 0.0) & (wx != 0.0)
            weight = wx * wy
            value = image[yy - oy, xx - ox].astype(np.float64)
            acc = np.where(take, acc + value * weight, acc)
            wsum = np.where(take, wsum + weight, wsum)
    out = np.zeros(sx.shape, dtype=np.float32)
    ok = np.abs(wsum) > 1.0e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        out[ok] = (acc[ok] / wsum[ok]).astype(np.float32)
    return out


def _lanczos_resident(image, sx, sy, clamping_threshold):
    """tests/test_gpu_warp_vs_cpu.py:80-128 vectorised."""
    bx = np.floor(sx).astype(np.int64)
    by = np.floor(sy).astype(np.int64)
    acc = np.zeros(sx.shape, dtype=np.float64)
    wsum = np.zeros(sx.shape, dtype=np.float64)
    vmin = np.full(sx.shape, np.inf)
    vmax = np.full(sx.shape, -np.inf)
    for dy in range(-2, 4):
        yy = by + dy
        wy = lanczos3_weight(sy - yy.astype(np.float64), 1.0e-6)
        for dx in range(-2, 4):
            xx = bx + dx
            value = image[yy, xx].astype(np.float64)
            fin = np.isfinite(value)
            weight = wy * lanczos3_weight(sx - xx.astype(np.float64), 1.0e-6)
            acc = np.where(fin, acc + np.where(fin, value, 0.0) * weight, acc)
            wsum = np.where(fin, wsum + weight, wsum)
            vmin = np.where(fin, np.minimum(vmin, value), vmin)
            vmax = np.where(fin, np.maximum(vmax, value), vmax)
    ok = np.abs(wsum) > 1.0e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        value = acc / wsum
    if clamping_threshold >= 0.0:
        rng = vmax - vmin
        clamped = np.minimum(vmax + clamping_threshold * rng, np.maximum(vmin - clamping_threshold * rng, value))
        value = np.where(np.isfinite(vmin), clamped, value)
    return value.astype(np.float32), ok


def warp_matrix(image, matrix, fill=0.0, interpolation="bilinear"):
    """CPU-path matrix warp of a whole frame (engine/warp.py:155-228).

    Returns (warped f32, coverage f32 0/1)."""
    img = np.asarray(image, dtype=np.float32)
    h, w = img.shape
    out = np.full((h, w), fill, dtype=np.float32)
    cov = np.zeros((h, w), dtype=np.float32)
    sx, sy = source_coordinates(invert_matrix(matrix), h, w)
    valid = np.isfinite(sx) & np.isfinite(sy)
    if interpolation == "bilinear":
        valid &= (sx >= 0.0) & (sx <= float(w - 1)) & (sy >= 0.0) & (sy <= float(h - 1))
    elif interpolation == "lanczos3":
        valid &= (sx >= 2.0) & (sx < float(w - 3)) & (sy >= 2.0) & (sy < float(h - 3))
    else:
        raise ValueError(f"unsupported warp interpolation: {interpolation}")
    if not np.any(valid):
        return out, cov
    vx, vy = sx[valid], sy[valid]
    if interpolation == "bilinear":
        sampled = _bilinear(img, vx, vy)
    else:
        sampled = _lanczos_cpu(img, vx, vy)
    out[valid] = sampled
    cov[valid] = 1.0
    return out, cov


def warp_matrix_window(image, matrix, window, origin=(0, 0), full_shape=None, fill=0.0, interpolation="lanczos3"):
    """``warp_matrix`` restricted to the output window (x0, y0, width,
    height) of a frame of `full_shape` (H, W): the same per-pixel arithmetic
    (coordinates from the absolute pixel indices, footprint against the full
    frame), with `image` holding the source from `origin` (x, y) on, a crop
    that must contain every tap of the window's in-footprint pixels.
    Returns (warped f32, coverage f32) of the window."""
    img = np.asarray(image, dtype=np.float32)
    h, w = full_shape if full_shape is not None else img.shape
    x0, y0, ww, wh = (int(v) for v in window)
    out = np.full((wh, ww), fill, dtype=np.float32)
    cov = np.zeros((wh, ww), dtype=np.float32)
    sx, sy = source_coordinates(invert_matrix(matrix), wh, ww, x0, y0)
    valid = np.isfinite(sx) & np.isfinite(sy)
    if interpolation == "bilinear":
        valid &= (sx >= 0.0) & (sx <= float(w - 1)) & (sy >= 0.0) & (sy <= float(h - 1))
        reach = (0, 1)
    elif interpolation == "lanczos3":
        valid &= (sx >= 2.0) & (sx < float(w - 3)) & (sy >= 2.0) & (sy < float(h - 3))
        reach = (-2, 3)
    else:
        raise ValueError(f"unsupported warp interpolation: {interpolation}")
    if not np.any(valid):
        return out, cov
    vx, vy = sx[valid], sy[valid]
    ox, oy = origin
    ix, iy = np.floor(vx).astype(np.int64), np.floor(vy).astype(np.int64)
    if (ix.min() + reach[0] < ox or iy.min() + reach[0] < oy or ix.max() + reach[1] >= ox + img.shape[1]
            or iy.max() + reach[1] >= oy + img.shape[0]):
        raise ValueError("source crop does not hold every tap of the window")
    if interpolation == "bilinear":
        out[valid] = _bilinear(img, vx, vy, origin, (h, w))
    else:
        out[valid] = _lanczos_cpu(img, vx, vy, origin)
    cov[valid] = 1.0
    return out, cov


def source_box(matrix, window, full_shape, interpolation="lanczos3"):
    """Source rows/columns (x0, y0, x1, y1), clipped to the frame, that hold
    every tap of the output window of an affine matrix (corners in FP64)."""
    inv = invert_matrix(matrix)
    x0, y0, ww, wh = window
    h, w = full_shape
    xs = np.array([x0, x0 + ww - 1, x0, x0 + ww - 1], np.float64)
    ys = np.array([y0, y0, y0 + wh - 1, y0 + wh - 1], np.float64)
    sx = (inv[0, 0] * xs + inv[0, 1] * ys) + inv[0, 2]
    sy = (inv[1, 0] * xs + inv[1, 1] * ys) + inv[1, 2]
    pad = 4 if interpolation == "lanczos3" else 2
    return (max(int(np.floor(sx.min())) - pad, 0), max(int(np.floor(sy.min())) - pad, 0),
            min(int(np.floor(sx.max())) + pad + 1, w), min(int(np.floor(sy.max())) + pad + 1, h))


def warp_matrix_resident(image, matrix, fill=np.nan, interpolation="lanczos3", clamping_threshold=-1.0):
    """Resident-backend semantics (tests/test_gpu_warp_vs_cpu.py:35-128).

    Bilinear is identical to the CPU path; Lanczos3 skips non-finite taps,
    uses the 1e-6 sinc threshold and the optional clamp.  A pixel whose finite
    tap weights sum to <= 1e-12 keeps the fill value and coverage 0.
    """
    img = np.asarray(image, dtype=np.float32)
    h, w = img.shape
    out = np.full((h, w), fill, dtype=np.float32)
    cov = np.zeros((h, w), dtype=np.float32)
    sx, sy = source_coordinates(invert_matrix(matrix), h, w)
    valid = np.isfinite(sx) & np.isfinite(sy)
    if interpolation == "bilinear":
        valid &= (sx >= 0.0) & (sx <= float(w - 1)) & (sy >= 0.0) & (sy <= float(h - 1))
        if np.any(valid):
            out[valid] = _bilinear(img, sx[valid], sy[valid])
            cov[valid] = 1.0
        return out, cov
    valid &= (sx >= 2.0) & (sx < float(w - 3)) & (sy >= 2.0) & (sy < float(h - 3))
    if not np.any(valid):
        return out, cov
    value, ok = _lanczos_resident(img, sx[valid], sy[valid], float(clamping_threshold))
    idx = np.flatnonzero(valid.ravel())[ok]
    out.ravel()[idx] = value[ok]
    cov.ravel()[idx] = 1.0
    return out, cov


def warp_lanczos3_source_map(image, source_map, fill=0.0, dtype=np.float64):
    """CPU-path Lanczos3 (engine/warp.py:121-152, no tap skipping) for an
    affine source map s = B (x, y, 1), B = source_map (2x3, reference pixel
    -> source pixel), returned in `dtype` (FP64 for finite differences).
    Footprint 2 <= s < dim-3 as engine/warp.py:207-212; `fill` elsewhere."""
    img = np.asarray(image, dtype=np.float32)
    h, w = img.shape
    b = np.asarray(source_map, dtype=np.float64).reshape(2, 3)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    sx = (b[0, 0] * xs + b[0, 1] * ys) + b[0, 2]
    sy = (b[1, 0] * xs + b[1, 1] * ys) + b[1, 2]
    ok = (sx >= 2.0) & (sx < w - 3) & (sy >= 2.0) & (sy < h - 3)
    out = np.full((h, w), fill, dtype=np.float64)
    if np.any(ok):
        px, py = sx[ok], sy[ok]
        bx, by = np.floor(px).astype(np.int64), np.floor(py).astype(np.int64)
        acc = np.zeros(px.shape)
        wsum = np.zeros(px.shape)
        for dy in range(-2, 4):
            wy = lanczos3_weight(py - (by + dy), 1.0e-8)
            for dx in range(-2, 4):
                wgt = wy * lanczos3_weight(px - (bx + dx), 1.0e-8)
                acc += img[by + dy, bx + dx].astype(np.float64) * wgt
                wsum += wgt
        out[ok] = acc / wsum
    return out.astype(dtype)


def translate_integer(image, dx: int, dy: int, fill=np.nan):
    """Whole-pixel translation (cpp/cuda/warp_kernels.cu:103-127):
    out[y, x] = in[y - dy, x - dx] inside the frame, `fill` elsewhere.
    Returns (out, coverage 0/1)."""
    img = np.asarray(image, dtype=np.float32)
    h, w = img.shape
    out = np.full((h, w), fill, dtype=np.float32)
    cov = np.zeros((h, w), dtype=np.float32)
    ys, xs = np.indices((h, w))
    sx, sy = xs - int(dx), ys - int(dy)
    ok = (sx >= 0) & (sx < w) & (sy >= 0) & (sy < h)
    out[ok] = img[sy[ok], sx[ok]]
    cov[ok] = 1.0
    return out, cov


def translate_bilinear_f32(image, dx: float, dy: float, fill=np.nan):
    """Bilinear translation with float32 coordinates
    (cpp/cuda/warp_kernels.cu:147-176): s = float32(x) - float32(dx), valid
    0 <= s <= dim - 1, far neighbour clamped to the edge, lerps in float32.
    Returns (out, coverage 0/1)."""
    img = np.asarray(image, dtype=np.float32)
    h, w = img.shape
    f32 = np.float32
    ys, xs = np.indices((h, w))
    sx = (xs.astype(f32) - f32(dx)).astype(f32)
    sy = (ys.astype(f32) - f32(dy)).astype(f32)
    ok = ~((sx < 0) | (sx > f32(w - 1)) | (sy < 0) | (sy > f32(h - 1)))
    out = np.full((h, w), fill, dtype=f32)
    cov = ok.astype(f32)
    x0 = np.floor(sx[ok]).astype(np.int64)
    y0 = np.floor(sy[ok]).astype(np.int64)
    x1 = np.where(x0 + 1 < w, x0 + 1, x0)
    y1 = np.where(y0 + 1 < h, y0 + 1, y0)
    tx = (sx[ok] - x0.astype(f32)).astype(f32)
    ty = (sy[ok] - y0.astype(f32)).astype(f32)
    one = f32(1)
    top = (img[y0, x0] * (one - tx) + img[y0, x1] * tx).astype(f32)
    bottom = (img[y1, x0] * (one - tx) + img[y1, x1] * tx).astype(f32)
    out[ok] = (top * (one - ty) + bottom * ty).astype(f32)
    return out, cov
