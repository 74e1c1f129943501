"""Mean / sigma-clip integration oracle (test infrastructure).

Restates, with float32 arithmetic in frame order, the resident sigma-clip
kernel cpp/cuda/integration_kernels.cu:1743-1900 and the fused matrix-warped
kernels :3840-4123 (driven by cpp/src/native_bindings.cpp:13227-13393 and
:14632-15115):

* frames of positive finite weight only, in order;
* mean: sum w*x and sum w over finite samples, master = sum / sum w (0 when
  sum w = 0), coverage = finite count;
* sigma clip: mean and population std (sqrt(sum (x - mean)^2 / n)) of the
  finite samples, thresholds mean -/+ k*std; with winsorize the samples are
  clipped to them and the clipped mean / std give new thresholds; samples
  below / above are counted as low / high rejections and the weighted mean of
  the rest is the master;
* fused forms: the samples are the frames resampled with the resident warp
  (``warp_matrix_resident``: non-finite Lanczos3 taps skipped, optional
  clamp); a sample inside the footprint adds to the geometric coverage and
  counts only when finite.

Every product and sum is rounded to float32 separately (no fused
multiply-add), as the library computes them.

Pinned by (tests/test_oracle_sigma_clip.py): the reference's fixtures
tests/test_cuda_resident_stack.py:2744-2866 with their values worked out by
hand, an FP64 brute force per pixel, and the warp-then-integrate composition
of tests/test_cuda_resident_stack.py:649-682.  The tile-local forms: the
fixtures of tests/test_cuda_resident_stack.py:470-603 with hand-worked pixel
values, and the invariants multiplier 1 = the plain integration, multiplier 0
= the integration without the target frames, multiplier 2 on every frame =
the same master with twice the weight.
"""

from __future__ import annotations

import numpy as np

from .warp import warp_matrix_resident

F32 = np.float32


def clip_integrate(samples, weights=None, rejection="sigma_clip", low_sigma=3.0, high_sigma=3.0, winsorize=True):
    """samples [F, H, W] float32 (NaN: not counted).  Returns a dict of
    float32 maps: master, weight, coverage, low, high."""
    x = np.asarray(samples, dtype=np.float32)
    f = x.shape[0]
    w = np.ones(f, np.float32) if weights is None else np.asarray(weights, dtype=np.float32)
    if w.ndim == 1:
        active = [k for k in range(f) if w[k] > 0.0 and np.isfinite(w[k])]
    else:
        # per-pixel weights [F, H, W]: the caller has made every sample whose
        # weight does not count NaN
        active = list(range(f))
    shape = x.shape[1:]
    zeros = lambda: np.zeros(shape, np.float32)  # noqa: E731
    out = {k: zeros() for k in ("master", "weight", "coverage", "low", "high")}
    if rejection == "mean":
        s, ws, cov = zeros(), zeros(), zeros()
        for k in active:
            m = np.isfinite(x[k])
            s = np.where(m, s + x[k] * w[k], s).astype(F32)
            ws = np.where(m, ws + w[k], ws).astype(F32)
            cov = np.where(m, cov + F32(1), cov).astype(F32)
        with np.errstate(invalid="ignore", divide="ignore"):
            out["master"] = np.where(ws > 0, s / np.where(ws > 0, ws, F32(1)), F32(0)).astype(F32)
        out["weight"], out["coverage"] = ws, cov
        return out
    lo_k, hi_k = F32(low_sigma), F32(high_sigma)
    mean, count = zeros(), zeros()
    for k in active:
        m = np.isfinite(x[k])
        mean = np.where(m, mean + x[k], mean).astype(F32)
        count = np.where(m, count + F32(1), count).astype(F32)
    has = count > 0
    safe = np.where(has, count, F32(1))
    with np.errstate(invalid="ignore"):
        mean = (mean / safe).astype(F32)
        var = zeros()
        for k in active:
            m = np.isfinite(x[k])
            d = (x[k] - mean).astype(F32)
            var = np.where(m, var + d * d, var).astype(F32)
        center, scale = mean, np.sqrt((var / safe).astype(F32)).astype(F32)
        lo = (center - lo_k * scale).astype(F32)
        hi = (center + hi_k * scale).astype(F32)
        if winsorize:
            wm = zeros()
            for k in active:
                m = np.isfinite(x[k])
                c = np.where(x[k] < lo, lo, np.where(x[k] > hi, hi, x[k])).astype(F32)
                wm = np.where(m, wm + c, wm).astype(F32)
            wm = (wm / safe).astype(F32)
            wv = zeros()
            for k in active:
                m = np.isfinite(x[k])
                c = np.where(x[k] < lo, lo, np.where(x[k] > hi, hi, x[k])).astype(F32)
                d = (c - wm).astype(F32)
                wv = np.where(m, wv + d * d, wv).astype(F32)
            center, scale = wm, np.sqrt((wv / safe).astype(F32)).astype(F32)
            lo = (center - lo_k * scale).astype(F32)
            hi = (center + hi_k * scale).astype(F32)
        s, ws, cov, nlo, nhi = zeros(), zeros(), zeros(), zeros(), zeros()
        for k in active:
            m = np.isfinite(x[k])
            low = m & (x[k] < lo)
            high = m & ~low & (x[k] > hi)
            keep = m & ~low & ~high
            nlo = np.where(low, nlo + F32(1), nlo).astype(F32)
            nhi = np.where(high, nhi + F32(1), nhi).astype(F32)
            s = np.where(keep, s + x[k] * w[k], s).astype(F32)
            ws = np.where(keep, ws + w[k], ws).astype(F32)
            cov = np.where(keep, cov + F32(1), cov).astype(F32)
        master = np.where(ws > 0, s / np.where(ws > 0, ws, F32(1)), F32(0)).astype(F32)
    for key, val in (("master", master), ("weight", ws), ("coverage", cov), ("low", nlo), ("high", nhi)):
        out[key] = np.where(has, val, F32(0)).astype(F32)
    return out


def tile_multiplier_map(shape, tile_extents, tile_multipliers):
    """Per-pixel multiplier: that of the first tile (x0, y0, x1, y1), half
    open, holding the pixel; 1 outside every tile
    (cpp/cuda/integration_kernels.cu:1678-1690)."""
    m = np.ones(shape, np.float32)
    done = np.zeros(shape, bool)
    for (x0, y0, x1, y1), mult in zip(np.asarray(tile_extents).reshape(-1, 4), np.asarray(tile_multipliers, F32)):
        region = np.zeros(shape, bool)
        region[int(y0):int(y1), int(x0):int(x1)] = True
        region &= ~done
        m[region] = mult
        done |= region
    return m


def tile_local_integrate(stack, weights, target_mask, tile_extents, tile_multipliers, rejection="mean",
                         low_sigma=3.0, high_sigma=3.0, winsorize=True):
    """integrate_tile_local_mean / integrate_tile_local_sigma_clip
    (cpp/cuda/integration_kernels.cu:1658-1741, :3622-3795): the weight of
    frame k at a pixel is weights[k], times the pixel's tile multiplier when
    target_mask[k] is set (one float32 product); a sample counts only where
    that weight is positive and finite (in the statistics as well as the
    weighted sum).  Returns the maps of ``clip_integrate``."""
    x = np.asarray(stack, dtype=np.float32)
    f = x.shape[0]
    w = np.ones(f, np.float32) if weights is None else np.asarray(weights, dtype=np.float32)
    mult = tile_multiplier_map(x.shape[1:], tile_extents, tile_multipliers)
    target = np.asarray(target_mask).reshape(f) != 0
    eff = np.stack([(w[k] * mult).astype(F32) if target[k] else np.full(x.shape[1:], w[k], F32) for k in range(f)])
    with np.errstate(invalid="ignore"):
        counts = (eff > 0) & np.isfinite(eff)
    samples = np.where(counts, x, np.float32(np.nan)).astype(F32)
    return clip_integrate(samples, eff, rejection, low_sigma, high_sigma, winsorize)


def matrix_warped_samples(frames, matrices, interpolation="bilinear", clamping_threshold=-1.0):
    """(samples [F, H, W] with NaN where a frame does not count, footprint
    flags [F, H, W]) of the fused forms."""
    samples, footprint = [], []
    for frame, matrix in zip(frames, matrices):
        warped, cov = warp_matrix_resident(frame, matrix, fill=np.nan, interpolation=interpolation,
                                           clamping_threshold=clamping_threshold)
        fp = cov > 0.5
        samples.append(np.where(fp & np.isfinite(warped), warped, np.nan).astype(F32))
        footprint.append(fp)
    return np.stack(samples), np.stack(footprint)


def matrix_warped_integrate(frames, matrices, weights=None, interpolation="bilinear", clamping_threshold=-1.0,
                            rejection="mean", low_sigma=3.0, high_sigma=3.0, winsorize=True):
    samples, footprint = matrix_warped_samples(frames, matrices, interpolation, clamping_threshold)
    f = samples.shape[0]
    w = np.ones(f, np.float32) if weights is None else np.asarray(weights, dtype=np.float32)
    out = clip_integrate(samples, w, rejection, low_sigma, high_sigma, winsorize)
    active = (w > 0) & np.isfinite(w)
    out["geometric"] = np.sum(footprint[active], axis=0, dtype=np.float32) if active.any() else \
        np.zeros(samples.shape[1:], np.float32)
    return out
