"""Star candidate catalogue oracle (test infrastructure).

Restates cpp/cuda/star_detect_kernels.cu as driven by
cpp/src/native_bindings.cpp:11356-12122:

* candidates (``is_local_maximum``, :41-69): interior pixels (not on the frame
  edge), finite, strictly above the threshold and not below any of the 8
  neighbours (``center < neighbour`` disqualifies; a NaN neighbour does not);
* rank order (``star_candidate_better``, :71-85): flux descending, then y,
  then x ascending;
* top-N: the first N in rank order; grid: cell (min(x*cols // W, cols-1),
  min(y*rows // H, rows-1)), the first k of each cell in rank order, merged
  in rank order (:236-315, sorted :500-605);
* separation filter (``star_catalog_nms_kernel``, :607-648): greedy in rank
  order, kept when the float32 squared distance to every kept entry is >=
  min_separation^2, entries of flux <= -1e30 skipped, up to max_output;
* catalogue translation (cpp/cuda/registration_kernels.cu:265-520): pair
  offsets inside the limits and prior vote, the first offset with the most
  offsets within tolerance wins, mutual nearest matches under it refine it;
* centroids (``star_refine_centroids_kernel``, :650-742): window radius
  clamped to 1..8 and to the frame, background the float32 median of the
  window's finite samples (middle pair averaged as 0.5 * (a + b)) unless an
  override is finite, weights max(v - background, 0) > 0, FP64 sums in raster
  order.

Pinned by tests/test_oracle_star_candidates.py: the literal cases of the
reference's tests (tests/test_gpu_registration_search.py:1339-1560 -- close
peaks, saturated plateau ties, spatial grid cells, a non-power-of-two grid
worked out by hand; tests/test_cuda_resident_stack.py:45-59 star field), and
closed forms of the centroid (symmetric stars centre exactly; a two-pixel
case by hand); the catalogue translation by the reference's star-field case
(tests/test_cuda_resident_stack.py:1866-1920: shift (4, -3) gives (-4, 3)
with rms 0) and hand-made catalogues (an outlier pair, limits, prior).
"""

from __future__ import annotations

import numpy as np

F32 = np.float32


def local_maxima(image, threshold):
    """[(x, y, flux)] in raster order."""
    img = np.asarray(image, dtype=np.float32)
    h, w = img.shape
    out = []
    thr = F32(threshold)
    for y in range(1, h - 1):
        for x in range(1, w - 1):
            c = img[y, x]
            if not np.isfinite(c) or not c > thr:
                continue
            patch = img[y - 1:y + 2, x - 1:x + 2]
            with np.errstate(invalid="ignore"):
                if np.any(c < patch):
                    continue
            out.append((float(x), float(y), float(c)))
    return out


def rank(cands):
    return sorted(cands, key=lambda c: (-c[2], c[1], c[0]))


def top_candidates(image, threshold, n):
    c = local_maxima(image, threshold)
    return len(c), rank(c)[:n]


def grid_top(image, threshold, cols, rows, per_cell):
    img = np.asarray(image)
    h, w = img.shape
    c = local_maxima(image, threshold)
    cells = {}
    for item in rank(c):
        cx = min((int(item[0]) * cols) // w, cols - 1)
        cy = min((int(item[1]) * rows) // h, rows - 1)
        cells.setdefault(cy * cols + cx, []).append(item)
    merged = [it for lst in cells.values() for it in lst[:per_cell]]
    return len(c), rank(merged)


def separation_filter(ranked, max_output, min_separation):
    sep2 = F32(min_separation) * F32(min_separation)
    kept = []
    for x, y, f in ranked:
        if len(kept) >= max_output:
            break
        if not np.isfinite(f) or f <= -1.0e30:
            continue
        ok = True
        for kx, ky, _ in kept:
            dx, dy = F32(x) - F32(kx), F32(y) - F32(ky)
            if F32(dx * dx) + F32(dy * dy) < sep2:
                ok = False
                break
        if ok:
            kept.append((x, y, f))
    return kept


def window_median_f32(values):
    v = np.sort(np.asarray(values, dtype=np.float32), kind="stable")
    n = v.size
    return v[n // 2] if n & 1 else F32(F32(0.5) * F32(v[n // 2 - 1] + v[n // 2]))


def refine_centroid(image, x, y, flux, radius, background=np.nan):
    """(x, y, moved) of one catalogue entry."""
    img = np.asarray(image, dtype=np.float32)
    h, w = img.shape
    if not np.isfinite(flux) or flux <= -1.0e30:
        return x, y, False
    cx, cy = int(np.floor(F32(x) + F32(0.5))), int(np.floor(F32(y) + F32(0.5)))
    if not (0 <= cx < w and 0 <= cy < h):
        return x, y, False
    r = min(max(int(radius), 1), 8)
    x0, x1, y0, y1 = max(0, cx - r), min(w - 1, cx + r), max(0, cy - r), min(h - 1, cy + r)
    win = img[y0:y1 + 1, x0:x1 + 1]
    bg = F32(background)
    if not np.isfinite(bg):
        fin = win[np.isfinite(win)]
        if fin.size == 0:
            return x, y, False
        bg = window_median_f32(fin)
    sw = sx = sy = 0.0
    for yy in range(y0, y1 + 1):
        for xx in range(x0, x1 + 1):
            v = img[yy, xx]
            if not np.isfinite(v):
                continue
            wf = max(F32(v - bg), F32(0.0))
            if not wf > 0:
                continue
            sw += float(wf)
            sx += float(xx) * float(wf)
            sy += float(yy) * float(wf)
    if not sw > 0.0:
        return x, y, False
    return float(F32(sx / sw)), float(F32(sy / sw)), True


def global_mean_background(image):
    img = np.asarray(image, dtype=np.float32)
    fin = img[np.isfinite(img)].astype(np.float64)
    return F32(fin.sum() / fin.size) if fin.size else F32(np.nan)


def grid_brightest_catalogue(image, threshold, cols, rows):
    """The brightest candidate of each cell in rank order, padded with (0, 0)
    entries to min(count, cells) (the reference's cell slots after its sort,
    star_detect_kernels.cu:1173-1195 and native_bindings.cpp:12783-12784)."""
    count, merged = grid_top(image, threshold, cols, rows, 1)
    n = min(count, cols * rows)
    cat = [(x, y) for x, y, _ in merged]
    return count, (cat + [(0.0, 0.0)] * n)[:n]


def catalogue_translation(ref, mov, tolerance_px, max_abs_dx=-1.0, max_abs_dy=-1.0, prior_dx=0.0, prior_dy=0.0,
                          prior_radius_px=-1.0):
    """cpp/cuda/registration_kernels.cu:265-520 on catalogues [(x, y)]:
    (dx, dy, inliers, refined_dx, refined_dy, mutual, rms), float32
    throughout; the refinement sums in moving order."""
    f = F32
    mad = f(max_abs_dx) if max_abs_dx >= 0 else f(-1)
    mady = f(max_abs_dy) if max_abs_dy >= 0 else mad
    prr = f(prior_radius_px) if prior_radius_px >= 0 else f(-1)
    offs = []
    for rx, ry in ref:
        for mx, my in mov:
            dx, dy = f(rx) - f(mx), f(ry) - f(my)
            bad = (mad >= 0 and abs(dx) > mad) or (mady >= 0 and abs(dy) > mady)
            if not bad and prr >= 0:
                ex, ey = dx - f(prior_dx), dy - f(prior_dy)
                bad = f(ex * ex) + f(ey * ey) > f(prr * prr)
            offs.append((f(np.nan), f(np.nan)) if bad else (dx, dy))
    tol2 = f(tolerance_px) * f(tolerance_px)
    best, best_score = -1, -1
    for i, (dx, dy) in enumerate(offs):
        if not (np.isfinite(dx) and np.isfinite(dy)):
            continue
        score = sum(1 for ox, oy in offs if np.isfinite(ox) and np.isfinite(oy)
                    and f(f((ox - dx) * (ox - dx)) + f((oy - dy) * (oy - dy))) <= tol2)
        if score > best_score:
            best, best_score = i, score
    bdx, bdy = offs[best] if best >= 0 else (f(np.nan), f(np.nan))

    mbest = [-1] * len(mov)
    for b, (mx, my) in enumerate(mov):
        tx, ty = f(mx) + bdx, f(my) + bdy
        d2best = tol2
        for a, (rx, ry) in enumerate(ref):
            ex, ey = f(rx) - tx, f(ry) - ty
            d2 = f(f(ex * ex) + f(ey * ey))
            if d2 <= d2best:
                d2best, mbest[b] = d2, a
    rbest = [-1] * len(ref)
    for a, (rx, ry) in enumerate(ref):
        d2best = tol2
        for b, (mx, my) in enumerate(mov):
            ex, ey = f(rx) - (f(mx) + bdx), f(ry) - (f(my) + bdy)
            d2 = f(f(ex * ex) + f(ey * ey))
            if d2 <= d2best:
                d2best, rbest[a] = d2, b
    pairs = [(mbest[b], b) for b in range(len(mov)) if mbest[b] >= 0 and rbest[mbest[b]] == b]
    sx = sy = f(0)
    for a, b in pairs:
        sx = f(sx + (f(ref[a][0]) - f(mov[b][0])))
        sy = f(sy + (f(ref[a][1]) - f(mov[b][1])))
    n = len(pairs)
    rdx, rdy = (f(sx * f(f(1) / f(n))), f(sy * f(f(1) / f(n)))) if n else (bdx, bdy)
    s2 = f(0)
    for a, b in pairs:
        ex, ey = f(ref[a][0]) - (f(mov[b][0]) + rdx), f(ref[a][1]) - (f(mov[b][1]) + rdy)
        s2 = f(s2 + f(f(ex * ex) + f(ey * ey)))
    rms = f(np.sqrt(f(s2 / f(n)))) if n else f(np.nan)
    return float(bdx), float(bdy), (best_score if best >= 0 else 0), float(rdx), float(rdy), n, float(rms)
