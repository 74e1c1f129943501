"""Star detection and star-list transform estimation oracle (test
infrastructure; see oracle/__init__.py).

Restates the reference CPU path that produces the warp matrices:

* ``detect_stars`` -- src/glass/cpu/star_detect.py:121-168 with
  ``robust_background`` (:29-42), ``_measure_candidate`` (:45-100) and
  ``_dedupe_stars`` (:103-118).  numpy semantics the product must follow:
  float32 medians average the middle pair in float32, the FP64 threshold is
  compared against the float32 image after rounding to float32 (weak Python
  scalar), candidate sums are FP64 over the contiguous patch in numpy's
  pairwise order, sorts are stable.
* ``estimate_star_transform`` -- src/glass/cpu/registration.py:476-566 with
  ``_translation_candidates`` (:218-233), ``_asterism_candidates`` (:334-358),
  ``_triangle_descriptors`` (:306-331), ``_greedy_pairs`` (:177-205) and the
  model fits (:101-174) for the translation, similarity and affine models.

Pinned bit for bit (lstsq / svd last bits, tie orders) by the golden vectors of
the reference's own functions in tests/golden/registration_golden.npz
(tests/golden/make_registration_golden.py), besides the reference tests'
literal cases.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from itertools import combinations, permutations

import numpy as np


@dataclass
class Star:
    x: float
    y: float
    flux: float
    peak: float = 0.0
    fwhm_px: float | None = None
    eccentricity: float | None = None
    snr: float | None = None
    sharpness: float | None = None
    npix: int = 0


@dataclass
class Registration:
    transform_model: str
    matrix: list
    matched_stars: int
    inliers: int
    rms_px: float
    status: str
    pairs: list = field(default_factory=list)


def robust_background(image):
    """star_detect.py:29-42: float32 median / MAD of the finite samples,
    noise = 1.4826 MAD, falling back to the float32 std, then 1."""
    finite = image[np.isfinite(image)]
    if finite.size == 0:
        return 0.0, 0.0, 0.0, 1.0
    median = float(np.median(finite))
    rms = float(np.std(finite))
    mad = float(np.median(np.abs(finite - median)))
    noise = float(1.4826 * mad)
    if noise <= 0.0:
        noise = rms
    if noise <= 0.0:
        noise = 1.0
    return median, rms, mad, noise


def measure_candidate(image, x, y, background, noise, radius, origin_x=0, origin_y=0):
    """star_detect.py:45-100: FP64 flux, centroid and second moments of the
    background-subtracted, zero-clipped window around (x, y)."""
    h, w = image.shape
    y0, y1 = max(0, y - radius), min(h, y + radius + 1)
    x0, x1 = max(0, x - radius), min(w, x + radius + 1)
    patch = image[y0:y1, x0:x1]
    ok = np.isfinite(patch)
    if not ok.any():
        return None
    wts = np.where(ok, np.maximum(patch - np.float32(background), 0.0), 0.0)
    flux = float(np.sum(wts, dtype=np.float64))
    if flux <= 0.0:
        return None
    yy, xx = np.mgrid[y0:y1, x0:x1]
    cx = float(np.sum(xx * wts, dtype=np.float64) / flux)
    cy = float(np.sum(yy * wts, dtype=np.float64) / flux)
    dx = xx - cx
    dy = yy - cy
    mxx = float(np.sum(dx * dx * wts, dtype=np.float64) / flux)
    myy = float(np.sum(dy * dy * wts, dtype=np.float64) / flux)
    mxy = float(np.sum(dx * dy * wts, dtype=np.float64) / flux)
    trace = max(mxx + myy, 0.0)
    det_term = max((mxx - myy) * (mxx - myy) + 4.0 * mxy * mxy, 0.0)
    root = float(np.sqrt(det_term))
    major = max((trace + root) / 2.0, 0.0)
    minor = max((trace - root) / 2.0, 0.0)
    sigma = float(np.sqrt(max((major + minor) / 2.0, 0.0)))
    fwhm = 2.354820045 * sigma if sigma > 0.0 else None
    ecc = float(np.sqrt(max(0.0, 1.0 - minor / major))) if major > 0.0 else None
    peak = float(max(float(image[y, x]) - background, 0.0))
    npix = int(np.count_nonzero(wts > 0.0))
    snr = flux / max(float(noise) * np.sqrt(max(npix, 1)), 1.0e-6)
    return Star(x=cx + float(origin_x), y=cy + float(origin_y), flux=flux, peak=peak,
                fwhm_px=None if fwhm is None else float(fwhm), eccentricity=ecc, snr=float(snr),
                sharpness=float(peak / max(flux, 1.0e-6)), npix=npix)


def dedupe_stars(stars, min_separation_px):
    """star_detect.py:103-118: brightest first (stable), drop stars closer
    than the separation to a kept one."""
    if min_separation_px <= 0.0 or len(stars) <= 1:
        return stars
    kept = []
    limit = float(min_separation_px * min_separation_px)
    for s in sorted(stars, key=lambda item: item.flux, reverse=True):
        if all((s.x - k.x) * (s.x - k.x) + (s.y - k.y) * (s.y - k.y) >= limit for k in kept):
            kept.append(s)
    return kept


def detect_stars(data, threshold_sigma=5.0, max_stars=500, *, background=None, noise=None, window_radius=4,
                 min_separation_px=2.0, origin_x=0, origin_y=0):
    """star_detect.py:121-168."""
    image = np.asarray(data, dtype=np.float32)
    median, _rms, _mad, sigma_bg = robust_background(image)
    median = median if background is None else float(background)
    sigma = sigma_bg if noise is None else float(noise)
    threshold = median + threshold_sigma * max(sigma, 1.0e-6)
    h, w = image.shape
    if h < 3 or w < 3:
        return []
    core = image[1:-1, 1:-1]
    mask = np.isfinite(core) & (core > threshold)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dx or dy:
                mask &= core >= image[1 + dy:h - 1 + dy, 1 + dx:w - 1 + dx]
    ys, xs = np.nonzero(mask)
    radius = max(1, int(window_radius))
    stars = []
    for ly, lx in zip(ys, xs):
        s = measure_candidate(image, int(lx + 1), int(ly + 1), median, sigma, radius, int(origin_x), int(origin_y))
        if s is not None:
            stars.append(s)
    stars = dedupe_stars(stars, float(min_separation_px))
    stars.sort(key=lambda s: s.flux, reverse=True)
    return stars[:max_stars]


# ------------------------------------------------------------ registration
def _points(stars, max_stars):
    return np.asarray([(s.x, s.y) for s in stars[: max(0, int(max_stars))]], dtype=np.float64)


def _translation(dx, dy):
    return np.asarray([[1.0, 0.0, dx], [0.0, 1.0, dy], [0.0, 0.0, 1.0]], dtype=np.float64)


def apply_matrix(points, matrix):
    """registration.py:46-51."""
    hom = np.hstack([points, np.ones((points.shape[0], 1), dtype=np.float64)])
    out = hom @ matrix.T
    wcol = out[:, 2:3]
    return np.divide(out[:, :2], wcol, out=out[:, :2].copy(), where=np.abs(wcol) > 1.0e-12)


def fit_translation(moving, reference):
    """registration.py:101-103."""
    d = np.median(reference - moving, axis=0)
    return _translation(d[0], d[1])


def fit_affine(moving, reference):
    """registration.py:106-119 (FP64 least squares)."""
    if len(moving) < 3:
        raise ValueError("affine registration requires at least three matched stars")
    a = np.zeros((2 * len(moving), 6), dtype=np.float64)
    a[0::2, 0:2] = moving
    a[0::2, 2] = 1.0
    a[1::2, 3:5] = moving
    a[1::2, 5] = 1.0
    rhs = np.asarray(reference, dtype=np.float64).reshape(-1)
    p = np.linalg.lstsq(a, rhs, rcond=None)[0]
    return np.asarray([[p[0], p[1], p[2]], [p[3], p[4], p[5]], [0.0, 0.0, 1.0]], dtype=np.float64)


def fit_similarity(moving, reference):
    """registration.py:140-162 (orthogonal Procrustes with scale)."""
    if len(moving) < 2:
        raise ValueError("similarity registration requires at least two matched stars")
    mm, rm = np.mean(moving, axis=0), np.mean(reference, axis=0)
    mc, rc = moving - mm, reference - rm
    var = float(np.sum(mc * mc))
    if var <= 1.0e-12:
        raise ValueError("degenerate moving stars for similarity registration")
    u, sv, vt = np.linalg.svd(mc.T @ rc)
    rot = vt.T @ u.T
    if np.linalg.det(rot) < 0:
        vt[-1, :] *= -1
        rot = vt.T @ u.T
    lin = float(np.sum(sv) / var) * rot
    t = rm - mm @ lin.T
    return np.asarray([[lin[0, 0], lin[0, 1], t[0]], [lin[1, 0], lin[1, 1], t[1]], [0.0, 0.0, 1.0]], dtype=np.float64)


def fit_model(model, moving, reference):
    if model == "translation":
        return fit_translation(moving, reference)
    if model == "similarity":
        return fit_similarity(moving, reference)
    if model == "affine":
        return fit_affine(moving, reference)
    raise ValueError(f"unsupported star registration model: {model}")


def greedy_pairs(reference, transformed, tolerance_px):
    """registration.py:177-205: candidate pairs within the tolerance, nearest
    first (stable over the row-major candidate order), each star used once."""
    if len(reference) == 0 or len(transformed) == 0:
        return [], float("inf")
    d = transformed[:, None, :] - reference[None, :, :]
    dist = np.sqrt(np.sum(d * d, axis=2))
    cand = sorted(((int(i), int(j)) for i, j in np.argwhere(dist <= tolerance_px)), key=lambda p: float(dist[p]))
    used_m, used_r, pairs, res = set(), set(), [], []
    for i, j in cand:
        if i in used_m or j in used_r:
            continue
        used_m.add(i)
        used_r.add(j)
        pairs.append((i, j))
        res.append(float(dist[i, j]))
    rms = float(np.sqrt(np.mean(np.square(res)))) if res else float("inf")
    return pairs, rms


def translation_candidates(reference, moving, brightest):
    """registration.py:218-233: one candidate per distinct rounded offset."""
    out, seen = [], set()
    for rp in reference[: min(len(reference), brightest)]:
        for mp in moving[: min(len(moving), brightest)]:
            d = rp - mp
            key = (int(round(d[0])), int(round(d[1])))
            if key not in seen:
                seen.add(key)
                out.append(_translation(d[0], d[1]))
    return out


def triangle_descriptors(points, max_points, max_triangles):
    """registration.py:306-331: (short/longest, middle/longest) of every
    non-degenerate triangle of the first max_points, largest area first."""
    pts = points[: min(len(points), max_points)]
    tris = []
    for i, j, k in combinations(range(len(pts)), 3):
        t = pts[[i, j, k]]
        ea, eb = t[1] - t[0], t[2] - t[0]
        area = abs(float(ea[0] * eb[1] - ea[1] * eb[0])) * 0.5
        if area < 2.0:
            continue
        sides = np.asarray([np.linalg.norm(t[0] - t[1]), np.linalg.norm(t[1] - t[2]), np.linalg.norm(t[2] - t[0])])
        longest = float(np.max(sides))
        if longest <= 1.0e-6:
            continue
        o = np.sort(sides)
        tris.append((np.asarray([o[0] / longest, o[1] / longest]), (i, j, k), area))
    tris.sort(key=lambda item: item[2], reverse=True)
    return [(d, idx) for d, idx, _a in tris[:max_triangles]]


def asterism_candidates(reference, moving, model, tolerance, max_triangles):
    """registration.py:334-358."""
    ref_t = triangle_descriptors(reference, 30, max_triangles)
    mov_t = triangle_descriptors(moving, 30, max_triangles)
    out = []
    for md, mi in mov_t:
        for rd, ri in ref_t:
            if float(np.linalg.norm(md - rd)) > tolerance:
                continue
            mt = moving[list(mi)]
            rt_raw = reference[list(ri)]
            for order in permutations(range(3)):
                try:
                    out.append(fit_model(model, mt, rt_raw[list(order)]))
                except ValueError:
                    continue
            if len(out) >= max_triangles:
                return out
    return out


def estimate_star_transform(reference, moving, transform_model="translation", min_inliers=6, tolerance_px=3.0,
                            max_stars=80):
    """registration.py:476-566 (translation / similarity / affine)."""
    ref, mov = _points(reference, max_stars), _points(moving, max_stars)
    minimum = 3 if transform_model == "affine" else 2 if transform_model == "similarity" else 1
    if len(ref) < minimum or len(mov) < minimum:
        return Registration(transform_model, _translation(0.0, 0.0).tolist(), 0, 0, float("nan"), "failed", [])
    cands = translation_candidates(ref, mov, 35)
    if transform_model in ("similarity", "affine"):
        cands = asterism_candidates(ref, mov, transform_model, 0.025, 600) + cands
    if not cands:
        cands = [np.eye(3, dtype=np.float64)]
    best, best_pairs, best_rms = cands[0], [], float("inf")
    for m in cands:
        pairs, rms = greedy_pairs(ref, apply_matrix(mov, m), tolerance_px)
        if len(pairs) > len(best_pairs) or (len(pairs) == len(best_pairs) and rms < best_rms):
            best, best_pairs, best_rms = m, pairs, rms
    for _ in range(2):
        if len(best_pairs) < minimum:
            break
        try:
            refined = fit_model(transform_model, mov[[p[0] for p in best_pairs]], ref[[p[1] for p in best_pairs]])
        except ValueError:
            break
        pairs, rms = greedy_pairs(ref, apply_matrix(mov, refined), tolerance_px)
        if len(pairs) > len(best_pairs) or (len(pairs) == len(best_pairs) and rms <= best_rms):
            best, best_pairs, best_rms = refined, pairs, rms
    status = "ok" if len(best_pairs) >= min_inliers else "failed"
    return Registration(transform_model, [[float(v) for v in row] for row in best], len(best_pairs),
                        len(best_pairs), float(best_rms) if np.isfinite(best_rms) else float("nan"), status,
                        best_pairs)


def ncc_translation_search(reference, moving, radius_x, radius_y, subpixel=False, center_dx=0.0, center_dy=0.0,
                           step=1.0, sample_stride=1):
    """Normalised cross-correlation translation search
    (cpp/cuda/registration_kernels.cu:10-264): candidates in row-major order,
    integer offsets or float32 offsets center + k * step sampled bilinearly in
    float32; FP64 sums over sampled reference pixels whose pair is finite;
    the first best score wins.  Returns (dx, dy, score as float32, scores)."""
    ref = np.asarray(reference, dtype=np.float32)
    mov = np.asarray(moving, dtype=np.float32)
    h, w = ref.shape
    ys, xs = np.meshgrid(np.arange(0, h, sample_stride), np.arange(0, w, sample_stride), indexing="ij")
    r_all = ref[ys, xs].astype(np.float64)
    f32 = np.float32
    scores = []
    offsets = []
    for oy in range(-radius_y, radius_y + 1):
        for ox in range(-radius_x, radius_x + 1):
            if subpixel:
                fdx = f32(f32(center_dx) + f32(ox) * f32(step))
                fdy = f32(f32(center_dy) + f32(oy) * f32(step))
                sx = (xs.astype(f32) - fdx).astype(f32)
                sy = (ys.astype(f32) - fdy).astype(f32)
                ok = ~((sx < 0) | (sx > f32(w - 1)) | (sy < 0) | (sy > f32(h - 1)))
                x0 = np.floor(sx[ok]).astype(np.int64)
                y0 = np.floor(sy[ok]).astype(np.int64)
                x1 = np.where(x0 + 1 < w, x0 + 1, x0)
                y1 = np.where(y0 + 1 < h, y0 + 1, y0)
                tx = (sx[ok] - x0.astype(f32)).astype(f32)
                ty = (sy[ok] - y0.astype(f32)).astype(f32)
                one = f32(1)
                top = (mov[y0, x0] * (one - tx) + mov[y0, x1] * tx).astype(f32)
                bot = (mov[y1, x0] * (one - tx) + mov[y1, x1] * tx).astype(f32)
                m = (top * (one - ty) + bot * ty).astype(f32)
                offsets.append((float(fdx), float(fdy)))
            else:
                sx, sy = xs - ox, ys - oy
                ok = (sx >= 0) & (sx < w) & (sy >= 0) & (sy < h)
                m = mov[sy[ok], sx[ok]]
                offsets.append((float(ox), float(oy)))
            r = r_all[ok]
            m = m.astype(np.float64)
            fin = np.isfinite(r) & np.isfinite(m)
            r, m = r[fin], m[fin]
            n = float(r.size)
            score = -float(np.finfo(np.float32).max)
            if n > 1:
                sr, sm = r.sum(), m.sum()
                num = (r * m).sum() - sr * sm / n
                vr = max((r * r).sum() - sr * sr / n, 0.0)
                vm = max((m * m).sum() - sm * sm / n, 0.0)
                den = np.sqrt(vr * vm)
                if den > 0:
                    score = num / den
            scores.append(score)
    best = int(np.argmax(np.asarray(scores)))  # first maximum
    return offsets[best][0], offsets[best][1], float(np.float32(scores[best])), np.asarray(scores)


# ---- translated-matrix candidates scored by alignment metrics
# (native_bindings.cpp:1459-1700, :10079-10184; warp_kernels.cu:1051-1172)

def translation_offsets(cx, cy, radius, step):
    """Float32 grid, dx outer (native_bindings.cpp:1489-1506)."""
    f = np.float32
    out = []
    dx = f(f(cx) - f(radius))
    while dx <= f(f(cx) + f(radius)) + f(f(step) * f(0.5)):
        dy = f(f(cy) - f(radius))
        while dy <= f(f(cy) + f(radius)) + f(f(step) * f(0.5)):
            out.append((dx, dy))
            dy = f(dy + f(step))
        dx = f(dx + f(step))
    return out or [(f(cx), f(cy))]


def invert_matrix_f32(m):
    """FP64 adjugate inverse rounded to float32 (native_bindings.cpp:1422-1457)."""
    a, b, c, d, e, f_, g, h, i = (float(v) for v in np.asarray(m, np.float32).ravel())
    cof = [e * i - f_ * h, -(d * i - f_ * g), d * h - e * g, -(b * i - c * h), a * i - c * g, -(a * h - b * g),
           b * f_ - c * e, -(a * f_ - c * d), a * e - b * d]
    det = a * cof[0] + b * cof[1] + c * cof[2]
    k = 1.0 / det
    return np.array([cof[0] * k, cof[3] * k, cof[6] * k, cof[1] * k, cof[4] * k, cof[7] * k, cof[2] * k,
                     cof[5] * k, cof[8] * k], np.float64).astype(np.float32)


def alignment_sums(reference, moving, inverse, stride):
    """The seven FP64 sums and the count of the alignment-metric kernel
    (float32 coordinates and bilinear taps, one rounding per operation)."""
    f = np.float32
    ref = np.asarray(reference, np.float32)
    mov = np.asarray(moving, np.float32)
    h, w = ref.shape
    inv = np.asarray(inverse, np.float32)
    ys, xs = np.meshgrid(np.arange(0, h, stride), np.arange(0, w, stride), indexing="ij")
    fx, fy = xs.astype(np.float32), ys.astype(np.float32)
    with np.errstate(all="ignore"):
        den = (inv[6] * fx + inv[7] * fy) + inv[8]
        sx = ((inv[0] * fx + inv[1] * fy) + inv[2]) / den
        sy = ((inv[3] * fx + inv[4] * fy) + inv[5]) / den
        ok = (np.abs(den) > f(1e-12)) & (sx >= 0) & (sx <= f(w - 1)) & (sy >= 0) & (sy <= f(h - 1))
        x0 = np.floor(np.where(ok, sx, 0)).astype(np.int64)
        y0 = np.floor(np.where(ok, sy, 0)).astype(np.int64)
        x1 = np.where(x0 + 1 < w, x0 + 1, x0)
        y1 = np.where(y0 + 1 < h, y0 + 1, y0)
        tx = (np.where(ok, sx, 0) - x0.astype(np.float32)).astype(np.float32)
        ty = (np.where(ok, sy, 0) - y0.astype(np.float32)).astype(np.float32)
        top = mov[y0, x0] * (f(1) - tx) + mov[y0, x1] * tx
        bottom = mov[y1, x0] * (f(1) - tx) + mov[y1, x1] * tx
        warped = top * (f(1) - ty) + bottom * ty
        r = ref[ys, xs]
        ok &= np.isfinite(r) & np.isfinite(warped)
    r = r[ok].astype(np.float64)
    m = warped[ok].astype(np.float64)
    d = m - r
    return np.array([r.sum(), m.sum(), (r * r).sum(), (m * m).sum(), (r * m).sum(), (d * d).sum(),
                     np.abs(d).sum(), float(ok.sum())])


def alignment_metrics(s):
    n = s[7]
    out = {"valid_pixels": int(n), "rms": np.nan, "mean_abs_diff": np.nan, "ncc": np.nan}
    if n == 0:
        return out
    out["rms"] = float(np.sqrt(s[5] / n))
    out["mean_abs_diff"] = float(s[6] / n)
    if n > 1:
        num = s[4] - s[0] * s[1] / n
        den = np.sqrt(max(s[2] - s[0] * s[0] / n, 0.0) * max(s[3] - s[1] * s[1] / n, 0.0))
        if den > 0:
            out["ncc"] = float(num / den)
    return out


def _better(c, b):
    cr, br = np.isfinite(c["rms"]), np.isfinite(b["rms"])
    if cr != br:
        return cr
    if cr and abs(c["rms"] - b["rms"]) > 1e-12:
        return c["rms"] < b["rms"]
    cn, bn = np.isfinite(c["ncc"]), np.isfinite(b["ncc"])
    if cn != bn:
        return cn
    if cn and abs(c["ncc"] - b["ncc"]) > 1e-12:
        return c["ncc"] > b["ncc"]
    return c["valid_pixels"] > b["valid_pixels"]


def _score(reference, moving, base, offsets, stride):
    best = None
    for dx, dy in offsets:
        m = np.asarray(base, np.float32).reshape(9).copy()
        m[2] = np.float32(m[2] + dx)
        m[5] = np.float32(m[5] + dy)
        c = alignment_metrics(alignment_sums(reference, moving, invert_matrix_f32(m), stride))
        c.update(dx=float(dx), dy=float(dy), matrix=m.reshape(3, 3))
        if best is None or _better(c, best):
            best = c
    return best


def refine_matrix_translation_candidates(reference, moving, seeds, search_radius_px=1.0, coarse_step_px=0.25,
                                         fine_radius_px=0.25, fine_step_px=0.0625, coarse_sample_stride=4,
                                         final_sample_stride=1):
    """(per-seed bests, selected index) of
    refine_matrix_translation_candidates_to_reference."""
    coarse = translation_offsets(0.0, 0.0, search_radius_px, coarse_step_px)
    results, sel = [], -1
    for base in np.asarray(seeds, np.float32).reshape(-1, 3, 3):
        cb = _score(reference, moving, base, coarse, coarse_sample_stride)
        best = cb
        if fine_radius_px > 0:
            best = _score(reference, moving, base, translation_offsets(cb["dx"], cb["dy"], fine_radius_px,
                                                                       fine_step_px), final_sample_stride)
        elif final_sample_stride != coarse_sample_stride:
            best = _score(reference, moving, base, [(np.float32(cb["dx"]), np.float32(cb["dy"]))],
                          final_sample_stride)
        results.append(best)
        if sel < 0 or _better(best, results[sel]):
            sel = len(results) - 1
    return results, sel
