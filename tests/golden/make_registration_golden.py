"""Write tests/golden/registration_golden.npz: input/output vectors of the
reference's numpy-only CPU registration functions.

The last bits of ``np.linalg.lstsq`` / ``np.linalg.svd`` inside the
similarity and affine fits, and the tie orders of star detection and star
matching, are fixed by no reference test, fixture or closed form.  This script
imports the reference's own ``glass.cpu.star_detect`` and
``glass.cpu.registration`` (pure numpy; src/glass/cpu/star_detect.py:121-168,
src/glass/cpu/registration.py:361-566) in a container where the reference is
mounted, runs them on seeded synthetic inputs and stores inputs and outputs.
Nothing else reads the reference: the fixture file is data only.

Usage (from the repository root):
    python tests/golden/make_registration_golden.py /root/reference
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

STAR_FIELDS = ("x", "y", "flux", "peak", "fwhm_px", "eccentricity", "snr", "sharpness", "npix")


def star_field(rng, height, width, n_stars, sky=100.0, noise=3.0, fwhm=(2.0, 3.5), peak=(30.0, 900.0)):
    yy, xx = np.mgrid[0:height, 0:width]
    img = np.full((height, width), sky, np.float64)
    xs = rng.uniform(4, width - 5, n_stars)
    ys = rng.uniform(4, height - 5, n_stars)
    for x, y in zip(xs, ys):
        s = rng.uniform(*fwhm) / 2.3548
        a = np.exp(rng.uniform(np.log(peak[0]), np.log(peak[1])))
        img += a * np.exp(-((xx - x) ** 2 + (yy - y) ** 2) / (2 * s * s))
    img += rng.normal(0.0, noise, img.shape)
    return img.astype(np.float32)


def stars_to_arrays(stars):
    out = {}
    for f in STAR_FIELDS:
        vals = [getattr(s, f) for s in stars]
        if f == "npix":
            out[f] = np.asarray(vals, np.int64)
        else:
            out[f] = np.asarray([np.nan if v is None else float(v) for v in vals], np.float64)
    return out


def detection_cases(rng):
    cases = []
    img = star_field(rng, 96, 128, 40)
    cases.append(("plain", img, dict(threshold_sigma=5.0, max_stars=500)))
    img = star_field(rng, 80, 112, 30)
    img[10:13, 20:40] = np.nan
    img[50, 60] = np.inf
    # a close pair and a flat-topped plateau (ties of the local-maximum test)
    img[30:32, 70:72] = 2000.0
    img[60, 15] = img[60, 17] = 1500.0
    cases.append(("nan_pairs_plateau", img, dict(threshold_sigma=4.0, max_stars=500, min_separation_px=3.0)))
    img = star_field(rng, 64, 64, 25)
    cases.append(("given_background_truncated", img, dict(threshold_sigma=5.0, max_stars=10, background=100.0,
                                                          noise=3.0, window_radius=3, origin_x=7, origin_y=-3)))
    return cases


def transform_pairs(rng):
    """(name, reference xy, moving xy, true moving->reference matrix)."""
    out = []
    for name, n, angle_deg, scale, t, jitter, outliers, drop in (
        ("translation_only", 60, 0.0, 1.0, (12.3, -7.8), 0.05, 4, 5),
        ("small_rotation", 70, 0.45, 1.0002, (-18.6, 9.1), 0.08, 6, 8),
        ("rotation_scale", 50, 2.0, 0.997, (25.2, 14.7), 0.1, 5, 4),
        ("sheared_affine", 55, 1.0, 1.0, (-4.4, 3.3), 0.05, 3, 6),
    ):
        ref = rng.uniform(10, 1000, (n, 2))
        a = np.deg2rad(angle_deg)
        lin = scale * np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
        if name == "sheared_affine":
            lin = lin @ np.array([[1.0, 0.01], [0.0, 0.995]])
        m = np.eye(3)
        m[:2, :2] = lin
        m[:2, 2] = t
        inv = np.linalg.inv(m)
        mov = (np.c_[ref, np.ones(n)] @ inv.T)[:, :2] + rng.normal(0, jitter, (n, 2))
        keep = np.ones(n, bool)
        keep[rng.choice(n, drop, replace=False)] = False
        mov = mov[keep]
        mov = np.vstack([mov, rng.uniform(10, 1000, (outliers, 2))])
        order = rng.permutation(len(mov))
        out.append((name, ref, mov[order], m))
    return out


def main(reference_root: str) -> None:
    sys.path.insert(0, str(Path(reference_root) / "src"))
    from glass.cpu.registration import estimate_star_transform, estimate_triangle_asterism_transform
    from glass.cpu.star_detect import Star, detect_stars

    rng = np.random.default_rng(20260412)
    out: dict[str, np.ndarray] = {}
    det_names = []
    for name, img, kw in detection_cases(rng):
        det_names.append(name)
        stars = detect_stars(img, **kw)
        out[f"det/{name}/image"] = img
        for k in ("threshold_sigma", "max_stars", "background", "noise", "window_radius", "min_separation_px",
                  "origin_x", "origin_y"):
            out[f"det/{name}/{k}"] = np.asarray(np.nan if kw.get(k) is None else kw[k], np.float64)
        for f, arr in stars_to_arrays(stars).items():
            out[f"det/{name}/{f}"] = arr
    out["det_names"] = np.asarray(det_names)

    reg_names = []
    for name, ref, mov, truth in transform_pairs(rng):
        flux_r = np.linspace(1000, 10, len(ref))
        flux_m = np.linspace(900, 9, len(mov))
        ref_stars = [Star(x=float(x), y=float(y), flux=float(f)) for (x, y), f in zip(ref, flux_r)]
        mov_stars = [Star(x=float(x), y=float(y), flux=float(f)) for (x, y), f in zip(mov, flux_m)]
        out[f"reg/{name}/reference_xy"] = ref
        out[f"reg/{name}/moving_xy"] = mov
        out[f"reg/{name}/truth"] = truth
        runs = [(f"star_{m}", lambda m=m: estimate_star_transform(ref_stars, mov_stars, m))
                for m in ("translation", "similarity", "affine")]
        runs += [(f"triangle_{m}", lambda m=m: estimate_triangle_asterism_transform(ref_stars, mov_stars, m))
                 for m in ("similarity", "affine")]
        for tag, fn in runs:
            r = fn()
            key = f"reg/{name}/{tag}"
            out[f"{key}/matrix"] = np.asarray(r.matrix, np.float64)
            out[f"{key}/pairs"] = np.asarray(r.pairs, np.int64).reshape(-1, 2)
            out[f"{key}/rms_px"] = np.asarray(r.rms_px, np.float64)
            out[f"{key}/status"] = np.asarray(r.status)
        reg_names.append(name)
    out["reg_names"] = np.asarray(reg_names)
    np.savez_compressed(HERE / "registration_golden.npz", **out)
    print(f"wrote {HERE / 'registration_golden.npz'}: {len(det_names)} detection cases, "
          f"{len(reg_names)} star-list pairs x 5 estimators")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
