"""Resident local normalisation oracle (test infrastructure).

Restates the reference's resident local normalisation, which runs between
warp and integration:

* ``pair_grid_stats``: per tile of the ceil(H/th) x ceil(W/tw) grid, over the
  pixels where both the source and the reference are finite, FP64 sums of
  s, s^2, r, r^2 and the count (cpp/cuda/local_norm_kernels.cu:220-304), then
  mean = sum * (1/n), var = max(sum2 * (1/n) - mean^2, 0), both rounded to
  float32; an empty tile reports zeros (cpp/src/native_bindings.cpp:10740-10768).
* ``grid_coefficients``: src/glass/engine/resident_cuda.py:1897-1946 -- per
  tile in float32: empty -> (1, 0); std <= eps on either side -> offset only
  (1, ref_mean - src_mean); else scale = ref_std / src_std, offset =
  ref_mean - src_mean * scale.
* ``global_coefficients``: resident_cuda.py:12978-12996 -- the same rule on the
  FP64 ``frame_global_stats`` of both frames, in FP64 (the application rounds
  the pair to float32, native_bindings.cpp:11023).
* ``apply_grid``: cpp/cuda/local_norm_kernels.cu:30-52 -- x * scale + offset
  with the tile (min(y // th, rows-1), min(x // tw, cols-1)); product and sum
  rounded separately as numpy float32 does.
* ``resident_local_normalization``: the frame loop of resident_cuda.py:
  12928-13100 (statuses, zero weight for frames without paired pixels,
  reference and zero-weight frames untouched).

Pinned by (tests/test_oracle_local_norm.py): the reference's own fixtures
(tests/test_cuda_resident_stack.py:1319-1478, tests/test_cpu_local_norm.py:
55-84, tests/test_gpu_local_norm_vs_cpu.py:56-88), brute-force per-tile
statistics with math.fsum, the explicit edge-tile partition, closed forms
(affine-related pairs map onto the reference, constant tiles are offset-only,
all-NaN tiles are empty).
"""

from __future__ import annotations

import numpy as np

from .quality import frame_global_stats

EPS = 1.0e-6


def grid_shape(height: int, width: int, tile_height: int, tile_width: int) -> tuple[int, int]:
    if tile_height <= 0 or tile_width <= 0:
        raise ValueError("tile dimensions must be positive")
    return (height + tile_height - 1) // tile_height, (width + tile_width - 1) // tile_width


def pair_grid_stats(source, reference, tile_height: int, tile_width: int) -> dict:
    src = np.asarray(source, dtype=np.float32)
    ref = np.asarray(reference, dtype=np.float32)
    if src.shape != ref.shape:
        raise ValueError("source and reference must have the same shape")
    height, width = src.shape
    rows, cols = grid_shape(height, width, tile_height, tile_width)
    out = {k: np.zeros((rows, cols), dtype=np.float32)
           for k in ("source_mean", "source_std", "reference_mean", "reference_std")}
    valid = np.zeros((rows, cols), dtype=np.uint64)
    both = np.isfinite(src) & np.isfinite(ref)
    for gy in range(rows):
        y0, y1 = gy * tile_height, min(height, (gy + 1) * tile_height)
        for gx in range(cols):
            x0, x1 = gx * tile_width, min(width, (gx + 1) * tile_width)
            m = both[y0:y1, x0:x1]
            n = int(np.count_nonzero(m))
            valid[gy, gx] = n
            if n == 0:
                continue
            s = src[y0:y1, x0:x1][m].astype(np.float64)
            r = ref[y0:y1, x0:x1][m].astype(np.float64)
            inv = 1.0 / float(n)
            s_mean = float(np.sum(s)) * inv
            r_mean = float(np.sum(r)) * inv
            s_var = max(float(np.sum(s * s)) * inv - s_mean * s_mean, 0.0)
            r_var = max(float(np.sum(r * r)) * inv - r_mean * r_mean, 0.0)
            out["source_mean"][gy, gx] = np.float32(s_mean)
            out["source_std"][gy, gx] = np.float32(np.sqrt(s_var))
            out["reference_mean"][gy, gx] = np.float32(r_mean)
            out["reference_std"][gy, gx] = np.float32(np.sqrt(r_var))
    out.update({"valid_pixels": valid, "grid_rows": rows, "grid_cols": cols,
                "valid_pixel_total": int(valid.sum(dtype=np.uint64))})
    return out


def grid_coefficients(stats: dict, eps: float = EPS) -> dict:
    sm = np.asarray(stats["source_mean"], dtype=np.float32)
    sd = np.asarray(stats["source_std"], dtype=np.float32)
    rm = np.asarray(stats["reference_mean"], dtype=np.float32)
    rd = np.asarray(stats["reference_std"], dtype=np.float32)
    valid = np.asarray(stats["valid_pixels"], dtype=np.uint64)
    scales = np.ones(sm.shape, dtype=np.float32)
    offsets = np.zeros(sm.shape, dtype=np.float32)
    statuses = []
    counts = {"empty": 0, "offset_only": 0, "ok": 0}
    for gy in range(sm.shape[0]):
        row = []
        for gx in range(sm.shape[1]):
            if int(valid[gy, gx]) == 0:
                status = "empty"
            elif float(sd[gy, gx]) <= eps or float(rd[gy, gx]) <= eps:
                status = "offset_only"
                offsets[gy, gx] = np.float32(rm[gy, gx] - sm[gy, gx])
            else:
                status = "ok"
                scale = np.float32(rd[gy, gx] / sd[gy, gx])
                scales[gy, gx] = scale
                offsets[gy, gx] = np.float32(rm[gy, gx] - np.float32(sm[gy, gx] * scale))
            counts[status] += 1
            row.append(status)
        statuses.append(row)
    active = valid > 0
    return {
        "scales": scales, "offsets": offsets, "statuses": statuses, "valid_pixels": valid,
        "empty_tiles": counts["empty"], "offset_only_tiles": counts["offset_only"], "ok_tiles": counts["ok"],
        "scale_mean": float(np.mean(scales[active])) if active.any() else 1.0,
        "offset_mean": float(np.mean(offsets[active])) if active.any() else 0.0,
        "valid_pixel_total": int(valid.sum(dtype=np.uint64)),
    }


def global_coefficients(reference_stats: dict, source_stats: dict, eps: float = EPS):
    """(status, scale, offset) in FP64 from two frame_global_stats dicts."""
    if int(source_stats["valid_pixels"]) == 0:
        return "empty", 1.0, 0.0
    rm, rs = float(reference_stats["mean"]), float(reference_stats["std"])
    sm, ss = float(source_stats["mean"]), float(source_stats["std"])
    if ss <= eps or rs <= eps:
        return "offset_only", 1.0, rm - sm
    scale = rs / ss
    return "ok", scale, rm - sm * scale


def apply_grid(frame, scales, offsets, tile_height: int, tile_width: int) -> np.ndarray:
    src = np.asarray(frame, dtype=np.float32)
    sc = np.asarray(scales, dtype=np.float32)
    of = np.asarray(offsets, dtype=np.float32)
    height, width = src.shape
    if sc.shape != grid_shape(height, width, tile_height, tile_width) or of.shape != sc.shape:
        raise ValueError("coefficient grid shape does not match frame shape and tile dimensions")
    ty = np.minimum(np.arange(height) // tile_height, sc.shape[0] - 1)
    tx = np.minimum(np.arange(width) // tile_width, sc.shape[1] - 1)
    s = sc[ty[:, None], tx[None, :]]
    o = of[ty[:, None], tx[None, :]]
    return (src * s).astype(np.float32) + o


def apply_global(frame, scale: float, offset: float) -> np.ndarray:
    src = np.asarray(frame, dtype=np.float32)
    return (src * np.float32(scale)).astype(np.float32) + np.float32(offset)


def resident_local_normalization(frames, reference_index: int, weights, mode: str = "global_mean_std",
                                 tile_size: int = 512, eps: float = EPS):
    """(normalised frames, weights, per-frame records) of the resident frame
    loop.  frames: [F, H, W] float32 (copied); weights: [F]."""
    out = np.array(frames, dtype=np.float32, copy=True)
    w = np.array(weights, dtype=np.float32, copy=True)
    records = []
    ref = out[reference_index].copy()
    ref_stats = frame_global_stats(ref) if mode == "global_mean_std" else None
    for i in range(out.shape[0]):
        rec = {"index": i, "status": "ok", "scale": 1.0, "offset": 0.0}
        if w[i] <= 0.0:
            rec["status"] = "skipped_zero_weight"
        elif i == reference_index:
            rec["status"] = "reference"
        elif mode == "global_mean_std":
            status, scale, offset = global_coefficients(ref_stats, frame_global_stats(out[i]), eps)
            rec.update(status=status, scale=scale, offset=offset)
            if status == "empty":
                w[i] = 0.0
            else:
                out[i] = apply_global(out[i], scale, offset)
        else:
            stats = pair_grid_stats(out[i], ref, tile_size, tile_size)
            coeffs = grid_coefficients(stats, eps)
            rec.update(scale=coeffs["scale_mean"], offset=coeffs["offset_mean"], scales=coeffs["scales"],
                       offsets=coeffs["offsets"], valid_pixels=coeffs["valid_pixels"])
            if coeffs["valid_pixel_total"] == 0:
                rec["status"] = "empty"
                w[i] = 0.0
            else:
                out[i] = apply_grid(out[i], coeffs["scales"], coeffs["offsets"], tile_size, tile_size)
                if coeffs["empty_tiles"]:
                    rec["status"] = "partial"
        records.append(rec)
    return out, w, records
