"""Resident local normalisation driver (between warp and integration).

Mirrors the frame loop of the reference's resident run
(src/glass/engine/resident_cuda.py:12902-13110) on ``ResidentCalibratedStack``:

* ``global_mean_std``: one (scale, offset) per frame from the FP64
  ``frame_global_stats`` of the frame and of the reference frame (one batched
  statistics pass, one batched application);
* ``grid_mean_std``: per-tile coefficients from ``frame_pair_grid_stats_batch``
  (one statistics pass for every positive-weight source), then one batched
  in-place application (``apply_grid_normalization_frames``).

Coefficient rule (resident_cuda.py:1897-1946, in float32 per tile): a tile
without paired finite pixels keeps (1, 0); a tile whose source or reference
standard deviation is <= eps gets offset only, ref_mean - src_mean; otherwise
scale = ref_std / src_std and offset = ref_mean - src_mean * scale.  Frames
with no finite pixels (global) or no paired pixels (grid) get weight 0.
The arithmetic on pixels runs in libglass_hip.so; numpy handles only the
per-tile coefficient grids.
"""

from __future__ import annotations

import time
from typing import Any, Sequence

import numpy as np

from .resident import ResidentCalibratedStack

EPS = 1.0e-6
MODES = ("global_mean_std", "grid_mean_std")


def _coefficient_grids(sm, sd, rm, rd, valid, eps):
    """Elementwise coefficient rule over tiles of any array shape: (scales,
    offsets, ok mask, offset-only mask, active mask)."""
    active = valid > 0
    flat = (sd.astype(np.float64) <= eps) | (rd.astype(np.float64) <= eps)
    ok = active & ~flat
    offset_only = active & flat
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = (rd / sd).astype(np.float32)
    scales = np.where(ok, ratio, np.float32(1.0)).astype(np.float32)
    offsets = np.zeros(sm.shape, dtype=np.float32)
    offsets[offset_only] = (rm[offset_only] - sm[offset_only]).astype(np.float32)
    offsets[ok] = (rm[ok] - (sm[ok] * scales[ok]).astype(np.float32)).astype(np.float32)
    return scales, offsets, ok, offset_only, active


def _summary(scales, offsets, valid, ok, offset_only, active, statuses) -> dict[str, Any]:
    a_s, a_o = scales[active], offsets[active]
    return {
        "scales": scales,
        "offsets": offsets,
        "valid_pixels": valid,
        "statuses": statuses,
        "empty_tiles": int(np.count_nonzero(~active)),
        "offset_only_tiles": int(np.count_nonzero(offset_only)),
        "ok_tiles": int(np.count_nonzero(ok)),
        "scale_mean": float(np.mean(a_s)) if a_s.size else 1.0,
        "scale_min": float(np.min(a_s)) if a_s.size else 1.0,
        "scale_max": float(np.max(a_s)) if a_s.size else 1.0,
        "offset_mean": float(np.mean(a_o)) if a_o.size else 0.0,
        "offset_min": float(np.min(a_o)) if a_o.size else 0.0,
        "offset_max": float(np.max(a_o)) if a_o.size else 0.0,
        "valid_pixel_total": int(np.sum(valid, dtype=np.uint64)),
    }


def _status_names(ok, offset_only):
    return np.where(ok, "ok", np.where(offset_only, "offset_only", "empty")).tolist()


def grid_local_norm_coefficients(stats: dict, eps: float = EPS) -> dict[str, Any]:
    """Per-tile coefficients of one frame_pair_grid_stats result
    (resident_cuda.py:1897-1946), vectorised over the tiles."""
    f32 = {k: np.asarray(stats[k], dtype=np.float32)
           for k in ("source_mean", "source_std", "reference_mean", "reference_std")}
    valid = np.asarray(stats["valid_pixels"], dtype=np.uint64)
    scales, offsets, ok, offset_only, active = _coefficient_grids(
        f32["source_mean"], f32["source_std"], f32["reference_mean"], f32["reference_std"], valid, eps)
    return _summary(scales, offsets, valid, ok, offset_only, active, _status_names(ok, offset_only))


def _grid_coefficients_batch(frames: list[dict], eps: float) -> list[dict[str, Any]]:
    """grid_local_norm_coefficients of every frame of a batch, with the
    elementwise work and the exact reductions (counts, min, max) done once
    over the stacked grids; the float32 means stay per frame (np.mean of the
    active tiles, as the single-frame rule computes them)."""
    if not frames:
        return []
    st = {k: np.stack([np.asarray(fr[k], dtype=np.float32) for fr in frames])
          for k in ("source_mean", "source_std", "reference_mean", "reference_std")}
    valid = np.stack([np.asarray(fr["valid_pixels"], dtype=np.uint64) for fr in frames])
    scales, offsets, ok, offset_only, active = _coefficient_grids(
        st["source_mean"], st["source_std"], st["reference_mean"], st["reference_std"], valid, eps)
    names = _status_names(ok, offset_only)
    axes = (1, 2)
    n_active = np.count_nonzero(active, axis=axes)
    n_empty = np.count_nonzero(~active, axis=axes)
    n_off = np.count_nonzero(offset_only, axis=axes)
    n_ok = np.count_nonzero(ok, axis=axes)
    s_min = np.where(active, scales, np.float32(np.inf)).min(axis=axes)
    s_max = np.where(active, scales, np.float32(-np.inf)).max(axis=axes)
    o_min = np.where(active, offsets, np.float32(np.inf)).min(axis=axes)
    o_max = np.where(active, offsets, np.float32(-np.inf)).max(axis=axes)
    totals = np.sum(valid, axis=axes, dtype=np.uint64)
    out = []
    for k in range(len(frames)):
        has = bool(n_active[k])
        out.append({
            "scales": scales[k], "offsets": offsets[k], "valid_pixels": valid[k], "statuses": names[k],
            "empty_tiles": int(n_empty[k]), "offset_only_tiles": int(n_off[k]), "ok_tiles": int(n_ok[k]),
            "scale_mean": float(np.mean(scales[k][active[k]])) if has else 1.0,
            "scale_min": float(s_min[k]) if has else 1.0,
            "scale_max": float(s_max[k]) if has else 1.0,
            "offset_mean": float(np.mean(offsets[k][active[k]])) if has else 0.0,
            "offset_min": float(o_min[k]) if has else 0.0,
            "offset_max": float(o_max[k]) if has else 0.0,
            "valid_pixel_total": int(totals[k]),
        })
    return out


def resident_local_normalization(stack: ResidentCalibratedStack, reference_index: int, weights: Sequence[float],
                                 mode: str = "global_mean_std", tile_size: int = 512, eps: float = EPS,
                                 frame_ids: Sequence[str] | None = None) -> dict[str, Any]:
    """Normalise every positive-weight frame of ``stack`` to the reference
    frame in place.  Returns {"weights": float32[F] (zero for frames left
    without pixels), "frames": per-frame records, "mode", "timing"}."""
    if mode not in MODES:
        raise ValueError("resident_local_normalization_mode must be global_mean_std or grid_mean_std")
    if int(tile_size) <= 0:
        raise ValueError("resident_local_normalization_tile_size must be positive")
    f = stack.frame_count
    w = np.array(weights, dtype=np.float32, copy=True).reshape(-1)
    if w.size != f:
        raise ValueError("weights must have one entry per frame")
    ids = [str(i) for i in range(f)] if frame_ids is None else [str(x) for x in frame_ids]
    t0 = time.perf_counter()
    grid_stats: dict[int, dict] = {}
    grid_coeffs: dict[int, dict] = {}
    reference_mean = reference_std = 0.0
    reference_stats = None
    global_stats: dict[int, dict] = {}
    if mode == "global_mean_std":
        # one batched statistics pass over the reference and every source
        sources = [i for i in range(f) if float(w[i]) > 0.0 and i != reference_index]
        listed = [reference_index, *sources]
        global_stats = dict(zip(listed, stack.frames_global_stats(listed)))
        reference_stats = global_stats[reference_index]
        reference_mean = float(reference_stats["mean"])
        reference_std = float(reference_stats["std"])
    else:
        sources = [i for i in range(f) if float(w[i]) > 0.0 and i != reference_index]
        if sources:
            batch = stack.frame_pair_grid_stats_batch(reference_index, np.asarray(sources, dtype=np.int32),
                                                      tile_size, tile_size)
            grid_stats = {int(item["source_index"]): item for item in batch["frames"]}
            grid_coeffs = dict(zip(grid_stats, _grid_coefficients_batch(batch["frames"], eps)))
    t_stats = time.perf_counter()
    records: list[dict[str, Any]] = []
    pending: list[tuple[int, np.ndarray, np.ndarray]] = []
    pending_global: list[tuple[int, float, float]] = []
    for i in range(f):
        rec: dict[str, Any] = {"frame_id": ids[i], "reference_frame_id": ids[reference_index],
                               "model": f"resident_{mode}", "scale": 1.0, "offset": 0.0, "status": "ok",
                               "warnings": [], "reference_mean": reference_mean, "reference_std": reference_std}
        if float(w[i]) <= 0.0:
            rec["status"] = "skipped_zero_weight"
            rec["warnings"].append("frame was excluded or failed registration before local normalization")
        elif i == reference_index:
            rec["status"] = "reference"
        elif mode == "global_mean_std":
            src = global_stats[i]
            rec.update(source_mean=float(src["mean"]), source_std=float(src["std"]),
                       valid_pixels=int(src["valid_pixels"]))
            if int(src["valid_pixels"]) == 0:
                rec["status"] = "empty"
                w[i] = 0.0
                rec["warnings"].append("frame had no finite pixels for resident global LN")
            else:
                if float(src["std"]) <= eps or reference_std <= eps:
                    rec["status"] = "offset_only"
                    scale, offset = 1.0, reference_mean - float(src["mean"])
                else:
                    scale = reference_std / float(src["std"])
                    offset = reference_mean - float(src["mean"]) * scale
                rec.update(scale=scale, offset=offset)
                pending_global.append((i, scale, offset))
        else:
            coeffs = grid_coeffs[i]
            if coeffs["valid_pixel_total"] == 0:
                rec["status"] = "empty"
                w[i] = 0.0
                rec["warnings"].append("frame had no finite paired pixels for resident grid LN")
            else:
                pending.append((i, coeffs["scales"], coeffs["offsets"]))
                if coeffs["empty_tiles"]:
                    rec["status"] = "partial"
                    rec["warnings"].append(
                        f"{coeffs['empty_tiles']} local-normalization tiles had no paired finite pixels")
            rec.update(scale=coeffs["scale_mean"], offset=coeffs["offset_mean"], grid_coefficients={
                "model": "resident_grid_pair_mean_std", "tile_size": int(tile_size),
                "grid_rows": int(grid_stats[i]["grid_rows"]), "grid_cols": int(grid_stats[i]["grid_cols"]),
                **{k: coeffs[k] for k in ("scale_mean", "scale_min", "scale_max", "offset_mean", "offset_min",
                                          "offset_max", "valid_pixel_total", "empty_tiles", "offset_only_tiles",
                                          "ok_tiles", "scales", "offsets", "valid_pixels", "statuses")},
            })
        records.append(rec)
    if pending_global:
        stack.apply_global_normalization_frames([p[0] for p in pending_global], [p[1] for p in pending_global],
                                                [p[2] for p in pending_global])
    if pending:
        stack.apply_grid_normalization_frames(
            np.asarray([p[0] for p in pending], dtype=np.int32),
            np.stack([p[1] for p in pending]).astype(np.float32, copy=False),
            np.stack([p[2] for p in pending]).astype(np.float32, copy=False),
            tile_size, tile_size)
    t_end = time.perf_counter()
    return {"mode": f"resident_{mode}", "weights": w, "frames": records,
            "timing": {"stats_s": t_stats - t0, "apply_s": t_end - t_stats}}
