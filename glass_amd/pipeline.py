"""Host driver of the resident stacking path (masters -> calibrate -> warp ->
weights -> reject/integrate -> DQ), mirroring the stage order of the
reference's ``run_resident_calibration_integration`` (src/glass/engine/
resident_cuda.py:7119) on top of ``ResidentCalibratedStack``.

Semantics follow the CPU path, which is the oracle:
* masters: src/glass/engine/pipeline.py:517-1047 -- winsorized mean stack with
  the master rejection policy, per-tile statistics choice; darks minus bias
  when they exclude it; flats reduced by the bias master, divided by their own
  exact median, stacked, floored;
* light calibration: src/glass/cpu/calibration.py:51-76;
* warp: identity matrices skip resampling (resident_cuda.py:2751-2754), others
  are resampled on the reference grid; geometric coverage is accumulated;
* weights: none, variance_aware (engine/integration.py:80-106) or the resident
  simple_snr rule (resident_cuda.py:1948-1955), zero weight for masked frames;
* optional resident local normalisation (glass_amd.local_norm);
* integration: cpu/integration.py:36-118 and the DQ map (resident_cuda.py:
  3110-3297) in one launch.
All arithmetic runs in libglass_hip.so; numpy only shapes host inputs/outputs.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from .local_norm import MODES as LN_MODES, resident_local_normalization
from .resident import ResidentCalibratedStack

DEFAULT_POLICY = {
    "master_dark_includes_bias": True,
    "dark_scaling_enabled": True,
    "flat_floor": 1.0e-6,
    "pedestal": 0.0,
    "master_rejection_low_sigma": 3.0,
    "master_rejection_high_sigma": 3.0,
    "master_rejection_min_samples": 3,
    "master_rejection_max_fraction": 0.5,
    "tile_size": 512,
}


def _policy(policy: dict | None) -> dict:
    out = dict(DEFAULT_POLICY)
    out.update(policy or {})
    return out


def _stack_frames(frames) -> ResidentCalibratedStack:
    frames = list(frames)
    if not frames:
        raise ValueError("cannot stack an empty frame list")
    h, w = np.asarray(frames[0]).shape if not hasattr(frames[0], "shape") else tuple(frames[0].shape)
    stack = ResidentCalibratedStack(len(frames), int(h), int(w))
    for i, frame in enumerate(frames):
        stack.upload_calibrated_frame(i, frame)
    return stack


def _master_integrate(stack: ResidentCalibratedStack, pol: dict) -> np.ndarray:
    master, *_ = stack.integrate_stack_engine(
        low_sigma=float(pol["master_rejection_low_sigma"]),
        high_sigma=float(pol["master_rejection_high_sigma"]),
        min_samples=int(pol["master_rejection_min_samples"]),
        max_reject_fraction=float(pol["master_rejection_max_fraction"]),
        tile_size=int(pol["tile_size"]),
    )
    return master


def build_master_bias(frames, policy: dict | None = None) -> np.ndarray:
    """pipeline.py:1186-1200 -> _stack_mean_master (:647, :529-610)."""
    pol = _policy(policy)
    with _stack_frames(frames) as stack:
        return _master_integrate(stack, pol)


def build_master_dark(frames, master_bias=None, policy: dict | None = None) -> np.ndarray:
    """pipeline.py:1222-1246: darks stacked, then minus the bias master when
    the darks do not include bias (f32, :589-593)."""
    pol = _policy(policy)
    with _stack_frames(frames) as stack:
        if not pol["master_dark_includes_bias"] and master_bias is not None:
            master = stack_minus_master(stack, pol, master_bias)
        else:
            master = _master_integrate(stack, pol)
    return master


def stack_minus_master(stack: ResidentCalibratedStack, pol: dict, subtract) -> np.ndarray:
    master = _master_integrate(stack, pol)
    # the subtraction of the stacked dark and the bias master runs on device
    with ResidentCalibratedStack(1, stack.height, stack.width) as tmp:
        tmp.set_calibration_masters(bias=subtract)
        tmp.calibrate_frame(0, master, 0.0, None, {"master_dark_includes_bias": False})
        return tmp.download_frame(0)


def build_master_flat(frames, master_bias=None, policy: dict | None = None):
    """pipeline.py:828-955: per flat (flat - bias) / exact median (FP64
    partition median, :721-775), stacked with the master rejection policy,
    then floored.  Returns (master_flat, per-flat scalars)."""
    pol = _policy(policy)
    frames = list(frames)
    h, w = np.asarray(frames[0]).shape
    floor = float(pol["flat_floor"])
    with ResidentCalibratedStack(len(frames), h, w) as stack:
        stack.set_calibration_masters(bias=master_bias)
        for i, frame in enumerate(frames):
            stack.calibrate_frame(i, frame, 0.0, None, {"master_dark_includes_bias": True})
        scalars = []
        for i in range(len(frames)):
            median = stack.frame_global_stats(i, exact_median=True)["median"]
            if abs(median) < floor:
                raise ValueError("flat normalization is below flat_floor")
            scalars.append(float(median))
        stack.normalize_frames(list(range(len(frames))), np.zeros(len(frames)), np.asarray(scalars, np.float32))
        master = _master_integrate(stack, pol)
    return np.maximum(master, np.float32(floor)).astype(np.float32), scalars


def build_masters_device(bias_raw, dark_raw, flat_raw, height: int, width: int, policy: dict | None = None,
                         raw_format: str = "fits_u16be_bzero") -> dict:
    """Masters from raw frames already in HBM (torch tensors [n, H*W]), same
    semantics as build_master_* but without a host round trip; the masters
    stay on the device as torch float32 tensors."""
    import torch

    pol = _policy(policy)
    device = bias_raw.device
    out: dict = {}

    def stack_of(raw):
        n = int(raw.shape[0])
        s = ResidentCalibratedStack(n, height, width)
        s.calibrate_frames_device(list(range(n)), raw, raw_format, [0.0] * n, None, pol)
        return s

    def integrate(s):
        master = torch.empty((height, width), dtype=torch.float32, device=device)
        s.integrate_stack_device(master, low_sigma=float(pol["master_rejection_low_sigma"]),
                                 high_sigma=float(pol["master_rejection_high_sigma"]),
                                 min_samples=int(pol["master_rejection_min_samples"]),
                                 max_reject_fraction=float(pol["master_rejection_max_fraction"]),
                                 tile_size=int(pol["tile_size"]))
        return master

    with stack_of(bias_raw) as s:
        out["bias"] = integrate(s)
    with stack_of(dark_raw) as s:
        dark = integrate(s)
    if not pol["master_dark_includes_bias"]:
        with ResidentCalibratedStack(1, height, width) as tmp:
            tmp.upload_calibrated_frame(0, dark)
            tmp.set_calibration_masters(bias=out["bias"])
            tmp.calibrate_frames_device([0], dark, "float32", [0.0], None, {"master_dark_includes_bias": False})
            dark = torch.from_numpy(tmp.download_frame(0)).to(device)
    out["dark"] = dark
    n = int(flat_raw.shape[0])
    with ResidentCalibratedStack(n, height, width) as s:
        s.set_calibration_masters(bias=out["bias"])
        s.calibrate_frames_device(list(range(n)), flat_raw, raw_format, [0.0] * n, None, pol)
        scalars = [s.frame_global_stats(i, exact_median=True)["median"] for i in range(n)]
        s.normalize_frames(list(range(n)), np.zeros(n), np.asarray(scalars, np.float32))
        flat = integrate(s)
    out["flat"] = torch.clamp_min(flat, float(np.float32(pol["flat_floor"])))
    out["flat_scalars"] = scalars
    return out


def variance_aware_weights(rms: Sequence[float], active: Sequence[bool] | None = None) -> np.ndarray:
    """engine/integration.py:80-106 ``variance_aware`` from per-frame rms
    (``background_rms``, quality.py:294): 1/max(rms^2, 1e-12), floored at 1e-6,
    divided by the f32 median of the positive raw weights; inactive frames get
    weight 0 (resident_cuda.py:12759-12790)."""
    raw = [max(1.0 / max(float(r) ** 2, 1.0e-12), 1.0e-6) for r in rms]
    positive = np.asarray([v for v in raw if np.isfinite(v) and v > 0.0], dtype=np.float32)
    scale = max(float(np.median(positive)) if positive.size else 1.0, 1.0e-6)
    weights = np.asarray([max(v / scale, 1.0e-6) for v in raw], dtype=np.float32)
    if active is not None:
        weights = np.where(np.asarray(active, dtype=bool), weights, np.float32(0.0)).astype(np.float32)
    return weights


def simple_snr_weights(stats: Sequence[dict], active: Sequence[bool] | None = None,
                       eps: float = 1.0e-6) -> np.ndarray:
    """The resident ``simple_snr`` rule (``_simple_snr_weight_from_stats``,
    resident_cuda.py:1948-1955, frame loop :12762-12779) on frame_global_stats
    of the calibrated, unwarped frames: 0 without valid pixels, 1/eps when
    std <= eps, else max(|mean|/std, eps); inactive frames keep weight 0 and
    their statistics are not read."""
    out = np.zeros(len(stats), dtype=np.float32)
    for i, st in enumerate(stats):
        if active is not None and not bool(active[i]):
            continue
        if int(st.get("valid_pixels") or 0) <= 0:
            continue
        std = float(st.get("std") or 0.0)
        mean = abs(float(st.get("mean") or 0.0))
        out[i] = 1.0 / eps if std <= eps else max(mean / std, eps)
    return out


def _is_identity(matrix) -> bool:
    return bool(np.allclose(np.asarray(matrix, dtype=np.float64), np.eye(3), atol=0.0, rtol=0.0))


@dataclass
class StackResult:
    master: np.ndarray
    weight: np.ndarray
    coverage: np.ndarray
    low: np.ndarray
    high: np.ndarray
    dq: np.ndarray
    geometric: np.ndarray
    weights: np.ndarray
    timing: dict = field(default_factory=dict)


def run_resident_stack(
    lights,
    light_exposures,
    masters: dict,
    matrices,
    *,
    dark_exposure_s=None,
    interpolation: str = "lanczos3",
    weighting: str = "none",
    active=None,
    low_sigma: float = 3.0,
    high_sigma: float = 3.0,
    min_samples: int = 3,
    max_reject_fraction: float = 0.5,
    policy: dict | None = None,
    raw_format: str = "float32",
    warp_semantics: str = "cpu",
    local_normalization: str = "off",
    local_normalization_tile_size: int = 512,
    reference_index: int = 0,
) -> StackResult:
    """One resident stacking job from host frames (the drop-in driver).

    ``local_normalization``: "off", "global_mean_std" or "grid_mean_std" --
    frames are normalised in place to frame ``reference_index`` after the
    weights and before integration (resident_cuda.py:12902-13110); frames left
    without pixels get weight 0."""
    if local_normalization not in ("off", *LN_MODES):
        raise ValueError("local_normalization must be off, global_mean_std or grid_mean_std")
    pol = _policy(policy)
    lights = list(lights)
    f = len(lights)
    first = np.asarray(lights[0])
    if raw_format == "fits_u16be_bzero":
        raise ValueError("pass decoded frames or use calibrate_frames_* for raw bytes")
    h, w = first.shape
    t0 = time.perf_counter()
    stack = ResidentCalibratedStack(f, h, w)
    try:
        stack.set_calibration_masters(masters.get("bias"), masters.get("dark"), masters.get("flat"))
        for i, light in enumerate(lights):
            stack.calibrate_frame(i, light, light_exposures[i], dark_exposure_s, pol)
        t_cal = time.perf_counter()
        stats = [stack.frame_global_stats(i) for i in range(f)] if weighting != "none" else None
        stack.reset_warp_coverage()
        moving = [i for i in range(f) if not _is_identity(matrices[i])]
        for i in range(f):
            if i not in moving:
                stack.accumulate_full_warp_coverage_frame()
        if moving:
            mats = np.stack([np.asarray(matrices[i], dtype=np.float64) for i in moving])
            if warp_semantics == "cpu":
                stack.apply_matrix_frames_cpu_semantics(moving, mats, interpolation, fill=np.nan)
            elif interpolation == "lanczos3":
                stack.apply_matrix_lanczos3_frames(moving, mats)
            else:
                stack.apply_matrix_bilinear_frames(moving, mats)
        t_warp = time.perf_counter()
        if weighting == "none":
            weights = np.ones(f, dtype=np.float32)
            if active is not None:
                weights = np.where(np.asarray(active, bool), weights, np.float32(0.0)).astype(np.float32)
        elif weighting == "variance_aware":
            weights = variance_aware_weights([s["std"] for s in stats], active)
        elif weighting == "simple_snr":
            weights = simple_snr_weights(stats, active)
        else:
            raise ValueError(f"unsupported weighting: {weighting}")
        timing_ln = 0.0
        if local_normalization != "off":
            t_ln = time.perf_counter()
            ln = resident_local_normalization(stack, reference_index, weights, local_normalization,
                                              local_normalization_tile_size)
            weights = ln["weights"]
            timing_ln = time.perf_counter() - t_ln
        maps = integrate_with_dq(stack, weights, low_sigma, high_sigma, min_samples, max_reject_fraction)
        geometric = stack.warp_coverage_map()
        t_int = time.perf_counter()
        return StackResult(maps["master"], maps["weight"], maps["coverage"], maps["low"], maps["high"],
                           maps["dq"], geometric, weights,
                           {"calibrate_s": t_cal - t0, "warp_s": t_warp - t_cal, "local_norm_s": timing_ln,
                            "integrate_s": t_int - t_warp - timing_ln})
    finally:
        stack.close()


def write_resident_outputs(result: StackResult, output_dir, filter_name: str, frame_count: int | None = None,
                           max_workers: int | None = None) -> dict:
    """Write the output maps as simple FITS files the way the resident run
    does (resident_cuda.py:13464-13662, writers :4798-4848): master and weight
    float32, DQ int16, coverage / low / high counts rounded with np.rint to
    int16 (int32 beyond 32767 frames); one worker thread per file."""
    from concurrent.futures import ThreadPoolExecutor
    from pathlib import Path

    from . import fits as F

    out = Path(output_dir)
    out.mkdir(parents=True, exist_ok=True)
    n = int(frame_count if frame_count is not None else len(result.weights))
    count_dtype = np.int16 if n <= np.iinfo(np.int16).max else np.int32
    filt = str(filter_name)
    dtype_name = np.dtype(count_dtype).name
    specs = [
        ("master", out / f"resident_master_{filt}.fits", result.master, {"IMAGETYP": "master", "FILTER": filt},
         np.float32, False),
        ("weight", out / f"resident_weight_map_{filt}.fits", result.weight, {"IMAGETYP": "weight", "FILTER": filt},
         np.float32, False),
        ("dq", out / f"resident_dq_map_{filt}.fits", result.dq,
         {"IMAGETYP": "dq_mask", "FILTER": filt, "DQSTAGE": "integration",
          "DQFLAGS": "NO_DATA,WARP_EDGE,LOW_REJECTED,HIGH_REJECTED"}, np.int16, False),
        ("coverage", out / f"resident_coverage_map_{filt}.fits", result.coverage,
         {"IMAGETYP": "coverage", "FILTER": filt, "DTYPE": dtype_name}, count_dtype, True),
        ("low_rejection", out / f"resident_low_rejection_map_{filt}.fits", result.low,
         {"IMAGETYP": "rej_low", "FILTER": filt, "DTYPE": dtype_name}, count_dtype, True),
        ("high_rejection", out / f"resident_high_rejection_map_{filt}.fits", result.high,
         {"IMAGETYP": "rej_high", "FILTER": filt, "DTYPE": dtype_name}, count_dtype, True),
    ]

    def write(spec):
        name, path, data, header, dtype, counts = spec
        t = time.perf_counter()
        F.write_fits_data(path, F.count_map_for_write(data, dtype) if counts else data, header, dtype=dtype)
        return name, str(path), time.perf_counter() - t

    t0 = time.perf_counter()
    workers = max(1, min(len(specs), int(max_workers or len(specs))))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        done = list(pool.map(write, specs))
    return {"paths": {name: path for name, path, _ in done}, "breakdown_s": {name: s for name, _, s in done},
            "total_s": time.perf_counter() - t0, "workers": workers}


def integrate_with_dq(stack: ResidentCalibratedStack, weights, low_sigma=3.0, high_sigma=3.0, min_samples=3,
                      max_reject_fraction=0.5) -> dict:
    """One winsorized integration launch that returns the five maps of
    ``integrate_hardened_winsorized_sigma`` and the resident DQ map
    (resident_cuda.py:3110-3297, computed by the same kernel from its own
    counts, the geometric warp coverage and the active frame count)."""
    from . import _native as N

    if not (float(low_sigma) > 0 and float(high_sigma) > 0):
        raise ValueError("low_sigma and high_sigma must be positive")
    if int(min_samples) < 1:
        raise ValueError("min_samples must be at least 1")
    if not 0.0 <= float(max_reject_fraction) <= 1.0:
        raise ValueError("max_reject_fraction must be between 0 and 1")
    w = np.asarray(weights, dtype=np.float32)
    active = int(np.count_nonzero(np.isfinite(w) & (w > 0)))
    params = N.IntegrateParams(N.GH_REJECT_WINSORIZED_SIGMA, float(low_sigma), float(high_sigma), int(min_samples),
                               float(max_reject_fraction), N.GH_COUNT_F32, -1, 0, 0)
    return stack._integrate(w, params, {"master", "weight", "coverage", "low", "high"}, N.GH_COUNT_F32, dq=True,
                            dq_geometric=True, dq_active_frames=active)
