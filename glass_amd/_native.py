"""ctypes binding of libglass_hip.so (the C ABI in include/glass_hip.h).

The library is loaded from glass_amd/lib/ (built in-tree by glass_amd.build).
There is no fallback: if the library is missing or fails to load, every entry
point raises, so nothing silently runs on the CPU.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_size_t, c_uint8, c_uint32, c_uint64, c_void_p
from pathlib import Path

LIBRARY_PATH = Path(__file__).resolve().parent / "lib" / "libglass_hip.so"

GH_OK = 0
GH_ERR_VALUE = 1
GH_ERR_INDEX = 2
GH_ERR_RUNTIME = 3

GH_FMT_F32 = 0
GH_FMT_U16 = 1
GH_FMT_FITS_U16BE_BZERO = 2

GH_INTERP_BILINEAR = 1
GH_INTERP_LANCZOS3 = 3

GH_TRANSLATE_INTEGER = 0
GH_WARP_TRACK_COVERAGE = 1
GH_WARP_SKIP_NONFINITE_TAPS = 2
GH_WARP_AFFINE = 4

GH_REJECT_NONE = 0
GH_REJECT_WINSORIZED_SIGMA = 1
GH_REJECT_SIGMA_CLIP = 2
GH_CLIP_STACK = 0
GH_CANDIDATES_ALL, GH_CANDIDATES_TOP, GH_CANDIDATES_TOP_NMS, GH_CANDIDATES_GRID_NMS = 0, 1, 2, 3

GH_COUNT_F32 = 0
GH_COUNT_U16 = 1

GH_MODEL_TRANSLATION = 0
GH_MODEL_SIMILARITY = 1
GH_MODEL_AFFINE = 2


class CalibPolicy(ctypes.Structure):
    _fields_ = [
        ("master_dark_includes_bias", c_int32),
        ("flat_floor", c_float),
        ("pedestal", c_float),
        ("accumulate_stats", c_int32),
    ]


class IntegrateParams(ctypes.Structure):
    _fields_ = [
        ("rejection", c_int32),
        ("low_sigma", c_float),
        ("high_sigma", c_float),
        ("min_samples", c_int32),
        ("max_reject_fraction", c_float),
        ("count_dtype", c_int32),
        ("statistics_variant", c_int32),
        ("require_positive_scale", c_int32),
        ("tile_size", c_int32),
    ]


class IntegrateOutputs(ctypes.Structure):
    _fields_ = [
        ("master", c_void_p),
        ("weight", c_void_p),
        ("coverage", c_void_p),
        ("low", c_void_p),
        ("high", c_void_p),
        ("dq", c_void_p),
        ("dq_use_geometric", c_int32),
        ("dq_active_frames", c_int32),
    ]


class ClipParams(ctypes.Structure):
    _fields_ = [
        ("sampler", c_int32),
        ("rejection", c_int32),
        ("winsorize", c_int32),
        ("low_sigma", c_float),
        ("high_sigma", c_float),
        ("clamping_threshold", c_float),
    ]


class CandidateParams(ctypes.Structure):
    _fields_ = [
        ("mode", c_int32),
        ("threshold", c_float),
        ("max_candidates", c_int32),
        ("max_output_candidates", c_int32),
        ("grid_cols", c_int32),
        ("grid_rows", c_int32),
        ("candidates_per_cell", c_int32),
        ("min_separation_px", c_float),
        ("centroid_radius", c_int32),
        ("centroid_global_mean_background", c_int32),
    ]


class CandidateResult(ctypes.Structure):
    _fields_ = [
        ("count", c_int64),
        ("stored_count", c_int64),
        ("selected_count", c_int64),
        ("refined_count", c_int64),
        ("centroid_background", c_float),
        ("max_shift_px", c_float),
    ]


class StarTranslationParams(ctypes.Structure):
    _fields_ = [
        ("threshold", c_float),
        ("max_candidates", c_int32),
        ("tolerance_px", c_float),
        ("max_abs_dx", c_float),
        ("max_abs_dy", c_float),
        ("prior_dx", c_float),
        ("prior_dy", c_float),
        ("prior_radius_px", c_float),
        ("grid_cols", c_int32),
        ("grid_rows", c_int32),
    ]


class StarTranslationResult(ctypes.Structure):
    _fields_ = [(name, c_float) for name in ("dx", "dy", "refined_dx", "refined_dy", "rms_px", "max_abs_dx",
                                             "max_abs_dy", "prior_radius_px")] + \
        [(name, c_int64) for name in ("inliers", "mutual_inliers", "reference_count", "moving_count",
                                      "reference_total_count", "moving_total_count")]


class AlignSearchParams(ctypes.Structure):
    _fields_ = [("search_radius_px", c_float), ("coarse_step_px", c_float), ("fine_radius_px", c_float),
                ("fine_step_px", c_float), ("coarse_sample_stride", c_int32), ("final_sample_stride", c_int32)]


class AlignCandidate(ctypes.Structure):
    _fields_ = [("matrix", c_float * 9), ("dx", c_float), ("dy", c_float), ("valid_pixels", c_int64),
                ("sampled_pixels", c_int64), ("sample_stride", c_int32), ("rms", c_double),
                ("mean_abs_diff", c_double), ("ncc", c_double), ("coarse_candidates", c_int64),
                ("fine_candidates", c_int64)]


class FrameStatistics(ctypes.Structure):
    _fields_ = [
        ("mean", c_double),
        ("std", c_double),
        ("sum", c_double),
        ("sum_sq", c_double),
        ("valid_pixels", c_int64),
        ("total_pixels", c_int64),
        ("nonfinite_pixels", c_int64),
        ("median", c_double),
        ("mad", c_double),
        ("min", c_float),
        ("max", c_float),
    ]


class StarRecord(ctypes.Structure):
    _fields_ = [
        ("x", c_double),
        ("y", c_double),
        ("flux", c_double),
        ("peak", c_double),
        ("fwhm_px", c_double),
        ("eccentricity", c_double),
        ("snr", c_double),
        ("sharpness", c_double),
        ("npix", c_int64),
        ("has_fwhm", c_int32),
        ("has_eccentricity", c_int32),
    ]


class BatchJob(ctypes.Structure):
    _fields_ = [
        ("stack", c_void_p),
        ("warp_indices", POINTER(c_int64)),
        ("inverse", POINTER(c_double)),
        ("n_warp", c_int64),
        ("interpolation", c_int32),
        ("warp_flags", c_uint32),
        ("fill", c_float),
        ("clamping_threshold", c_float),
        ("weights", POINTER(c_float)),
        ("params", IntegrateParams),
        ("outputs", IntegrateOutputs),
    ]


class DeviceProperties(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("memory_total_bytes", c_int64),
        ("multiprocessor_count", c_int32),
        ("major", c_int32),
        ("minor", c_int32),
        ("wavefront_size", c_int32),
    ]


class TriangleInfo(ctypes.Structure):
    _fields_ = [("count", c_int64), ("raw_count", c_int64), ("neighbors", c_int32), ("reserved", c_int32)]


class SimilarityFit(ctypes.Structure):
    _fields_ = [
        ("matrix", c_float * 9),
        ("scale", c_float),
        ("rotation_rad", c_float),
        ("rms_px", c_float),
        ("inliers", c_int32),
        ("best_candidate_index", c_int64),
        ("candidate_count", c_int64),
        ("scored_candidates", c_int64),
        ("refined_inliers", c_int32),
        ("refit_status", c_int32),
        ("refit_rms_px", c_float),
        ("reserved", c_int32),
    ]


class CatalogSimilarityParams(ctypes.Structure):
    _fields_ = [
        ("tolerance_px", c_float),
        ("min_pair_distance", c_float),
        ("prior_dx", c_float),
        ("prior_dy", c_float),
        ("prior_radius_px", c_float),
        ("min_scale", c_float),
        ("max_scale", c_float),
        ("max_abs_rotation_rad", c_float),
        ("top_k", c_int32),
    ]


class SimilarityCandidate(ctypes.Structure):
    _fields_ = [
        ("candidate_index", c_int64),
        ("inliers", c_int32),
        ("rms_px", c_float),
        ("a", c_float),
        ("b", c_float),
        ("tx", c_float),
        ("ty", c_float),
    ]


GH_BATCH_USE_GRAPH = 1
GH_BATCH_AUTOTUNE = 2

# name -> (argtypes); every function returns int
_SIGNATURES = {
    "gh_last_error": [c_char_p, c_size_t],
    "gh_version": [],
    "gh_device_count": [POINTER(c_int)],
    "gh_device_info": [c_int, POINTER(DeviceProperties)],
    "gh_host_alloc": [c_size_t, POINTER(c_void_p)],
    "gh_host_free": [c_void_p],
    "gh_stack_create": [c_int64, c_int64, c_int64, POINTER(c_void_p)],
    "gh_stack_destroy": [c_void_p],
    "gh_stack_info": [c_void_p, POINTER(c_int64)],
    "gh_stack_frame_ptr": [c_void_p, c_int64, POINTER(c_void_p)],
    "gh_stack_frame_nonfinite": [c_void_p, c_int64, POINTER(c_int64)],
    "gh_set_masters": [c_void_p, c_void_p, c_void_p, c_void_p],
    "gh_upload_frame": [c_void_p, c_int64, c_void_p],
    "gh_download_tile": [c_void_p, POINTER(c_int64), c_int64, c_int64, c_int64, c_int64, c_int64, c_void_p],
    "gh_calibrate_frames": [
        c_void_p, POINTER(c_int64), c_int64, c_void_p, c_int64, c_int32, POINTER(c_float),
        POINTER(CalibPolicy), c_void_p,
    ],
    "gh_normalize_frames": [
        c_void_p, POINTER(c_int64), c_int64, POINTER(c_float), POINTER(c_float), c_void_p,
    ],
    "gh_warp_frames": [
        c_void_p, POINTER(c_int64), c_int64, POINTER(c_double), c_int32, c_float, c_float, c_uint32, c_void_p,
    ],
    "gh_reset_warp_coverage": [c_void_p],
    "gh_accumulate_full_warp_coverage": [c_void_p],
    "gh_warp_coverage_map": [c_void_p, c_void_p],
    "gh_integrate": [c_void_p, POINTER(c_float), POINTER(IntegrateParams), POINTER(IntegrateOutputs), c_void_p],
    "gh_integrate_noalloc": [c_void_p, POINTER(c_float), POINTER(IntegrateParams), POINTER(IntegrateOutputs),
                             c_void_p],
    "gh_frame_stats": [c_void_p, c_int64, c_int32, POINTER(FrameStatistics)],
    "gh_integrate_clip": [
        c_void_p, POINTER(c_float), POINTER(c_double), POINTER(ClipParams), c_void_p, c_void_p, c_void_p, c_void_p,
        c_void_p, c_void_p, c_void_p,
    ],
    "gh_star_candidates": [
        c_void_p, c_int64, POINTER(CandidateParams), c_void_p, c_void_p, c_void_p, c_void_p, POINTER(CandidateResult),
    ],
    "gh_estimate_translation_from_stars": [
        c_void_p, c_int64, c_int64, POINTER(StarTranslationParams), POINTER(StarTranslationResult),
    ],
    "gh_refine_matrix_translation_candidates": [
        c_void_p, c_int64, c_int64, POINTER(c_float), c_int64, POINTER(AlignSearchParams), POINTER(AlignCandidate),
        POINTER(c_int64),
    ],
    "gh_refine_matrix_translation_candidates_batch": [
        c_void_p, c_int64, POINTER(c_int64), c_int64, POINTER(c_float), POINTER(AlignSearchParams),
        POINTER(AlignCandidate),
    ],
    "gh_integrate_tile_local": [
        c_void_p, POINTER(c_float), POINTER(c_uint8), POINTER(c_int32), POINTER(c_float), c_int64,
        POINTER(ClipParams), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
    ],
    "gh_estimate_translation": [
        c_void_p, c_int64, c_int64, c_int32, c_int32, c_int32, c_float, c_float, c_float, c_int32,
        POINTER(c_float), POINTER(c_float), POINTER(c_float), POINTER(c_int64),
    ],
    "gh_translate_frame": [c_void_p, c_int64, c_int32, c_double, c_double, c_float, c_void_p],
    "gh_frames_stats": [c_void_p, POINTER(c_int64), c_int64, POINTER(FrameStatistics)],
    "gh_pair_grid_stats": [
        c_void_p, c_int64, POINTER(c_int64), c_int64, c_int32, c_int32, POINTER(c_float), POINTER(c_float),
        POINTER(c_float), POINTER(c_float), POINTER(c_uint64),
    ],
    "gh_apply_normalization": [
        c_void_p, POINTER(c_int64), c_int64, c_void_p, c_void_p, c_int32, c_int32, c_void_p,
    ],
    "gh_calibration_stats": [c_void_p, c_int64, POINTER(c_double), POINTER(c_double), POINTER(c_int64)],
    "gh_synchronize": [c_void_p],
    "gh_detect_stars": [c_void_p, c_int64, c_double, c_int64, c_double, c_double, c_int32, c_double, c_int64,
                        c_int64, POINTER(StarRecord), c_int64, POINTER(c_int64), POINTER(c_double)],
    "gh_affine_workspace_size": [c_int64, c_int64, POINTER(c_int64)],
    "gh_affine_lanczos3_forward": [c_void_p, c_int64, c_int64, POINTER(c_double), c_float, c_void_p, c_void_p],
    "gh_affine_lanczos3_backward": [c_void_p, c_int64, c_int64, POINTER(c_double), c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_int64, c_void_p],
    "gh_refine_affine": [c_void_p, c_int64, c_int64, POINTER(c_double), c_int32, c_double, POINTER(c_double),
                         POINTER(c_double)],
    "gh_batch_create": [POINTER(BatchJob), c_int64, c_uint32, POINTER(c_void_p)],
    "gh_batch_run": [c_void_p, c_void_p],
    "gh_batch_info": [c_void_p, POINTER(c_int64)],
    "gh_batch_set_schedule": [c_void_p, c_int64, c_int32, c_int32, c_int32],
    "gh_batch_destroy": [c_void_p],
    "gh_triangle_descriptors": [c_void_p, c_void_p, c_int64, c_int32, c_int32, c_int32, c_void_p, c_void_p,
                                c_void_p, POINTER(TriangleInfo)],
    "gh_triangle_descriptors_batch": [POINTER(c_void_p), POINTER(c_void_p), POINTER(c_int64), c_int64, c_int32,
                                      c_int32, c_int32, c_void_p, c_void_p, c_void_p, POINTER(TriangleInfo)],
    "gh_similarity_from_triangles": [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                     c_int64, c_void_p, c_void_p, c_int64, c_float, c_float,
                                     POINTER(SimilarityFit)],
    "gh_similarity_from_triangles_batch": [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_int64,
                                           POINTER(c_void_p), POINTER(c_void_p), POINTER(c_int64),
                                           POINTER(c_void_p), POINTER(c_void_p), POINTER(c_int64), c_float,
                                           c_float, POINTER(SimilarityFit)],
    "gh_similarity_from_catalogs": [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64,
                                    POINTER(CatalogSimilarityParams), POINTER(SimilarityFit),
                                    POINTER(SimilarityCandidate), POINTER(c_int32)],
    "gh_estimate_star_transform": [POINTER(c_double), c_int64, POINTER(c_double), c_int64, c_int32, c_int64,
                                   c_double, c_int64, POINTER(c_double), POINTER(c_int64), POINTER(c_int64),
                                   POINTER(c_double), POINTER(c_int32)],
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_lib = None


def load_library(path: Path | str | None = None) -> ctypes.CDLL:
    """Load libglass_hip.so (cached).  Raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    target = Path(path) if path is not None else LIBRARY_PATH
    if not target.exists():
        raise RuntimeError(
            f"libglass_hip.so not found at {target}; build it with `python -m glass_amd.build` "
            "(the resident stacking path has no CPU fallback)"
        )
    # One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64.so.7.
    # Loading torch first makes our library bind to that copy (same SONAME), so
    # device pointers and streams are shared with torch tensors; loading ours
    # first would bring a second runtime and torch would then see no device.
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - torch is optional for the C ABI
        pass
    lib = ctypes.CDLL(str(target))
    for name, argtypes in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = c_int
    if path is None:
        _lib = lib
    return lib


class NativeError(RuntimeError):
    pass


def last_error() -> str:
    lib = load_library()
    buf = ctypes.create_string_buffer(4096)
    lib.gh_last_error(buf, len(buf))
    return buf.value.decode("utf-8", errors="replace")


def call(name: str, *args) -> None:
    """Call a gh_* entry point and map its status onto Python exceptions the
    way the reference bindings do (ValueError / IndexError / RuntimeError)."""
    lib = load_library()
    status = getattr(lib, name)(*args)
    if status == GH_OK:
        return
    message = last_error()
    if status == GH_ERR_VALUE:
        raise ValueError(message)
    if status == GH_ERR_INDEX:
        raise IndexError(message)
    raise RuntimeError(message)
