"""The module facade that the reference's ``glass_cuda._native()`` returns in
place of ``_glass_cuda_native`` (CPU: no compute calls).

The attribute lists below are what src/glass_cuda.py reads off the native
module and off the native stack object on the resident path, with the line of
each first use.  The facade must answer every one, and must NOT carry the
optional attributes whose ``hasattr`` probe selects a Python route the
reference keeps as its default.
"""

from __future__ import annotations

import inspect

import numpy as np
import pytest

import glass_amd
from glass_amd import native_module

# Module attributes: glass_cuda.py line of the call (and resident_cuda.py caller).
MODULE_ATTRIBUTES = [
    ("cuda_available", 66),          # resident_cuda.py:205, :6051
    ("list_devices", 189),
    ("get_device_info", 208),        # resident_cuda.py:548, :6053, :16227
    ("host_pinned_empty_f32", 219),  # resident_cuda.py host rings
    ("host_pinned_empty_u8", 226),
    ("ResidentCalibratedStack", 2665),  # resident_cuda.py:205 hasattr, stack construction
    ("estimate_similarity_from_catalogs_f32", 1753),  # resident_cuda.py:9985
    ("triangle_asterism_descriptors_f32", 1855),  # resident_cuda.py:10852 (required at :10311)
    ("triangle_asterism_descriptors_batch_f32", 1967),  # resident_cuda.py:11220
    ("estimate_similarity_from_triangle_descriptors_f32", 2084),  # resident_cuda.py:11399 (required at :10311)
    ("estimate_similarity_from_triangle_descriptors_batch_f32", 2202),  # resident_cuda.py:11294
]

# Probed with hasattr; absent means the reference's own default route.
ABSENT_OPTIONAL = [
    ("RawFitsReadQueue", 381),                   # raw_fits_read_queue_available()
    ("read_simple_fits_raw_batch_into_u8", 377),
    ("resident_dq_map_host_f32", 230),           # Python DQ is the default (:240-252)
    ("resident_dq_map_count_maps_i16", 257),
]

# glass_cuda.ResidentCalibratedStack forwards to these members of the native
# object (self._impl.<name>, or getattr by name), glass_cuda.py line of use.
STACK_MEMBERS = [
    ("frame_count", 2669), ("height", 2673), ("width", 2677), ("pixels_per_frame", 2681),
    ("loaded_count", 2685), ("bytes_allocated", 2689), ("warp_scratch_bytes", 2695),
    ("warp_copy_mode", 2701), ("warp_coverage_frame_count", 2725),
    ("set_calibration_masters", 2733), ("reset_warp_coverage", 2742),
    ("accumulate_full_warp_coverage_frame", 2749), ("warp_coverage_map", 2754),
    ("download_frame_tile", 2760), ("download_frames_tile", 2773),
    ("upload_calibrated_frame", 2789), ("calibrate_frame", 3175), ("calibrate_frame_timed", 3202),
    ("calibrate_frame_pinned_async", 3221), ("calibrate_frame_pinned_async_timed", 3239),
    ("calibrate_frame_host_async_timed", 3258), ("calibrate_frames_host_async_timed", 3277),
    ("calibrate_frames_host_async_multistream_timed", 3299),
    ("calibrate_frames_host_async_multistream_h2d_release_timed", 3323),
    ("finish_pending_calibration_timed", 3336),
    ("calibrate_frames_host_async_multistream_callback_release_timed", 3354),
    ("calibrate_frames_fits_u16be_bzero_host_async_multistream_callback_release_timed", 3377),
    ("calibrate_frames_fits_u16be_bzero_paths_multistream_timed", 3404),
    ("calibrate_frames_fits_u16be_bzero_paths_completion_queue_timed", 3433),
    ("begin_fits_u16be_bzero_paths_completion_queue_read_timed", 3461),
    ("finish_fits_u16be_bzero_paths_completion_queue_calibration_timed", 3483),
    ("cancel_fits_u16be_bzero_paths_completion_queue_read", 3495),
    ("apply_translation_frame", 3504), ("apply_translation_bilinear_frame", 3509),
    ("apply_matrix_bilinear_frame", 3514), ("apply_matrix_bilinear_frames", 3525),
    ("apply_matrix_lanczos3_frame", 3576), ("apply_matrix_lanczos3_frames", 3593),
    ("refine_matrix_translation_candidates_to_reference", 3762),
    ("refine_matrix_translation_candidates_batch_to_reference", 3821),
    ("estimate_translation_to_reference", 3847), ("estimate_translation_subpixel_to_reference", 3883),
    ("frame_global_stats", 3911), ("frame_pair_grid_stats", 4125), ("frame_pair_grid_stats_batch", 4157),
    ("apply_global_normalization_frame", 4211), ("apply_grid_normalization_frame", 4247),
    ("apply_grid_normalization_frames", 4297),
    ("star_candidates", 4359), ("star_top_candidates", 4371), ("star_top_nms_candidates", 4391),
    ("star_top_nms_candidates_centroid", 4425), ("star_grid_top_nms_candidates_centroid", 4509),
    ("star_grid_top_nms_candidates_deterministic_centroid", 4554),
    ("estimate_translation_from_stars_to_reference", 4813),
    ("integrate_mean", 4859), ("integrate_tile_local_mean", 4881), ("integrate_sigma_clip", 4911),
    ("integrate_hardened_winsorized_sigma", 4950), ("integrate_tile_local_sigma_clip", 5123),
    ("integrate_matrix_warped_mean", 5161), ("integrate_matrix_warped_sigma_clip", 5237),
]


@pytest.mark.parametrize("name,line", MODULE_ATTRIBUTES)
def test_facade_module_attribute(name, line):
    assert hasattr(native_module, name), f"glass_cuda.py:{line} reads native.{name}"


@pytest.mark.parametrize("name,line", ABSENT_OPTIONAL)
def test_optional_attribute_absent(name, line):
    assert not hasattr(native_module, name), f"glass_cuda.py:{line} would take the native route"


@pytest.mark.parametrize("name,line", STACK_MEMBERS)
def test_stack_member(name, line):
    assert hasattr(native_module.ResidentCalibratedStack, name), \
        f"glass_cuda.py:{line} calls the native stack's {name}"


def test_triangle_signatures_match_native_positional_order():
    """glass_cuda.py passes the native calls positionally (:1753-1767,
    :1855-1861, :1967-1973, :2084-2095, :2202-2213)."""
    order = {
        "estimate_similarity_from_catalogs_f32": [
            "reference_x", "reference_y", "moving_x", "moving_y", "tolerance_px", "min_pair_distance", "prior_dx",
            "prior_dy", "prior_radius_px", "min_scale", "max_scale", "max_abs_rotation_rad", "top_k"],
        "triangle_asterism_descriptors_f32": ["x", "y", "max_stars", "neighbors", "max_descriptors"],
        "triangle_asterism_descriptors_batch_f32": ["x_list", "y_list", "max_stars", "neighbors", "max_descriptors"],
        "estimate_similarity_from_triangle_descriptors_f32": [
            "reference_x", "reference_y", "moving_x", "moving_y", "reference_descriptors", "reference_indices",
            "moving_descriptors", "moving_indices", "tolerance_px", "descriptor_radius"],
        "estimate_similarity_from_triangle_descriptors_batch_f32": [
            "reference_x", "reference_y", "reference_descriptors", "reference_indices", "moving_x_list",
            "moving_y_list", "moving_descriptors_list", "moving_indices_list", "tolerance_px", "descriptor_radius"],
    }
    for name, params in order.items():
        assert list(inspect.signature(getattr(native_module, name)).parameters) == params, name


def test_integrate_signature_matches_reference_keywords():
    # native_bindings.cpp:21174-21184 argument names and defaults
    sig = inspect.signature(native_module.ResidentCalibratedStack.integrate_hardened_winsorized_sigma)
    params = sig.parameters
    assert [p for p in params][1:9] == ["weights", "low_sigma", "high_sigma", "min_samples",
                                        "max_reject_fraction", "count_map_dtype", "profile", "download_mode"]
    assert params["min_samples"].default == 3 and params["max_reject_fraction"].default == 0.5
    assert params["count_map_dtype"].default == "float32" and params["download_mode"].default == "full"


class _ReferenceLoader:
    """The reference's `_native()` caching loader (glass_cuda.py:31-43) with
    the drop-in's one-line binding: it returns glass_amd.native_module."""

    def __init__(self):
        self.attempted = False
        self.module = None

    def _native(self):
        if not self.attempted:
            self.attempted = True
            self.module = native_module
        return self.module

    # glass_cuda.cuda_available (glass_cuda.py:62-66)
    def cuda_available(self):
        native = self._native()
        if native is None:
            return False
        return bool(native.cuda_available())

    # glass_cuda.raw_fits_read_queue_available (glass_cuda.py:381-383)
    def raw_fits_read_queue_available(self):
        native = self._native()
        return native is not None and hasattr(native, "RawFitsReadQueue")


def test_reference_loader_binding_without_device():
    loader = _ReferenceLoader()
    # no HIP device in this container: False, never an exception
    assert loader.cuda_available() is glass_amd.cuda_available()
    assert loader.raw_fits_read_queue_available() is False


def test_pinned_arrays_reject_empty_shapes():
    with pytest.raises(ValueError, match="non-empty"):
        native_module.host_pinned_empty_f32(0, 4)
    with pytest.raises(ValueError, match="non-empty"):
        native_module.host_pinned_empty_u8(0)


@pytest.mark.gpu
def test_device_info_and_pinned_arrays(gpu):
    info = native_module.get_device_info(0)
    for key in ("device_id", "name", "compute_capability", "memory_total_mib", "multi_processor_count",
                "native_backend", "available_to_glass"):
        assert key in info
    assert info["arch"].startswith("gfx950")
    assert info["memory_total_mib"] > 200 * 1024 and info["multi_processor_count"] >= 256
    assert len(native_module.list_devices()) >= 1
    with pytest.raises(IndexError):
        native_module.get_device_info(1 << 20)
    a = native_module.host_pinned_empty_f32(3, 5)
    assert a.shape == (3, 5) and a.dtype == np.float32
    a[...] = 7.0
    b = native_module.host_pinned_empty_u8(17)
    b[...] = 3
    assert float(a.sum()) == 105.0 and int(b.sum()) == 51
    # the pinned buffer feeds the upload ring like any host array
    with native_module.ResidentCalibratedStack(1, 3, 5) as s:
        s.upload_calibrated_frame(0, a)
        assert np.array_equal(s.download_frame_tile(0, 0, 0, 5, 3), a)
