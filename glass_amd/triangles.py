"""Triangle-asterism and star-pair similarity registration (libglass_hip.so).

The resident default registration route ``similarity_cuda_triangle`` of the
reference: compact star catalogues -> local triangle descriptors -> a
similarity transform from matched descriptors (or, for the catalogue route,
from ordered two-star pair seeds with an FP64 refit).  Same names, argument
order, defaults, errors and result keys as the reference's native bindings
(src/glass_cuda.py:1726-2240; cpp/src/native_bindings.cpp:17044-18678), so
the reference's wrappers can call these through the module facade
(``glass_amd.native_module``):

* ``triangle_asterism_descriptors_f32(x, y, max_stars=80, neighbors=5,
  max_descriptors=1200)``
* ``triangle_asterism_descriptors_batch_f32(x_list, y_list, ...)``
* ``estimate_similarity_from_triangle_descriptors_f32(reference_x,
  reference_y, moving_x, moving_y, reference_descriptors, reference_indices,
  moving_descriptors, moving_indices, tolerance_px=2.0,
  descriptor_radius=0.1)``
* ``estimate_similarity_from_triangle_descriptors_batch_f32(reference_x,
  reference_y, reference_descriptors, reference_indices, moving_x_list,
  moving_y_list, moving_descriptors_list, moving_indices_list, ...)``
* ``estimate_similarity_from_catalogs_f32(reference_x, reference_y, moving_x,
  moving_y, tolerance_px=2.0, min_pair_distance=2.0, prior_dx=None,
  prior_dy=None, prior_radius_px=None, min_scale=None, max_scale=None,
  max_abs_rotation_rad=None, top_k=0)`` -- also accepts the binding's
  positional floats (0 / -1 / FLT_MAX for "unset").

Every call runs on the GPU (HIP kernels of glass_amd/csrc/triangles.hip);
there is no CPU route.
"""

from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _native as N

_FLT_MAX = float(np.finfo(np.float32).max)


def _coords(v) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1))


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a.size else None


def _table(arrays) -> ctypes.Array:
    table = (ctypes.c_void_p * max(len(arrays), 1))()
    for i, a in enumerate(arrays):
        table[i] = a.ctypes.data if a.size else None
    return table


def _descriptor_rows(v, name: str) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float32))
    if a.ndim != 2 or a.shape[1] != 2:
        raise ValueError(f"{name}triangle descriptors must have shape (N, 2)")
    return a


def _index_rows(v, name: str) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32))
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError(f"{name}triangle indices must have shape (N, 3)")
    return a


# ---------------------------------------------------------------- descriptors


def _descriptor_result(info, desc, idx, areas, max_stars, model) -> dict:
    n = int(info.count)
    return {
        "count": n,
        "raw_count": int(info.raw_count),
        "max_stars": int(max_stars),
        "neighbors": int(info.neighbors),
        "descriptors": desc[:n].copy(),
        "indices": idx[:n].copy(),
        "areas": areas[:n].copy(),
        "model": model,
    }


def triangle_asterism_descriptors_f32(x, y, max_stars: int = 80, neighbors: int = 5,
                                      max_descriptors: int = 1200) -> dict:
    """Local triangle descriptors of one catalogue (glass_cuda.py:1835-1901,
    native_bindings.cpp:17044-17253)."""
    x, y = _coords(x), _coords(y)
    if x.shape != y.shape:
        raise ValueError("catalog coordinate arrays must have the same shape")
    max_stars, neighbors, max_descriptors = int(max_stars), int(neighbors), int(max_descriptors)
    cap = max(max_descriptors, 0)
    desc = np.empty((cap, 2), np.float32)
    idx = np.empty((cap, 3), np.int32)
    areas = np.empty(cap, np.float32)
    info = N.TriangleInfo()
    N.call("gh_triangle_descriptors", _ptr(x), _ptr(y), x.size, max_stars, neighbors, max_descriptors, _ptr(desc),
           _ptr(idx), _ptr(areas), ctypes.byref(info))
    return _descriptor_result(info, desc, idx, areas, max_stars, "triangle_asterism_descriptors_hip")


def triangle_asterism_descriptors_batch_f32(x_list, y_list, max_stars: int = 80, neighbors: int = 5,
                                            max_descriptors: int = 1200) -> list[dict]:
    """Descriptors of many catalogues in one launch (glass_cuda.py:1935-1986,
    native_bindings.cpp:17255-17595); element b equals the single call on
    catalogue b."""
    start = time.perf_counter()
    if len(x_list) != len(y_list):
        raise ValueError("catalog x/y batch lists must have the same length")
    if max_stars < 0:
        raise ValueError("max_stars must be non-negative")
    if max_descriptors < 0:
        raise ValueError("max_descriptors must be non-negative")
    if len(x_list) == 0:
        return []
    xs = [_coords(v) for v in x_list]
    ys = [_coords(v) for v in y_list]
    for a, b in zip(xs, ys):
        if a.shape != b.shape:
            raise ValueError("catalog coordinate arrays must have matching x/y shapes")
    batch = len(xs)
    max_stars, neighbors, max_descriptors = int(max_stars), int(neighbors), int(max_descriptors)
    desc = np.empty((batch, max_descriptors, 2), np.float32)
    idx = np.empty((batch, max_descriptors, 3), np.int32)
    areas = np.empty((batch, max_descriptors), np.float32)
    infos = (N.TriangleInfo * batch)()
    sizes = np.asarray([a.size for a in xs], np.int64)
    prepared = time.perf_counter()
    N.call("gh_triangle_descriptors_batch", _table(xs), _table(ys), sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
           batch, max_stars, neighbors, max_descriptors, _ptr(desc), _ptr(idx), _ptr(areas), infos)
    called = time.perf_counter()
    out = []
    for b in range(batch):
        r = _descriptor_result(infos[b], desc[b], idx[b], areas[b], max_stars,
                               "triangle_asterism_descriptors_hip_batch_one_launch")
        r.update(batch_index=b, batch_count=batch, batch_model="triangle_asterism_descriptors_hip_batch_one_launch",
                 batch_timing_model="one_native_call_upload_kernel_download_one_sync",
                 batch_host_prepare_s=prepared - start, batch_upload_s=0.0, batch_kernel_sync_s=called - prepared,
                 batch_output_download_s=0.0, batch_total_elapsed_s_at_result=time.perf_counter() - start)
        out.append(r)
    return out


# ---------------------------------------------------------------- similarity


def _matrix(fit) -> list:
    m = np.frombuffer(bytes(fit.matrix), dtype=np.float32).reshape(3, 3)
    return m.tolist()


def _triangle_fit_result(fit, nr, nm, nrd, nmd, tolerance_px, descriptor_radius) -> dict:
    return {
        "matrix": _matrix(fit),
        "scale": float(fit.scale),
        "rotation_rad": float(fit.rotation_rad),
        "rms_px": float(fit.rms_px),
        "inliers": int(fit.inliers),
        "best_candidate_index": int(fit.best_candidate_index),
        "candidate_count": int(fit.candidate_count),
        "scored_candidates": int(fit.scored_candidates),
        "reference_count": int(nr),
        "moving_count": int(nm),
        "reference_descriptor_count": int(nrd),
        "moving_descriptor_count": int(nmd),
        "tolerance_px": float(np.float32(tolerance_px)),
        "descriptor_radius": float(np.float32(descriptor_radius)),
        "status": "ok" if fit.inliers > 0 else "failed",
        "model": "triangle_descriptor_similarity_hip",
        "best_reduction_mode": "per_block_best_then_single_block",
    }


def estimate_similarity_from_triangle_descriptors_f32(reference_x, reference_y, moving_x, moving_y,
                                                      reference_descriptors, reference_indices, moving_descriptors,
                                                      moving_indices, tolerance_px: float = 2.0,
                                                      descriptor_radius: float = 0.1) -> dict:
    """Similarity from matched triangle descriptors (glass_cuda.py:2050-2140,
    native_bindings.cpp:17597-17848)."""
    rx, ry, mx, my = _coords(reference_x), _coords(reference_y), _coords(moving_x), _coords(moving_y)
    if rx.shape != ry.shape or mx.shape != my.shape:
        raise ValueError("catalog coordinate arrays must have matching x/y shapes")
    rd, md = _descriptor_rows(reference_descriptors, ""), _descriptor_rows(moving_descriptors, "")
    ri, mi = _index_rows(reference_indices, ""), _index_rows(moving_indices, "")
    if rd.shape[0] != ri.shape[0] or md.shape[0] != mi.shape[0]:
        raise ValueError("descriptor and index row counts must match")
    if tolerance_px < 0.0:
        raise ValueError("tolerance_px must be non-negative")
    if descriptor_radius < 0.0:
        raise ValueError("descriptor_radius must be non-negative")
    fit = N.SimilarityFit()
    N.call("gh_similarity_from_triangles", _ptr(rx), _ptr(ry), rx.size, _ptr(mx), _ptr(my), mx.size, _ptr(rd),
           _ptr(ri), rd.shape[0], _ptr(md), _ptr(mi), md.shape[0], float(tolerance_px), float(descriptor_radius),
           ctypes.byref(fit))
    return _triangle_fit_result(fit, rx.size, mx.size, rd.shape[0], md.shape[0], tolerance_px, descriptor_radius)


def estimate_similarity_from_triangle_descriptors_batch_f32(reference_x, reference_y, reference_descriptors,
                                                            reference_indices, moving_x_list, moving_y_list,
                                                            moving_descriptors_list, moving_indices_list,
                                                            tolerance_px: float = 2.0,
                                                            descriptor_radius: float = 0.1) -> list[dict]:
    """Triangle similarity of many moving catalogues against one reference
    (glass_cuda.py:2143-2240, native_bindings.cpp:17850-18252): one upload,
    every frame's kernels on one stream, one synchronisation."""
    start = time.perf_counter()
    lengths = {len(moving_x_list), len(moving_y_list), len(moving_descriptors_list), len(moving_indices_list)}
    if len(lengths) != 1:
        raise ValueError("moving batch lists must have the same length")
    if len(moving_x_list) == 0:
        return []
    if tolerance_px < 0.0:
        raise ValueError("tolerance_px must be non-negative")
    if descriptor_radius < 0.0:
        raise ValueError("descriptor_radius must be non-negative")
    rx, ry = _coords(reference_x), _coords(reference_y)
    if rx.shape != ry.shape:
        raise ValueError("catalog coordinate arrays must have matching x/y shapes")
    rd, ri = _descriptor_rows(reference_descriptors, "reference "), _index_rows(reference_indices, "reference ")
    if rd.shape[0] != ri.shape[0]:
        raise ValueError("reference descriptor and index row counts must match")
    mxs = [_coords(v) for v in moving_x_list]
    mys = [_coords(v) for v in moving_y_list]
    mds = [_descriptor_rows(v, "moving ") for v in moving_descriptors_list]
    mis = [_index_rows(v, "moving ") for v in moving_indices_list]
    for a, b, d, i in zip(mxs, mys, mds, mis):
        if a.shape != b.shape:
            raise ValueError("catalog coordinate arrays must have matching x/y shapes")
        if d.shape[0] != i.shape[0]:
            raise ValueError("moving descriptor and index row counts must match")
    batch = len(mxs)
    nm = np.asarray([a.size for a in mxs], np.int64)
    nmd = np.asarray([d.shape[0] for d in mds], np.int64)
    fits = (N.SimilarityFit * batch)()
    i64p = ctypes.POINTER(ctypes.c_int64)
    prepared = time.perf_counter()
    N.call("gh_similarity_from_triangles_batch", _ptr(rx), _ptr(ry), rx.size, _ptr(rd), _ptr(ri), rd.shape[0], batch,
           _table(mxs), _table(mys), nm.ctypes.data_as(i64p), _table(mds), _table(mis), nmd.ctypes.data_as(i64p),
           float(tolerance_px), float(descriptor_radius), fits)
    called = time.perf_counter()
    reference_bytes = rx.nbytes + ry.nbytes + rd.nbytes + ri.nbytes
    out = []
    for b in range(batch):
        r = _triangle_fit_result(fits[b], rx.size, mxs[b].size, rd.shape[0], mds[b].shape[0], tolerance_px,
                                 descriptor_radius)
        r.update(batch_index=b, batch_count=batch,
                 batch_model="triangle_descriptor_similarity_hip_batch_shared_reference_device",
                 reference_device_reuse=True, reference_device_bytes=int(reference_bytes),
                 moving_device_reuse=True,
                 moving_device_bytes=int(mxs[b].nbytes + mys[b].nbytes + mds[b].nbytes + mis[b].nbytes),
                 output_device_reuse=True, output_device_bytes=int(ctypes.sizeof(N.SimilarityFit)),
                 batch_timing_model="one_native_call_all_frames_one_stream_one_sync",
                 batch_host_prepare_s=prepared - start, batch_frame_total_s=(called - prepared) / batch,
                 batch_total_elapsed_s_at_result=time.perf_counter() - start)
        out.append(r)
    return out


def estimate_similarity_from_catalogs_f32(reference_x, reference_y, moving_x, moving_y, tolerance_px: float = 2.0,
                                          min_pair_distance: float = 2.0, prior_dx=None, prior_dy=None,
                                          prior_radius_px=None, min_scale=None, max_scale=None,
                                          max_abs_rotation_rad=None, top_k: int = 0) -> dict:
    """Similarity from two compact star catalogues by ordered two-star pair
    seeds, scored by mutual nearest neighbours, with the FP64 refit
    (glass_cuda.py:1726-1824, native_bindings.cpp:18254-18678)."""
    rx, ry, mx, my = _coords(reference_x), _coords(reference_y), _coords(moving_x), _coords(moving_y)
    if rx.shape != ry.shape or mx.shape != my.shape:
        raise ValueError("catalog coordinate arrays must have matching x/y shapes")
    p = N.CatalogSimilarityParams(
        float(tolerance_px), float(min_pair_distance), float(0.0 if prior_dx is None else prior_dx),
        float(0.0 if prior_dy is None else prior_dy), float(-1.0 if prior_radius_px is None else prior_radius_px),
        float(0.0 if min_scale is None else min_scale), float(_FLT_MAX if max_scale is None else max_scale),
        float(-1.0 if max_abs_rotation_rad is None else max_abs_rotation_rad), int(top_k))
    k = max(int(top_k), 0)
    top = (N.SimilarityCandidate * max(k, 1))()
    top_count = ctypes.c_int32(0)
    fit = N.SimilarityFit()
    N.call("gh_similarity_from_catalogs", _ptr(rx), _ptr(ry), rx.size, _ptr(mx), _ptr(my), mx.size, ctypes.byref(p),
           ctypes.byref(fit), top, ctypes.byref(top_count))
    prior_on = p.prior_radius_px >= 0.0
    candidates = []
    for i in range(top_count.value):
        c = top[i]
        a, b = np.float32(c.a), np.float32(c.b)
        candidates.append({
            "candidate_index": int(c.candidate_index),
            "inliers": int(c.inliers),
            "rms_px": float(c.rms_px),
            "matrix": [[float(a), float(-b), float(c.tx)], [float(b), float(a), float(c.ty)], [0.0, 0.0, 1.0]],
            "scale": float(np.sqrt(a * a + b * b)),
            "rotation_rad": float(np.arctan2(b, a)),
        })
    status = int(fit.refit_status)
    return {
        "matrix": _matrix(fit),
        "scale": float(fit.scale),
        "rotation_rad": float(fit.rotation_rad),
        "rms_px": float(fit.rms_px),
        "inliers": int(fit.inliers),
        "refined_inliers": int(fit.refined_inliers),
        "refit_status": "ok" if status == 0 else ("rejected" if status == 3 else "failed"),
        "refit_status_code": status,
        "refit_rms_px": float(fit.refit_rms_px),
        "best_candidate_index": int(fit.best_candidate_index),
        "candidate_count": int(fit.candidate_count),
        "scored_candidates": int(fit.scored_candidates),
        "reference_count": int(rx.size),
        "moving_count": int(mx.size),
        "tolerance_px": float(p.tolerance_px),
        "min_pair_distance": float(p.min_pair_distance),
        "prior_dx": float(p.prior_dx) if prior_on else 0.0,
        "prior_dy": float(p.prior_dy) if prior_on else 0.0,
        "prior_radius_px": float(p.prior_radius_px),
        "min_scale": float(p.min_scale),
        "max_scale": float(p.max_scale),
        "max_abs_rotation_rad": float(p.max_abs_rotation_rad) if p.max_abs_rotation_rad >= 0.0 else -1.0,
        "top_k": int(top_k),
        "top_candidates": candidates,
        "status": "ok" if fit.inliers > 0 else "failed",
        "model": "catalog_pair_similarity_hip",
    }


# ---------------------------------------------------------------- resident route


def register_resident_frames(stack, reference_index: int, indices, threshold: float, max_stars: int = 64,
                             neighbors: int = 5, max_descriptors: int = 1200, tolerance_px: float = 1.0,
                             descriptor_radius: float = 0.1, scan_candidates: int | None = None,
                             min_separation_px: float | None = None, centroid_radius: int = 4,
                             centroid_global_mean_background: bool = True) -> dict:
    """The catalogue -> descriptor -> similarity chain of the resident default
    registration (resident_cuda.py:10820-10860 catalogues, :11200-11300 batched
    descriptors and fits; defaults of resident_cuda.py:7134-7141 and
    :10320-10510) on resident frames of a ``ResidentCalibratedStack``:

    * catalogue of each frame: the brightest ``scan_candidates`` (default
      4 x max_stars) local maxima above ``threshold``, greedy separation
      ``min_separation_px`` (default max(32, min(h, w) / 100)), ``max_stars``
      kept, window centroids of radius ``centroid_radius``;
    * descriptors of every catalogue in one batched call;
    * similarity of every moving frame against the reference in one batched
      call.

    Returns {"fits": {index: fit dict (matrix moving -> reference)},
    "catalogue_counts": {index: stars}, "timing_s": {...}}.  Frames whose
    catalogue has fewer than 3 stars or no descriptor get no fit (the
    reference skips them the same way, resident_cuda.py:11209-11212,
    :11283-11287)."""
    h, w = stack.height, stack.width
    scan = int(scan_candidates) if scan_candidates is not None else max(int(max_stars), int(max_stars) * 4)
    sep = float(min_separation_px) if min_separation_px is not None else max(32.0, float(min(h, w)) / 100.0)
    frames = [int(reference_index)] + [int(i) for i in indices if int(i) != int(reference_index)]
    t0 = time.perf_counter()
    catalogs = {}
    for i in frames:
        catalogs[i] = stack.star_top_nms_candidates_centroid(i, float(threshold), scan, int(max_stars), sep,
                                                             int(centroid_radius),
                                                             bool(centroid_global_mean_background))
    t1 = time.perf_counter()
    usable = [i for i in frames if int(catalogs[i]["stored_count"]) >= 3]
    desc = {}
    if usable:
        for i, d in zip(usable, triangle_asterism_descriptors_batch_f32(
                [catalogs[i]["x"] for i in usable], [catalogs[i]["y"] for i in usable], max_stars=max_stars,
                neighbors=neighbors, max_descriptors=max_descriptors)):
            desc[i] = d
    t2 = time.perf_counter()
    fits = {}
    ref = int(reference_index)
    moving = [i for i in usable if i != ref and desc[i]["count"] > 0]
    if ref in desc and desc[ref]["count"] > 0 and moving:
        rc = catalogs[ref]
        batch = estimate_similarity_from_triangle_descriptors_batch_f32(
            rc["x"], rc["y"], desc[ref]["descriptors"], desc[ref]["indices"], [catalogs[i]["x"] for i in moving],
            [catalogs[i]["y"] for i in moving], [desc[i]["descriptors"] for i in moving],
            [desc[i]["indices"] for i in moving], tolerance_px=tolerance_px, descriptor_radius=descriptor_radius)
        fits = dict(zip(moving, batch))
    t3 = time.perf_counter()
    return {
        "fits": fits,
        "catalogs": catalogs,
        "descriptors": desc,
        "catalogue_counts": {i: int(catalogs[i]["stored_count"]) for i in frames},
        "timing_s": {"catalogues": t1 - t0, "descriptors": t2 - t1, "similarity": t3 - t2, "total": t3 - t0},
    }
