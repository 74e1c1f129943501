"""Batched multi-job stacking: many resident jobs in one launch sequence.

``StackingBatch(jobs, use_graph=True, autotune=True)`` takes a list of job
dicts, each naming a loaded ``ResidentCalibratedStack`` and what to do with it:

    {"stack": stack,                       # frames already calibrated
     "warp_indices": [...], "matrices": [...],   # moving -> reference, affine
     "interpolation": "lanczos3" | "bilinear",
     "semantics": "cpu" | "resident", "fill": nan, "clamping_threshold": -1.0,
     "weights": None | array[frame_count],
     "low_sigma": 3.0, "high_sigma": 3.0, "min_samples": 3,
     "max_reject_fraction": 0.5, "rejection": "winsorized_sigma" | "none",
     "count_map_dtype": "uint16" | "float32",
     "outputs": {"master": t, "weight": t, "coverage": t, "low": t, "high": t, "dq": t},
     "dq_geometric": False, "dq_active_frames": 0}

``run(stream=None)`` executes every job: one batched warp launch over all
jobs' (frame group, tile) items, one launch deciding each job's statistics
arithmetic from the warped frames' non-finite flags, and one integration launch
per sort-network capacity class over the pixels of all its jobs (a device job
table) -- captured into a hipGraph on the first run (after the autotuner has
timed candidate schedules: warp grid, item order and frames per item;
integration capacity classes) and replayed.  ``set_schedule`` pins a
schedule instead.  Outputs equal the per-job ``warp_frames`` +
``integrate_to_device`` calls whatever the schedule; the jobs' stacks are
read, not modified.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native as N
from .resident import invert_affine


def _job_struct(job, keep):
    import torch

    stack = job["stack"]
    idx = np.ascontiguousarray(np.asarray(job.get("warp_indices", []), dtype=np.int64).reshape(-1))
    mats = np.asarray(job.get("matrices", np.zeros((0, 3, 3))), dtype=np.float64).reshape(-1, 3, 3)
    if mats.shape[0] != idx.size:
        raise ValueError("matrices must have shape (len(warp_indices), 3, 3)")
    inv = np.ascontiguousarray(np.stack([invert_affine(m) for m in mats]) if idx.size else np.zeros((0, 3, 3)))
    keep += [idx, inv]
    interp = N.GH_INTERP_LANCZOS3 if job.get("interpolation", "lanczos3") == "lanczos3" else N.GH_INTERP_BILINEAR
    flags = N.GH_WARP_SKIP_NONFINITE_TAPS if job.get("semantics", "cpu") == "resident" else 0
    weights = job.get("weights")
    wptr = None
    if weights is not None:
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float32).reshape(-1))
        if w.size != stack.frame_count:
            raise ValueError("weights must have shape (frame_count,)")
        keep.append(w)
        wptr = w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    cdt = N.GH_COUNT_U16 if job.get("count_map_dtype", "uint16") == "uint16" else N.GH_COUNT_F32
    rej = N.GH_REJECT_NONE if job.get("rejection", "winsorized_sigma") == "none" else N.GH_REJECT_WINSORIZED_SIGMA
    params = N.IntegrateParams(rej, float(job.get("low_sigma", 3.0)), float(job.get("high_sigma", 3.0)),
                               int(job.get("min_samples", 3)), float(job.get("max_reject_fraction", 0.5)), cdt, -1, 0, 0)
    outs = job.get("outputs", {})
    ptr = {}
    for name in ("master", "weight", "coverage", "low", "high", "dq"):
        t = outs.get(name)
        if t is None:
            ptr[name] = None
            continue
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
            raise ValueError(f"output {name!r} must be a contiguous GPU tensor")
        if t.numel() != stack.pixels_per_frame:
            raise ValueError(f"output {name!r} must hold height*width elements")
        keep.append(t)
        ptr[name] = ctypes.c_void_p(t.data_ptr())
    o = N.IntegrateOutputs(ptr["master"], ptr["weight"], ptr["coverage"], ptr["low"], ptr["high"], ptr["dq"],
                           1 if (ptr["dq"] is not None and job.get("dq_geometric", False)) else 0,
                           int(job.get("dq_active_frames", 0)))
    return N.BatchJob(stack._h, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                      inv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), idx.size, interp, flags,
                      float(job.get("fill", math.nan)), float(job.get("clamping_threshold", -1.0)), wptr, params, o)


class StackingBatch:
    def __init__(self, jobs, use_graph: bool = True, autotune: bool = True):
        keep = []
        arr = (N.BatchJob * len(jobs))(*[_job_struct(j, keep) for j in jobs])
        flags = (N.GH_BATCH_USE_GRAPH if use_graph else 0) | (N.GH_BATCH_AUTOTUNE if autotune else 0)
        handle = ctypes.c_void_p()
        N.call("gh_batch_create", arr, len(jobs), flags, ctypes.byref(handle))
        self._h = handle
        self._keep = [j["stack"] for j in jobs] + [j.get("outputs", {}) for j in jobs]

    def run(self, stream=None) -> None:
        """Enqueue one execution on `stream` (default: torch's current stream).
        A null stream handle (torch's default stream) runs the batch on its
        own stream and returns when it has finished."""
        if stream is None:
            import torch

            stream = torch.cuda.current_stream().cuda_stream
        N.call("gh_batch_run", self._h, ctypes.c_void_p(stream or None))

    def info(self) -> dict:
        out = (ctypes.c_int64 * 10)()
        N.call("gh_batch_info", self._h, out)
        keys = ("jobs", "warp_items", "integrate_launches", "graph", "warp_grid", "warp_item_stride",
                "device_variant_jobs", "autotune_trials", "warp_frames_per_item", "integrate_classes")
        return {k: int(v) for k, v in zip(keys, out)}

    def set_schedule(self, warp_grid: int, warp_item_stride: int = 1, frames_per_item: int = 8,
                     integrate_classes: int = 0) -> None:
        """Pin the schedule (no autotuning): warp workgroups, work-item order
        stride (1 or coprime to the item count), frames per warp work item (2,
        4 or 8) and integration capacity classes (0 smallest capacity per job,
        1 powers of two, 2 one launch).  Results do not depend on it."""
        N.call("gh_batch_set_schedule", self._h, int(warp_grid), int(warp_item_stride), int(frames_per_item),
               int(integrate_classes))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            N.call("gh_batch_destroy", self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
