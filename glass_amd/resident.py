"""Drop-in ``ResidentCalibratedStack`` for the resident stacking path.

Mirrors the hot-path surface of the reference operator API
``glass_cuda.ResidentCalibratedStack`` (src/glass_cuda.py:2653) / the pybind11
class of cpp/src/native_bindings.cpp:20624-21213: same method names, argument
order, defaults, return shapes and exception types.  Every call goes through
the C ABI of libglass_hip.so (glass_amd/_native.py); there is no CPU path.

Inputs may be numpy arrays (copied in, like pybind ``c_style | forcecast``) or
torch tensors already on the GPU (zero-copy: their device pointer is passed).
"""

from __future__ import annotations

import ctypes
import math
import time
from typing import Any, Iterable, Sequence

import numpy as np

from . import _native as N

_DEFAULT_POLICY = {
    "master_dark_includes_bias": True,
    "dark_scaling_enabled": True,
    "flat_floor": 1.0e-6,
    "pedestal": 0.0,
}


def _is_torch_cuda(obj) -> bool:
    return type(obj).__module__.startswith("torch") and getattr(obj, "is_cuda", False)


def _buffer(obj, dtype, shape=None, name="array"):
    """(pointer, keepalive) for a host array or a device tensor."""
    if _is_torch_cuda(obj):
        import torch

        want = {np.float32: torch.float32, np.uint16: torch.int16, np.uint8: torch.uint8}.get(np.dtype(dtype).type)
        if not obj.is_contiguous():
            raise ValueError(f"{name} device tensor must be contiguous")
        if want is not None and obj.dtype not in (want, torch.uint8) and not (
            np.dtype(dtype) == np.uint16 and obj.dtype in (torch.int16, torch.uint16 if hasattr(torch, "uint16") else torch.int16)
        ):
            raise ValueError(f"{name} device tensor has dtype {obj.dtype}, expected {np.dtype(dtype)}")
        if shape is not None and obj.numel() * obj.element_size() != int(np.prod(shape)) * np.dtype(dtype).itemsize:
            raise ValueError(f"{name} must have shape {tuple(shape)}")
        return ctypes.c_void_p(obj.data_ptr()), obj
    if type(obj).__module__.startswith("torch"):
        obj = obj.detach().cpu().numpy()
    arr = np.ascontiguousarray(obj, dtype=dtype)
    if shape is not None and arr.size != int(np.prod(shape)):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {arr.shape}")
    if shape is not None and arr.ndim == 2 and tuple(arr.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {arr.shape}")
    return ctypes.c_void_p(arr.ctypes.data), arr


def _index_array(indices) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(list(indices) if not isinstance(indices, np.ndarray) else indices), dtype=np.int64)
    if arr.ndim != 1:
        raise ValueError("indices must be a 1-D sequence")
    return arr


def _i64p(arr: np.ndarray):
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def dark_scale(light_exposure_s, dark_exposure_s, policy: dict | None, has_dark: bool) -> float:
    """Dark scaling factor of the CPU path (src/glass/cpu/calibration.py:63-66):
    FP64 ratio rounded to f32, 1.0 without scaling, without a dark exposure or
    with a zero dark exposure."""
    pol = dict(_DEFAULT_POLICY)
    pol.update(policy or {})
    if not has_dark:
        return 1.0
    if pol["dark_scaling_enabled"] and dark_exposure_s not in (None, 0):
        return float(np.float32(float(light_exposure_s) / float(dark_exposure_s)))
    return 1.0


def invert_affine(matrix) -> np.ndarray:
    """FP64 inverse of a 3x3 moving->reference matrix, as the CPU warp does
    (src/glass/engine/warp.py:169-176).  Singular matrices raise ValueError."""
    m = np.asarray(matrix, dtype=np.float64).reshape(3, 3)
    if not np.all(np.isfinite(m)):
        raise ValueError("warp matrix must be finite")
    if abs(float(np.linalg.det(m))) <= 1.0e-20:
        raise ValueError("warp matrix is singular")
    try:
        return np.linalg.inv(m)
    except np.linalg.LinAlgError as exc:
        raise ValueError("warp matrix is singular") from exc


class ResidentCalibratedStack:
    """F x H x W float32 calibrated light stack resident in HBM."""

    def __init__(self, frame_count: int, height: int, width: int):
        handle = ctypes.c_void_p()
        N.call("gh_stack_create", int(frame_count), int(height), int(width), ctypes.byref(handle))
        self._h = handle
        self._frame_count = int(frame_count)
        self._height = int(height)
        self._width = int(width)
        self._has = {"bias": False, "dark": False, "flat": False}
        self._pending = []       # queued asynchronous host calibrations
        self._fits_queue = None  # FITS completion queue between begin and finish

    # ------------------------------------------------------------ lifecycle
    def close(self) -> None:
        if getattr(self, "_fits_queue", None) is not None:
            self.cancel_fits_u16be_bzero_paths_completion_queue_read()
        if getattr(self, "_pending", None):
            self._finish()
        if getattr(self, "_h", None) is not None and self._h.value:
            N.call("gh_stack_destroy", self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _info(self) -> np.ndarray:
        out = (ctypes.c_int64 * 8)()
        N.call("gh_stack_info", self._h, out)
        return np.frombuffer(out, dtype=np.int64).copy()

    # ----------------------------------------------------------- properties
    @property
    def frame_count(self) -> int:
        return self._frame_count

    @property
    def height(self) -> int:
        return self._height

    @property
    def width(self) -> int:
        return self._width

    @property
    def pixels_per_frame(self) -> int:
        return self._height * self._width

    @property
    def loaded_count(self) -> int:
        return int(self._info()[4])

    @property
    def bytes_allocated(self) -> int:
        return int(self._info()[5])

    @property
    def warp_coverage_frame_count(self) -> int:
        return int(self._info()[6])

    @property
    def warp_scratch_bytes(self) -> int:
        return int(self._info()[7])

    @property
    def warp_copy_mode(self) -> str:
        return "slot_swap"

    def frame_device_ptr(self, index: int) -> int:
        """Device address of frame `index` (valid until that frame is warped)."""
        ptr = ctypes.c_void_p()
        N.call("gh_stack_frame_ptr", self._h, int(index), ctypes.byref(ptr))
        return int(ptr.value or 0)

    def frame_has_nonfinite(self, index: int) -> bool:
        out = ctypes.c_int64()
        N.call("gh_stack_frame_nonfinite", self._h, int(index), ctypes.byref(out))
        return bool(out.value)

    # -------------------------------------------------------------- masters
    def set_calibration_masters(self, bias=None, dark=None, flat=None) -> None:
        shape = (self._height, self._width)
        keep = []
        ptrs = []
        for name, value in (("bias", bias), ("dark", dark), ("flat", flat)):
            if value is None:
                ptrs.append(None)
                self._has[name] = False
                continue
            ptr, ref = _buffer(value, np.float32, shape, name)
            keep.append(ref)
            ptrs.append(ptr)
            self._has[name] = True
        N.call("gh_set_masters", self._h, *ptrs)

    # ---------------------------------------------------------- upload/down
    def upload_calibrated_frame(self, index: int, frame) -> None:
        ptr, ref = _buffer(frame, np.float32, (self._height, self._width), "frame")
        N.call("gh_upload_frame", self._h, int(index), ptr)

    def download_frame_tile(self, index: int, x0: int, y0: int, x1: int, y1: int) -> np.ndarray:
        out = self.download_frames_tile([index], x0, y0, x1, y1)
        return out[0]

    def download_frames_tile(self, indices, x0: int, y0: int, x1: int, y1: int) -> np.ndarray:
        idx = _index_array(indices)
        if idx.size == 0:
            raise ValueError("indices must contain at least one frame index")
        if x0 < 0 or y0 < 0 or x0 >= x1 or y0 >= y1 or x1 > self._width or y1 > self._height:
            raise IndexError("resident frame tile bounds are invalid")
        out = np.empty((idx.size, int(y1) - int(y0), int(x1) - int(x0)), dtype=np.float32)
        N.call("gh_download_tile", self._h, _i64p(idx), int(idx.size), int(x0), int(y0), int(x1), int(y1),
               ctypes.c_void_p(out.ctypes.data))
        return out

    def download_frame(self, index: int) -> np.ndarray:
        return self.download_frame_tile(index, 0, 0, self._width, self._height)

    # ------------------------------------------------------------ calibrate
    def _policy(self, policy: dict | None, accumulate_stats: bool = False) -> N.CalibPolicy:
        pol = dict(_DEFAULT_POLICY)
        pol.update(policy or {})
        return N.CalibPolicy(
            int(bool(pol["master_dark_includes_bias"])),
            float(pol["flat_floor"]),
            float(pol["pedestal"] or 0.0),
            int(bool(accumulate_stats)),
        )

    def _calibrate(self, indices, raw_ptr, stride, fmt, light_exposures, dark_exposures, policy, accumulate_stats,
                   stream=None):
        idx = _index_array(indices)
        lights = list(light_exposures)
        if isinstance(dark_exposures, (list, tuple, np.ndarray)):
            darks = list(dark_exposures)
        else:
            darks = [dark_exposures] * idx.size
        if len(lights) != idx.size or len(darks) != idx.size:
            raise ValueError("light_exposures and dark_exposures must match indices")
        scales = np.asarray(
            [dark_scale(lights[k], darks[k], policy, self._has["dark"]) for k in range(idx.size)], dtype=np.float32
        )
        pol = self._policy(policy, accumulate_stats)
        N.call(
            "gh_calibrate_frames", self._h, _i64p(idx), int(idx.size), raw_ptr, int(stride), int(fmt),
            scales.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(pol),
            ctypes.c_void_p(stream) if stream else None,
        )

    def calibrate_frame(self, index, light, light_exposure_s, dark_exposure_s=None, policy=None) -> None:
        """calibrate_frame (native_bindings.cpp:5979) -- f32 light frame."""
        ptr, ref = _buffer(light, np.float32, (self._height, self._width), "light")
        self._calibrate([index], ptr, self.pixels_per_frame * 4, N.GH_FMT_F32, [light_exposure_s],
                        [dark_exposure_s], policy, False)

    def calibrate_frame_timed(self, index, light, light_exposure_s, dark_exposure_s=None, policy=None) -> dict:
        start = time.perf_counter()
        self.calibrate_frame(index, light, light_exposure_s, dark_exposure_s, policy)
        total = time.perf_counter() - start
        return {"schema_version": 1, "h2d_mode": "pageable", "total_s": total}

    def calibrate_frames_host_async_timed(self, indices, lights, light_exposures, dark_exposures, policy=None) -> dict:
        start = time.perf_counter()
        for k, index in enumerate(_index_array(indices)):
            darks = dark_exposures[k] if isinstance(dark_exposures, (list, tuple, np.ndarray)) else dark_exposures
            self.calibrate_frame(int(index), lights[k], light_exposures[k], darks, policy)
        return {"schema_version": 1, "frame_count": len(lights), "total_s": time.perf_counter() - start}

    def calibrate_frames_fits_u16be_bzero_host_async_multistream_callback_release_timed(
        self, indices, raw_lights, light_exposures, dark_exposures, stream_count, wave_frames, release_callback,
        policy=None,
    ) -> dict:
        """Raw FITS BITPIX=16/BZERO=32768 big-endian light frames
        (native_bindings.cpp:6839; kernel calibration_kernels.cu:110-153).
        Frames are queued asynchronously (uploads overlap calibration);
        release_callback(k) runs once frame k's host buffer has been consumed,
        in order."""
        import torch

        idx = _index_array(indices)
        start = time.perf_counter()
        stream = torch.cuda.Stream()
        pending = []
        for k, index in enumerate(idx):
            ptr, ref = _buffer(raw_lights[k], np.uint8, (self.pixels_per_frame * 2,), "raw_light")
            darks = dark_exposures[k] if isinstance(dark_exposures, (list, tuple, np.ndarray)) else dark_exposures
            self._calibrate([int(index)], ptr, self.pixels_per_frame * 2, N.GH_FMT_FITS_U16BE_BZERO,
                            [light_exposures[k]], [darks], policy, False, stream.cuda_stream)
            ev = torch.cuda.Event()
            ev.record(stream)
            pending.append((k, ev, ref))
            while pending and pending[0][1].query():
                done = pending.pop(0)
                if release_callback is not None:
                    release_callback(int(done[0]))
        for k, ev, _ in pending:
            ev.synchronize()
            if release_callback is not None:
                release_callback(int(k))
        stream.synchronize()
        return {"schema_version": 1, "frame_count": int(idx.size), "stream_count": int(stream_count),
                "wave_frames": int(wave_frames), "total_s": time.perf_counter() - start}

    # ---- asynchronous / multistream forms of the reference's calibration API
    def _queue_host(self, indices, lights, light_exposures, dark_exposures, policy, raw_format):
        from .staging import calibrate_host_frames

        idx = [int(i) for i in _index_array(indices)]
        if len(lights) != len(idx):
            raise ValueError("lights must match indices")
        t0 = time.perf_counter()
        keep, stream = calibrate_host_frames(self, idx, lights, list(light_exposures), dark_exposures, policy,
                                             raw_format)
        self._pending.append((keep, stream, idx, t0))
        return idx, t0

    def _finish(self) -> dict:
        t = time.perf_counter()
        frames = 0
        t0 = t
        for keep, stream, idx, start in self._pending:
            stream.synchronize()
            frames += len(idx)
            t0 = min(t0, start)
        pending = bool(self._pending)
        self._pending = []
        now = time.perf_counter()
        return {"schema_version": 1, "pending": pending, "frame_count": frames, "wait_sync_s": now - t,
                "total_s": now - t0 if pending else 0.0}

    def calibrate_frame_pinned_async(self, index, light, light_exposure_s, dark_exposure_s=None, policy=None):
        """calibrate_frame_pinned_async (native_bindings.cpp:6013): queued;
        finish_pending_calibration_timed() waits for it."""
        self._queue_host([index], [light], [light_exposure_s], dark_exposure_s, policy, "float32")

    def calibrate_frame_pinned_async_timed(self, index, light, light_exposure_s, dark_exposure_s=None, policy=None):
        _, t0 = self._queue_host([index], [light], [light_exposure_s], dark_exposure_s, policy, "float32")
        return {"schema_version": 1, "mode": "pinned_async", "enqueue_s": time.perf_counter() - t0}

    def calibrate_frame_host_async_timed(self, index, light, light_exposure_s, dark_exposure_s=None, policy=None):
        """calibrate_frame_host_async_timed (native_bindings.cpp:6047)."""
        _, t0 = self._queue_host([index], [light], [light_exposure_s], dark_exposure_s, policy, "float32")
        return {"schema_version": 1, "mode": "host_async", "enqueue_s": time.perf_counter() - t0}

    def calibrate_frames_host_async_timed(self, indices, lights, light_exposures, dark_exposures, policy=None):
        """calibrate_frames_host_async_timed (native_bindings.cpp:6065):
        every frame queued, then waited for."""
        idx, t0 = self._queue_host(indices, lights, light_exposures, dark_exposures, policy, "float32")
        out = self._finish()
        out.update(mode="host_async_batch", frame_count=len(idx), total_s=time.perf_counter() - t0)
        return out

    def calibrate_frames_host_async_multistream_timed(self, indices, lights, light_exposures, dark_exposures,
                                                      stream_count, policy=None):
        """calibrate_frames_host_async_multistream_timed (native_bindings.cpp:6179):
        the uploads already overlap calibration through the library's ring,
        so `stream_count` does not change the result."""
        out = self.calibrate_frames_host_async_timed(indices, lights, light_exposures, dark_exposures, policy)
        out.update(mode="host_async_multistream", stream_count=int(stream_count))
        return out

    def calibrate_frames_host_async_multistream_h2d_release_timed(self, indices, lights, light_exposures,
                                                                  dark_exposures, stream_count, policy=None):
        """Queued like calibrate_frames_host_async_multistream_timed; returns
        once the work is queued (the host buffers are held until
        finish_pending_calibration_timed())."""
        idx, t0 = self._queue_host(indices, lights, light_exposures, dark_exposures, policy, "float32")
        return {"schema_version": 1, "mode": "host_async_multistream_h2d_release", "frame_count": len(idx),
                "stream_count": int(stream_count), "enqueue_s": time.perf_counter() - t0}

    def calibrate_frames_host_async_multistream_callback_release_timed(
            self, indices, lights, light_exposures, dark_exposures, stream_count, wave_frames, release_callback,
            policy=None):
        """float32 form of the FITS callback-release entry: release_callback(k)
        runs, in order, once frame k's host buffer has been consumed."""
        import torch

        idx = _index_array(indices)
        start = time.perf_counter()
        stream = torch.cuda.Stream()
        darks = dark_exposures if isinstance(dark_exposures, (list, tuple, np.ndarray)) else [dark_exposures] * len(idx)
        pending = []
        for k, index in enumerate(idx):
            ptr, ref = _buffer(lights[k], np.float32, (self.height, self.width), "light")
            self._calibrate([int(index)], ptr, self.pixels_per_frame * 4, N.GH_FMT_F32, [light_exposures[k]],
                            [darks[k]], policy, False, stream.cuda_stream)
            ev = torch.cuda.Event()
            ev.record(stream)
            pending.append((k, ev, ref))
            while pending and pending[0][1].query():
                done = pending.pop(0)
                if release_callback is not None:
                    release_callback(int(done[0]))
        for k, ev, _ in pending:
            ev.synchronize()
            if release_callback is not None:
                release_callback(int(k))
        stream.synchronize()
        return {"schema_version": 1, "frame_count": int(idx.size), "stream_count": int(stream_count),
                "wave_frames": int(wave_frames), "total_s": time.perf_counter() - start}

    def finish_pending_calibration_timed(self) -> dict:
        """finish_pending_calibration_timed (native_bindings.cpp:6510)."""
        return self._finish()

    def calibrate_frames_fits_u16be_bzero_paths_multistream_timed(self, indices, paths, data_offsets, byte_counts,
                                                                  light_exposures, dark_exposures, stream_count,
                                                                  wave_frames, policy=None):
        """FITS light paths with known data offsets: read into the pinned
        ring and calibrated as they arrive."""
        from .staging import FitsCompletionQueue

        q = FitsCompletionQueue(self, _index_array(indices), paths, data_offsets, byte_counts,
                                max(2, int(wave_frames)), 1)
        out = q.calibrate(light_exposures, dark_exposures, policy)
        out.update(stream_count=int(stream_count), wave_frames=int(wave_frames))
        return out

    def calibrate_frames_fits_u16be_bzero_paths_completion_queue_timed(
            self, indices, paths, data_offsets, byte_counts, light_exposures, dark_exposures, stream_count,
            queue_buffer_count, worker_count, policy=None):
        """The FITS completion queue in one call: `worker_count` reader threads,
        `queue_buffer_count` pinned buffers, calibration in completion order."""
        self.begin_fits_u16be_bzero_paths_completion_queue_read_timed(
            indices, paths, data_offsets, byte_counts, stream_count, queue_buffer_count, worker_count, policy)
        return self.finish_fits_u16be_bzero_paths_completion_queue_calibration_timed(
            light_exposures, dark_exposures, stream_count, policy)

    def begin_fits_u16be_bzero_paths_completion_queue_read_timed(self, indices, paths, data_offsets, byte_counts,
                                                                 stream_count, queue_buffer_count, worker_count,
                                                                 policy=None):
        """Start the reader threads of the FITS completion queue."""
        from .staging import FitsCompletionQueue

        if self._fits_queue is not None:
            raise RuntimeError("a FITS completion queue read is already in progress")
        t0 = time.perf_counter()
        self._fits_queue = FitsCompletionQueue(self, _index_array(indices), paths, data_offsets, byte_counts,
                                               int(queue_buffer_count), int(worker_count))
        return {"schema_version": 1, "frame_count": len(self._fits_queue.indices),
                "queue_buffer_count": len(self._fits_queue.buffers), "worker_count": int(worker_count),
                "begin_s": time.perf_counter() - t0}

    def finish_fits_u16be_bzero_paths_completion_queue_calibration_timed(self, light_exposures, dark_exposures,
                                                                         stream_count, policy=None):
        """Calibrate the queued reads in completion order and wait for them."""
        if self._fits_queue is None:
            raise RuntimeError("no FITS completion queue read is in progress")
        q, self._fits_queue = self._fits_queue, None
        out = q.calibrate(light_exposures, dark_exposures, policy)
        out["stream_count"] = int(stream_count)
        return out

    def cancel_fits_u16be_bzero_paths_completion_queue_read(self) -> None:
        """Stop the reader threads of a queue that will not be calibrated."""
        if self._fits_queue is not None:
            q, self._fits_queue = self._fits_queue, None
            q.cancel()

    def calibrate_frames_host(self, indices, raw, raw_format: str, light_exposures, dark_exposures, policy=None,
                              frame_stride_bytes: int | None = None, accumulate_stats: bool = False,
                              stream: int | None = None) -> None:
        """Calibrate frames from one host buffer holding them back to back
        (numpy array or CPU tensor; pinned memory gives the full link rate).
        Uploads run through the device ring, overlapping calibration.  With a
        stream the call is asynchronous and `raw` must outlive that stream's
        work."""
        fmt = {"float32": N.GH_FMT_F32, "uint16": N.GH_FMT_U16, "fits_u16be_bzero": N.GH_FMT_FITS_U16BE_BZERO}[raw_format]
        sample = 4 if fmt == N.GH_FMT_F32 else 2
        stride = int(frame_stride_bytes or self.pixels_per_frame * sample)
        count = len(_index_array(indices))
        if _is_torch_cuda(raw):
            raise ValueError("calibrate_frames_host expects host memory; use calibrate_frames_device")
        if type(raw).__module__.startswith("torch"):
            if not raw.is_contiguous():
                raise ValueError("raw host tensor must be contiguous")
            nbytes, ptr = raw.numel() * raw.element_size(), ctypes.c_void_p(raw.data_ptr())
        else:
            arr = np.ascontiguousarray(raw)
            if stream and arr is not raw:
                raise ValueError("an asynchronous host calibration needs a contiguous raw buffer (no copy)")
            nbytes, ptr, raw = arr.nbytes, ctypes.c_void_p(arr.ctypes.data), arr
        if count and nbytes < (count - 1) * stride + self.pixels_per_frame * sample:
            raise ValueError("raw host buffer is smaller than the frames it should hold")
        self._calibrate(indices, ptr, stride, fmt, light_exposures, dark_exposures, policy, accumulate_stats, stream)

    def calibrate_fits_files(self, indices, paths, light_exposures, dark_exposures, policy=None,
                             ring_frames: int = 4) -> dict:
        """Read BITPIX=16/BZERO=32768 FITS lights straight into a ring of
        pinned host buffers and calibrate them as they arrive: the read of
        file k + 1 overlaps the upload and calibration of file k (the
        reference's pinned ring, resident_cuda.py:1159-1897)."""
        import torch

        from . import fits as F

        idx = _index_array(indices)
        paths = list(paths)
        if len(paths) != idx.size:
            raise ValueError("paths must match indices")
        lights = list(light_exposures)
        darks = (list(dark_exposures) if isinstance(dark_exposures, (list, tuple, np.ndarray))
                 else [dark_exposures] * idx.size)
        specs = [F.simple_fits_image_spec(p) for p in paths]
        for spec in specs:
            elig = F.u16_gpu_eligibility(spec, (self.height, self.width))
            if not elig["eligible"]:
                raise ValueError(f"{spec.path}: not a raw u16 light for this stack ({elig['reason']})")
        nbytes = self.pixels_per_frame * 2
        slots = max(1, min(int(ring_frames), int(idx.size)))
        ring = [torch.empty(nbytes, dtype=torch.uint8, pin_memory=True) for _ in range(slots)]
        done = [None] * slots
        stream = torch.cuda.Stream()
        t0 = time.perf_counter()
        read_s = 0.0
        for k in range(idx.size):
            slot = k % slots
            if done[slot] is not None:
                done[slot].synchronize()
            t = time.perf_counter()
            F.read_u16be_raw_into(specs[k], ring[slot].numpy())
            read_s += time.perf_counter() - t
            self._calibrate([int(idx[k])], ctypes.c_void_p(ring[slot].data_ptr()), nbytes,
                            N.GH_FMT_FITS_U16BE_BZERO, [lights[k]], [darks[k]], policy, False, stream.cuda_stream)
            ev = torch.cuda.Event()
            ev.record(stream)
            done[slot] = ev
        stream.synchronize()
        total = time.perf_counter() - t0
        return {"schema_version": 1, "frame_count": int(idx.size), "ring_frames": slots,
                "bytes_read": int(nbytes * idx.size), "file_read_s": read_s, "total_s": total}

    def calibrate_frames_device(self, indices, raw, raw_format: str, light_exposures, dark_exposures, policy=None,
                                frame_stride_bytes: int | None = None, accumulate_stats: bool = False,
                                stream: int | None = None) -> None:
        """Resident-raw extension: calibrate frames whose raw samples already
        sit in HBM (a torch tensor), without a host round trip."""
        fmt = {"float32": N.GH_FMT_F32, "uint16": N.GH_FMT_U16, "fits_u16be_bzero": N.GH_FMT_FITS_U16BE_BZERO}[raw_format]
        sample = 4 if fmt == N.GH_FMT_F32 else 2
        stride = int(frame_stride_bytes or self.pixels_per_frame * sample)
        if not _is_torch_cuda(raw):
            raise ValueError("calibrate_frames_device expects a GPU tensor")
        self._calibrate(indices, ctypes.c_void_p(raw.data_ptr()), stride, fmt, light_exposures, dark_exposures,
                        policy, accumulate_stats, stream)

    def calibration_stats(self, index: int) -> dict:
        s, ss, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        N.call("gh_calibration_stats", self._h, int(index), ctypes.byref(s), ctypes.byref(ss), ctypes.byref(n))
        count = int(n.value)
        mean = s.value / count if count else 0.0
        var = max(ss.value / count - mean * mean, 0.0) if count else 0.0
        return {"sum": s.value, "sum_sq": ss.value, "valid_pixels": count, "mean": mean, "std": math.sqrt(var)}

    def normalize_frames(self, indices, offsets, divisors, stream: int | None = None) -> None:
        """x = (x - offset) / divisor per frame, f32 (per-flat normalisation)."""
        idx = _index_array(indices)
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.float32).reshape(-1))
        div = np.ascontiguousarray(np.asarray(divisors, dtype=np.float32).reshape(-1))
        if off.size != idx.size or div.size != idx.size:
            raise ValueError("offsets and divisors must match indices")
        N.call("gh_normalize_frames", self._h, _i64p(idx), int(idx.size),
               off.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), div.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
               ctypes.c_void_p(stream) if stream else None)

    # ----------------------------------------------------------------- warp
    def _warp(self, indices, matrices, interpolation, fill, clamping_threshold, track_coverage, skip_nonfinite,
              stream=None) -> dict:
        idx = _index_array(indices)
        mats = np.asarray(matrices, dtype=np.float64).reshape(-1, 3, 3)
        if mats.shape[0] != idx.size:
            raise ValueError("matrices must have shape (len(indices), 3, 3)")
        inv = np.ascontiguousarray(np.stack([invert_affine(m) for m in mats]) if idx.size else np.zeros((0, 3, 3)))
        affine = bool(idx.size) and bool(np.all(inv[:, 2, 0] == 0.0) and np.all(inv[:, 2, 1] == 0.0)
                                         and np.all(inv[:, 2, 2] == 1.0))
        flags = (N.GH_WARP_TRACK_COVERAGE if track_coverage else 0) | (N.GH_WARP_AFFINE if affine else 0)
        if skip_nonfinite:
            flags |= N.GH_WARP_SKIP_NONFINITE_TAPS
        start = time.perf_counter()
        N.call("gh_warp_frames", self._h, _i64p(idx), int(idx.size),
               inv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), int(interpolation), float(fill),
               float(clamping_threshold), int(flags), ctypes.c_void_p(stream) if stream else None)
        if stream is None:
            N.call("gh_synchronize", self._h)
        return {
            "schema_version": 1,
            "frame_count": int(idx.size),
            "interpolation": "lanczos3" if interpolation == N.GH_INTERP_LANCZOS3 else "bilinear",
            "track_coverage": bool(track_coverage),
            "copy_mode": "slot_swap",
            "total_s": time.perf_counter() - start,
        }

    def apply_matrix_bilinear_frame(self, index, matrix, fill=math.nan) -> None:
        self._warp([index], [matrix], N.GH_INTERP_BILINEAR, fill, -1.0, False, True)

    def apply_matrix_lanczos3_frame(self, index, matrix, fill=math.nan, clamping_threshold=-1.0) -> None:
        self._warp([index], [matrix], N.GH_INTERP_LANCZOS3, fill, clamping_threshold, False, True)

    def apply_matrix_bilinear_frames(self, indices, matrices, fill=math.nan, max_chunk_capacity_frames=0,
                                     track_coverage=True) -> dict:
        return self._warp(indices, matrices, N.GH_INTERP_BILINEAR, fill, -1.0, track_coverage, True)

    def apply_matrix_lanczos3_frames(self, indices, matrices, fill=math.nan, clamping_threshold=-1.0,
                                     max_chunk_capacity_frames=0, track_coverage=True) -> dict:
        return self._warp(indices, matrices, N.GH_INTERP_LANCZOS3, fill, clamping_threshold, track_coverage, True)

    def apply_matrix_frames_cpu_semantics(self, indices, matrices, interpolation="lanczos3", fill=0.0,
                                          track_coverage=True, stream=None) -> dict:
        """Matrix warp with the CPU-path tap semantics (engine/warp.py:121-152:
        non-finite taps are not skipped, sinc threshold 1e-8)."""
        interp = N.GH_INTERP_LANCZOS3 if interpolation == "lanczos3" else N.GH_INTERP_BILINEAR
        return self._warp(indices, matrices, interp, fill, -1.0, track_coverage, False, stream)

    def apply_translation_frame(self, index: int, dx: int, dy: int, fill: float = math.nan,
                                stream: int | None = None) -> None:
        """apply_translation_frame (native_bindings.cpp:8778): whole-pixel shift
        in place, `fill` outside, geometric coverage accumulated."""
        if int(dx) != dx or int(dy) != dy:
            raise TypeError("apply_translation_frame takes integer offsets")
        N.call("gh_translate_frame", self._h, int(index), N.GH_TRANSLATE_INTEGER, float(int(dx)), float(int(dy)),
               float(fill), ctypes.c_void_p(stream) if stream else None)

    def apply_translation_bilinear_frame(self, index: int, dx: float, dy: float, fill: float = math.nan,
                                         stream: int | None = None) -> None:
        """apply_translation_bilinear_frame (native_bindings.cpp:8808): bilinear
        sample at (x - dx, y - dy) in float32, in place."""
        N.call("gh_translate_frame", self._h, int(index), N.GH_INTERP_BILINEAR, float(np.float32(dx)),
               float(np.float32(dy)), float(fill), ctypes.c_void_p(stream) if stream else None)

    def _ncc(self, reference_index, moving_index, subpixel, rx, ry, cdx, cdy, step, stride):
        dx, dy, score, n = ctypes.c_float(), ctypes.c_float(), ctypes.c_float(), ctypes.c_int64()
        N.call("gh_estimate_translation", self._h, int(reference_index), int(moving_index), int(subpixel), int(rx),
               int(ry), float(cdx), float(cdy), float(step), int(stride), ctypes.byref(dx), ctypes.byref(dy),
               ctypes.byref(score), ctypes.byref(n))
        return dx.value, dy.value, score.value, int(n.value)

    def estimate_translation_to_reference(self, reference_index: int, moving_index: int, max_shift_x: int,
                                          max_shift_y: int, sample_stride: int = 1) -> dict:
        """Integer translation (moving -> reference) maximising the normalised
        cross-correlation (native_bindings.cpp:10418-10485)."""
        dx, dy, score, n = self._ncc(reference_index, moving_index, 0, max_shift_x, max_shift_y, 0.0, 0.0, 1.0,
                                     sample_stride)
        return {"dx": int(dx), "dy": int(dy), "score": score, "search_count": n, "sample_stride": int(sample_stride),
                "reference_index": int(reference_index), "moving_index": int(moving_index),
                "model": "resident_translation_integer_ncc"}

    def estimate_translation_subpixel_to_reference(self, reference_index: int, moving_index: int, center_dx: float,
                                                   center_dy: float, radius_steps: int, step: float,
                                                   sample_stride: int = 1) -> dict:
        """Subpixel refinement of a translation on a (2r+1)^2 grid of float32
        offsets around the centre (native_bindings.cpp:10487-10568)."""
        dx, dy, score, n = self._ncc(reference_index, moving_index, 1, radius_steps, radius_steps, center_dx,
                                     center_dy, step, sample_stride)
        return {"dx": dx, "dy": dy, "score": score, "candidate_count": n, "center_dx": float(np.float32(center_dx)),
                "center_dy": float(np.float32(center_dy)), "radius_steps": int(radius_steps),
                "step": float(np.float32(step)), "sample_stride": int(sample_stride),
                "reference_index": int(reference_index), "moving_index": int(moving_index),
                "model": "resident_translation_subpixel_ncc"}

    def reset_warp_coverage(self) -> None:
        N.call("gh_reset_warp_coverage", self._h)

    def accumulate_full_warp_coverage_frame(self) -> None:
        N.call("gh_accumulate_full_warp_coverage", self._h)

    def warp_coverage_map(self) -> np.ndarray:
        out = np.empty((self._height, self._width), dtype=np.float32)
        N.call("gh_warp_coverage_map", self._h, ctypes.c_void_p(out.ctypes.data))
        return out

    # ---------------------------------------------------------------- stars
    def detect_stars(self, index: int, threshold_sigma: float = 5.0, max_stars: int = 500, *, background=None,
                     noise=None, window_radius: int = 4, min_separation_px: float = 2.0, origin_x: int = 0,
                     origin_y: int = 0) -> list:
        """detect_stars (src/glass/cpu/star_detect.py:121-168) on resident frame
        `index`, no host copy of the frame."""
        from .registration import detect_resident

        stars, _bg = detect_resident(self._h, index, threshold_sigma, max_stars, background, noise, window_radius,
                                     min_separation_px, origin_x, origin_y)
        return stars

    def refine_affine(self, reference_index: int, moving_index: int, matrix, max_iterations: int = 20,
                      tolerance_px: float = 1.0e-4):
        """Gauss-Newton photometric refinement of the moving -> reference affine
        matrix of two resident frames (see glass_amd.refine.refine_affine)."""
        from .refine import refine_affine

        return refine_affine(self, reference_index, moving_index, matrix, max_iterations, tolerance_px)

    # ---------------------------------------------------------------- stats
    def frame_global_stats(self, index: int, exact_median: bool = False) -> dict:
        st = N.FrameStatistics()
        N.call("gh_frame_stats", self._h, int(index), int(bool(exact_median)), ctypes.byref(st))
        out = {
            "mean": st.mean,
            "std": st.std,
            "valid_pixels": int(st.valid_pixels),
            "total_pixels": int(st.total_pixels),
            "nonfinite_pixels": int(st.nonfinite_pixels),
            "model": "resident_global_mean_std",
            "min": float(st.min),
            "max": float(st.max),
        }
        if exact_median:
            out["median"] = st.median
            out["mad"] = st.mad
        return out

    def frames_global_stats(self, indices) -> list[dict]:
        """frame_global_stats of many frames in one pass (same values as the
        per-frame call without exact_median)."""
        idx = _index_array(indices)
        arr = (N.FrameStatistics * max(int(idx.size), 1))()
        N.call("gh_frames_stats", self._h, _i64p(idx), int(idx.size), arr)
        return [{
            "mean": st.mean, "std": st.std, "valid_pixels": int(st.valid_pixels),
            "total_pixels": int(st.total_pixels), "nonfinite_pixels": int(st.nonfinite_pixels),
            "model": "resident_global_mean_std", "min": float(st.min), "max": float(st.max),
        } for st in arr[:idx.size]]

    # ------------------------------------------------- local normalisation
    def _grid_shape(self, tile_height: int, tile_width: int) -> tuple[int, int]:
        if int(tile_height) <= 0 or int(tile_width) <= 0:
            raise ValueError("tile dimensions must be positive")
        return ((self.height + int(tile_height) - 1) // int(tile_height),
                (self.width + int(tile_width) - 1) // int(tile_width))

    def _pair_grid_stats(self, reference_index: int, sources: np.ndarray, tile_height: int, tile_width: int):
        rows, cols = self._grid_shape(tile_height, tile_width)
        n = int(sources.size)
        out = {k: np.zeros((n, rows, cols), dtype=np.float32)
               for k in ("source_mean", "source_std", "reference_mean", "reference_std")}
        valid = np.zeros((n, rows, cols), dtype=np.uint64)
        fp = ctypes.POINTER(ctypes.c_float)
        t0 = time.perf_counter()
        N.call("gh_pair_grid_stats", self._h, int(reference_index), _i64p(sources), n, int(tile_height),
               int(tile_width), out["source_mean"].ctypes.data_as(fp), out["source_std"].ctypes.data_as(fp),
               out["reference_mean"].ctypes.data_as(fp), out["reference_std"].ctypes.data_as(fp),
               valid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        return out, valid, rows, cols, time.perf_counter() - t0

    def frame_pair_grid_stats(self, reference_index: int, source_index: int, tile_height: int,
                              tile_width: int) -> dict:
        """Per-tile float32 mean / std of a (source, reference) pair over the
        pixels finite in both (native_bindings.cpp:10645-10784)."""
        src = np.asarray([int(source_index)], dtype=np.int64)
        out, valid, rows, cols, _ = self._pair_grid_stats(reference_index, src, tile_height, tile_width)
        result = {k: v[0] for k, v in out.items()}
        result.update({
            "valid_pixels": valid[0], "grid_rows": rows, "grid_cols": cols, "tile_height": int(tile_height),
            "tile_width": int(tile_width), "valid_pixel_total": int(valid[0].sum(dtype=np.uint64)),
            "model": "resident_grid_pair_mean_std",
        })
        return result

    def frame_pair_grid_stats_batch(self, reference_index: int, source_indices, tile_height: int,
                                    tile_width: int) -> dict:
        """frame_pair_grid_stats for many sources in one pass over the
        reference (native_bindings.cpp:10786-11021)."""
        arr = np.asarray(source_indices)
        if arr.ndim != 1:
            raise ValueError("source_indices must be a one-dimensional array")
        src = np.ascontiguousarray(arr, dtype=np.int64)
        t0 = time.perf_counter()
        out, valid, rows, cols, call_s = self._pair_grid_stats(reference_index, src, tile_height, tile_width)
        frames = []
        for pos in range(src.size):
            frame = {k: v[pos] for k, v in out.items()}
            frame.update({
                "valid_pixels": valid[pos], "grid_rows": rows, "grid_cols": cols,
                "tile_height": int(tile_height), "tile_width": int(tile_width),
                "valid_pixel_total": int(valid[pos].sum(dtype=np.uint64)),
                "model": "resident_grid_pair_mean_std", "source_index": int(src[pos]),
                "reference_index": int(reference_index), "batch_position": pos,
            })
            frames.append(frame)
        grid = rows * cols
        return {
            "schema_version": 1, "model": "resident_grid_pair_mean_std_batch",
            "batch_model": "single_kernel_source_frame_tile_grid", "reference_index": int(reference_index),
            "source_count": int(src.size), "grid_rows": rows, "grid_cols": cols, "grid_count": grid,
            "tile_height": int(tile_height), "tile_width": int(tile_width),
            "source_indices": arr.astype(np.int32, copy=False),
            "allocation_s": 0.0, "index_upload_s": 0.0, "kernel_enqueue_s": call_s, "sync_s": 0.0,
            "download_s": 0.0, "total_s": time.perf_counter() - t0,
            "download_bytes": int(src.size) * grid * (4 * 8 + 8), "index_bytes": int(src.size) * 4,
            "frames": frames,
        }

    def _apply_normalization(self, indices: np.ndarray, scales, offsets, tile_height: int, tile_width: int,
                             stream: int | None) -> float:
        count = int(indices.size)
        rows, cols = self._grid_shape(tile_height, tile_width)
        n = count * rows * cols
        if not _is_torch_cuda(scales):
            scales = np.asarray(scales, dtype=np.float32).reshape(-1)
        if not _is_torch_cuda(offsets):
            offsets = np.asarray(offsets, dtype=np.float32).reshape(-1)
        sp, keep_s = _buffer(scales, np.float32, (n,), "scales")
        op, keep_o = _buffer(offsets, np.float32, (n,), "offsets")
        t0 = time.perf_counter()
        N.call("gh_apply_normalization", self._h, _i64p(indices), count, sp, op, int(tile_height), int(tile_width),
               ctypes.c_void_p(stream) if stream else None)
        del keep_s, keep_o
        return time.perf_counter() - t0

    def apply_global_normalization_frame(self, index: int, scale: float, offset: float,
                                         stream: int | None = None) -> dict:
        """x * scale + offset in place over the whole frame, float32
        coefficients (native_bindings.cpp:11023-11052)."""
        t0 = time.perf_counter()
        idx = np.asarray([int(index)], dtype=np.int64)
        call_s = self._apply_normalization(idx, np.float32([[[scale]]]), np.float32([[[offset]]]), self.height,
                                           self.width, stream)
        return {
            "schema_version": 1, "mode": "in_place_device_update", "model": "resident_global_mean_std",
            "frame_index": int(index), "frame_bytes": self.pixels_per_frame * 4, "temporary_output_bytes": 0,
            "device_to_device_copy_s": 0.0, "kernel_enqueue_s": call_s, "sync_s": 0.0,
            "total_s": time.perf_counter() - t0,
        }

    def apply_global_normalization_frames(self, indices, scales, offsets, stream: int | None = None) -> dict:
        """apply_global_normalization_frame for many frames in one launch per
        batch of distinct frames: frame indices[k] becomes x * float32(scales[k])
        + float32(offsets[k])."""
        idx = _index_array(indices)
        sc = np.asarray(scales, dtype=np.float32).reshape(-1)
        of = np.asarray(offsets, dtype=np.float32).reshape(-1)
        if sc.size != idx.size or of.size != idx.size:
            raise ValueError("scales and offsets must have one entry per frame")
        t0 = time.perf_counter()
        call_s = self._apply_normalization(idx, sc, of, self.height, self.width, stream) if idx.size else 0.0
        return {"schema_version": 1, "mode": "in_place_device_update_batch", "model": "resident_global_mean_std",
                "frame_count": int(idx.size), "frame_bytes": self.pixels_per_frame * 4,
                "temporary_output_bytes": 0, "kernel_enqueue_s": call_s, "sync_s": 0.0,
                "total_s": time.perf_counter() - t0}

    @staticmethod
    def _coefficients(array, ndim: int, what: str):
        shape = tuple(array.shape) if _is_torch_cuda(array) else np.asarray(array).shape
        if len(shape) != ndim:
            raise ValueError(what)
        return shape

    def apply_grid_normalization_frame(self, index: int, scales, offsets, tile_height: int, tile_width: int,
                                       stream: int | None = None) -> dict:
        """x * scales[tile] + offsets[tile] in place, piecewise constant over
        the tiles (native_bindings.cpp:11054-11151)."""
        s_shape = self._coefficients(scales, 2, "scales and offsets must have shape (grid_rows, grid_cols)")
        o_shape = self._coefficients(offsets, 2, "scales and offsets must have shape (grid_rows, grid_cols)")
        if s_shape != o_shape:
            raise ValueError("scales and offsets shapes must match")
        rows, cols = self._grid_shape(tile_height, tile_width)
        if s_shape[0] <= 0 or s_shape[1] <= 0:
            raise ValueError("coefficient grid must not be empty")
        if s_shape != (rows, cols):
            raise ValueError("coefficient grid shape does not match resident frame shape and tile dimensions")
        t0 = time.perf_counter()
        idx = np.asarray([int(index)], dtype=np.int64)
        call_s = self._apply_normalization(idx, scales, offsets, tile_height, tile_width, stream)
        grid = rows * cols
        return {
            "schema_version": 1, "mode": "in_place_device_update", "model": "resident_grid_mean_std",
            "frame_index": int(index), "frame_bytes": self.pixels_per_frame * 4, "temporary_output_bytes": 0,
            "coefficient_count": grid, "coefficient_bytes": grid * 8, "coefficient_alloc_s": 0.0,
            "coefficient_upload_s": 0.0, "device_to_device_copy_s": 0.0, "kernel_enqueue_s": call_s,
            "sync_s": 0.0, "total_s": time.perf_counter() - t0,
        }

    def apply_grid_normalization_frames(self, frame_indices, scales, offsets, tile_height: int, tile_width: int,
                                        stream: int | None = None) -> dict:
        """apply_grid_normalization_frame for many frames in one launch per
        batch of distinct frames (native_bindings.cpp:11153-11290)."""
        arr = np.asarray(frame_indices)
        if arr.ndim != 1:
            raise ValueError("frame_indices must be a one-dimensional array")
        s_shape = self._coefficients(scales, 3, "scales and offsets must have shape (frame_count, grid_rows, grid_cols)")
        o_shape = self._coefficients(offsets, 3, "scales and offsets must have shape (frame_count, grid_rows, grid_cols)")
        if s_shape[0] != arr.size or o_shape[0] != arr.size:
            raise ValueError("coefficient frame count must match frame_indices length")
        if s_shape[1:] != o_shape[1:]:
            raise ValueError("scales and offsets grid shapes must match")
        if int(tile_height) <= 0 or int(tile_width) <= 0:
            raise ValueError("tile dimensions must be positive")
        count = int(arr.size)
        rows, cols = int(s_shape[1]), int(s_shape[2])
        base = {"schema_version": 1, "mode": "in_place_device_update_batch",
                "model": "resident_grid_mean_std_batch_apply", "frame_count": count, "grid_rows": rows,
                "grid_cols": cols}
        if count == 0:
            return {**base, "grid_count": 0, "coefficient_count": 0, "coefficient_bytes": 0,
                    "coefficient_alloc_s": 0.0, "coefficient_upload_s": 0.0, "kernel_enqueue_s": 0.0,
                    "sync_s": 0.0, "total_s": 0.0, "frames": []}
        if rows <= 0 or cols <= 0:
            raise ValueError("coefficient grid must not be empty")
        if (rows, cols) != self._grid_shape(tile_height, tile_width):
            raise ValueError("coefficient grid shape does not match resident frame shape and tile dimensions")
        idx = np.ascontiguousarray(arr, dtype=np.int64)
        t0 = time.perf_counter()
        call_s = self._apply_normalization(idx, scales, offsets, tile_height, tile_width, stream)
        total_s = time.perf_counter() - t0
        grid = rows * cols
        frames = [{
            "schema_version": 1, "mode": "in_place_device_update_batch", "model": "resident_grid_mean_std",
            "frame_index": int(idx[pos]), "batch_position": pos, "frame_bytes": self.pixels_per_frame * 4,
            "temporary_output_bytes": 0, "coefficient_count": grid, "coefficient_bytes": grid * 8,
            "coefficient_alloc_s": 0.0, "coefficient_upload_s": 0.0, "device_to_device_copy_s": 0.0,
            "kernel_enqueue_s": call_s / count, "sync_s": 0.0, "total_s": total_s / count,
        } for pos in range(count)]
        return {**base, "grid_count": grid, "coefficient_count": count * grid, "coefficient_bytes": count * grid * 8,
                "coefficient_alloc_s": 0.0, "coefficient_upload_s": 0.0, "kernel_enqueue_s": call_s,
                "sync_s": 0.0, "total_s": total_s, "frame_indices": arr, "frames": frames}

    # ------------------------------------------------------------ integrate
    def _integrate(self, weights, params: N.IntegrateParams, want, count_dtype, dq=False, dq_geometric=False,
                   dq_active_frames=0, device_outputs: dict | None = None, stream=None):
        if weights is None:
            w = None
            wptr = None
        else:
            w = np.ascontiguousarray(np.asarray(weights, dtype=np.float32))
            if w.shape != (self._frame_count,):
                raise ValueError("weights must have shape (frame_count,)")
            wptr = w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        shape = (self._height, self._width)
        host: dict[str, Any] = {}
        outs = N.IntegrateOutputs()
        cdt = np.uint16 if count_dtype == N.GH_COUNT_U16 else np.float32
        for name, dtype in (("master", np.float32), ("weight", np.float32), ("coverage", cdt), ("low", cdt),
                            ("high", cdt), ("dq", np.uint32)):
            if name == "dq" and not dq:
                continue
            if device_outputs and name in device_outputs:
                setattr(outs, name, device_outputs[name].data_ptr())
                continue
            if name in want or name == "dq":
                arr = np.empty(shape, dtype=dtype)
                host[name] = arr
                setattr(outs, name, arr.ctypes.data)
        outs.dq_use_geometric = int(bool(dq_geometric))
        outs.dq_active_frames = int(dq_active_frames)
        N.call("gh_integrate", self._h, wptr, ctypes.byref(params), ctypes.byref(outs),
               ctypes.c_void_p(stream) if stream else None)
        return host

    def integrate_mean(self, weights=None):
        """integrate_mean(weights=None) -> (master, weight_map)
        (native_bindings.cpp:12935)."""
        params = N.IntegrateParams(N.GH_REJECT_NONE, 3.0, 3.0, 3, 0.5, N.GH_COUNT_F32, -1, 0, 0)
        out = self._integrate(weights, params, {"master", "weight"}, N.GH_COUNT_F32)
        return out["master"], out["weight"]

    # ------------------------------------------------ sigma clip / fused warp
    def _clip(self, weights, matrices, interpolation, clamping_threshold, rejection, low_sigma, high_sigma,
              winsorize, download_mode, geometric: bool):
        if download_mode not in ("full", "master_weight", "master_only"):
            raise ValueError("download_mode must be full, master_weight, or master_only")
        f = self.frame_count
        w = None
        if weights is not None:
            w = np.ascontiguousarray(np.asarray(weights, dtype=np.float32))
            if w.ndim != 1 or w.size != f:
                raise ValueError("weights must have shape (frame_count,)")
        inv = None
        sampler = N.GH_CLIP_STACK
        if matrices is not None:
            if interpolation not in ("bilinear", "lanczos3"):
                raise ValueError("interpolation must be bilinear or lanczos3")
            mats = np.asarray(matrices, dtype=np.float64)
            if mats.shape != (f, 3, 3):
                raise ValueError("matrices must have shape (frame_count, 3, 3)")
            inv = np.ascontiguousarray(np.stack([invert_affine(m) for m in mats]), dtype=np.float64)
            sampler = N.GH_INTERP_LANCZOS3 if interpolation == "lanczos3" else N.GH_INTERP_BILINEAR
        prm = N.ClipParams(sampler, rejection, int(bool(winsorize)), float(low_sigma), float(high_sigma),
                           float(clamping_threshold))
        shape = (self.height, self.width)
        want_weight = download_mode != "master_only"
        full = download_mode == "full"
        master = np.empty(shape, np.float32)
        maps = {"weight": np.empty(shape, np.float32) if want_weight else None}
        for name in ("coverage", "low", "high"):
            maps[name] = np.empty(shape, np.float32) if full else None
        maps["geometric"] = np.empty(shape, np.float32) if (full and geometric) else None
        if rejection == N.GH_REJECT_NONE:
            maps["low"] = maps["high"] = None

        def ptr(a):
            return None if a is None else ctypes.c_void_p(a.ctypes.data)

        t0 = time.perf_counter()
        N.call("gh_integrate_clip", self._h,
               None if w is None else w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
               None if inv is None else inv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(prm),
               ptr(master), ptr(maps["weight"]), ptr(maps["coverage"]), ptr(maps["low"]), ptr(maps["high"]),
               ptr(maps["geometric"]), None)
        return master, maps, time.perf_counter() - t0

    def integrate_sigma_clip(self, weights=None, low_sigma=3.0, high_sigma=3.0, winsorize=True,
                             download_mode="full"):
        """integrate_sigma_clip (native_bindings.cpp:13227-13393) -> (master,
        weight|None, coverage|None, low|None, high|None): mean / std sigma clip
        with an optional winsorized re-estimate, float32 sums in frame order."""
        if not (float(low_sigma) > 0.0 and float(high_sigma) > 0.0):
            raise ValueError("sigma thresholds must be positive")
        master, m, _ = self._clip(weights, None, "bilinear", -1.0, N.GH_REJECT_SIGMA_CLIP, low_sigma, high_sigma,
                                  winsorize, download_mode, geometric=False)
        return master, m["weight"], m["coverage"], m["low"], m["high"]

    # ---- star candidate catalogues (native_bindings.cpp:11356-12122)
    def _candidates(self, index, mode, threshold, max_candidates=0, max_output_candidates=0, grid=(0, 0, 0),
                    min_separation_px=0.0, centroid_radius=0, centroid_global_mean_background=False):
        prm = N.CandidateParams(mode, float(threshold), int(max_candidates), int(max_output_candidates), int(grid[0]),
                                int(grid[1]), int(grid[2]), float(min_separation_px), int(centroid_radius or 0),
                                int(bool(centroid_global_mean_background)))
        cap = max(1, int(max_candidates) if mode in (N.GH_CANDIDATES_ALL, N.GH_CANDIDATES_TOP)
                  else int(max_output_candidates))
        xs, ys, fl = (np.zeros(cap, np.float32) for _ in range(3))
        status = np.zeros(cap, np.uint8)
        res = N.CandidateResult()
        N.call("gh_star_candidates", self._h, int(index), ctypes.byref(prm), ctypes.c_void_p(xs.ctypes.data),
               ctypes.c_void_p(ys.ctypes.data), ctypes.c_void_p(fl.ctypes.data), ctypes.c_void_p(status.ctypes.data),
               ctypes.byref(res))
        k = int(res.stored_count)
        out = {"count": int(res.count), "stored_count": k}
        if centroid_radius is not None and mode in (N.GH_CANDIDATES_TOP_NMS, N.GH_CANDIDATES_GRID_NMS):
            on = int(centroid_radius) > 0
            bg = float(res.centroid_background)
            out["centroid_refine"] = {
                "enabled": on, "radius": int(centroid_radius), "stored_count": k,
                "refined_count": int(res.refined_count) if on else 0,
                "failed_count": k - int(res.refined_count) if on else 0,
                "max_shift_px": float(res.max_shift_px),
                "mode": ("resident_gpu_global_mean_centroid" if centroid_global_mean_background
                         else "resident_gpu_window_centroid") if on else "off",
                "background_mode": ("global_mean" if centroid_global_mean_background else "local_median")
                if on else "off",
                "background": bg if np.isfinite(bg) else None,
                "status": status[:k].copy() if on else None}
        out.update(x=xs[:k].copy(), y=ys[:k].copy(), flux=fl[:k].copy())
        return out, int(res.selected_count)

    def star_candidates(self, index, threshold, max_candidates):
        """star_candidates (native_bindings.cpp:11356-11418): the 3x3 local
        maxima above `threshold`; the first max_candidates in raster order."""
        if int(max_candidates) <= 0:
            raise ValueError("max_candidates must be positive")
        return self._candidates(index, N.GH_CANDIDATES_ALL, threshold, max_candidates)[0]

    def star_top_candidates(self, index, threshold, max_candidates):
        """star_top_candidates (native_bindings.cpp:11420-11487): the best
        max_candidates local maxima by flux (descending), then y, then x."""
        if int(max_candidates) <= 0:
            raise ValueError("max_candidates must be positive")
        return self._candidates(index, N.GH_CANDIDATES_TOP, threshold, max_candidates)[0]

    def star_top_nms_candidates(self, index, threshold, scan_candidates, max_output_candidates, min_separation_px):
        """star_top_nms_candidates (native_bindings.cpp:11489-11615): the best
        scan_candidates, then greedy separation >= min_separation_px."""
        out, _ = self._candidates(index, N.GH_CANDIDATES_TOP_NMS, threshold, scan_candidates, max_output_candidates,
                                  min_separation_px=min_separation_px, centroid_radius=None)
        out.update(scan_candidates=int(scan_candidates), max_output_candidates=int(max_output_candidates),
                   min_separation_px=float(min_separation_px))
        return {k: out[k] for k in ("count", "stored_count", "scan_candidates", "max_output_candidates",
                                    "min_separation_px", "x", "y", "flux")}

    def star_top_nms_candidates_centroid(self, index, threshold, scan_candidates, max_output_candidates,
                                         min_separation_px, centroid_radius, centroid_global_mean_background=False):
        """star_top_nms_candidates_centroid (native_bindings.cpp:11617-11806):
        star_top_nms_candidates with window centroids of the kept positions."""
        out, _ = self._candidates(index, N.GH_CANDIDATES_TOP_NMS, threshold, scan_candidates, max_output_candidates,
                                  min_separation_px=min_separation_px, centroid_radius=centroid_radius,
                                  centroid_global_mean_background=centroid_global_mean_background)
        out.update(scan_candidates=int(scan_candidates), max_output_candidates=int(max_output_candidates),
                   min_separation_px=float(min_separation_px))
        return out

    def _grid_candidates(self, index, threshold, grid_cols, grid_rows, candidates_per_cell, max_output_candidates,
                         min_separation_px, deterministic, centroid_radius, centroid_global_mean_background):
        out, selected = self._candidates(index, N.GH_CANDIDATES_GRID_NMS, threshold, 0, max_output_candidates,
                                         (grid_cols, grid_rows, candidates_per_cell), min_separation_px,
                                         centroid_radius, centroid_global_mean_background)
        cells = int(grid_cols) * int(grid_rows)
        out.update(grid_cols=int(grid_cols), grid_rows=int(grid_rows), cell_count=cells,
                   candidates_per_cell=int(candidates_per_cell), grid_capacity=cells * int(candidates_per_cell),
                   max_output_candidates=int(max_output_candidates), min_separation_px=float(min_separation_px),
                   deterministic=bool(deterministic), cell_selected_count=selected,
                   catalog_sort_mode="device_radix_rank_keys", catalog_topk_mode="stable_cell_rank_cut")
        return out

    def star_grid_top_nms_candidates(self, index, threshold, grid_cols, grid_rows, candidates_per_cell,
                                     max_output_candidates, min_separation_px, deterministic=False):
        """star_grid_top_nms_candidates[_deterministic]
        (native_bindings.cpp:11808-12109): the best candidates_per_cell of each
        grid cell, ranked together, then greedy separation.  The ranking here
        is deterministic in both forms."""
        return self._grid_candidates(index, threshold, grid_cols, grid_rows, candidates_per_cell,
                                     max_output_candidates, min_separation_px, deterministic, 0, False)

    def star_grid_top_nms_candidates_deterministic(self, index, threshold, grid_cols, grid_rows, candidates_per_cell,
                                                   max_output_candidates, min_separation_px):
        return self.star_grid_top_nms_candidates(index, threshold, grid_cols, grid_rows, candidates_per_cell,
                                                 max_output_candidates, min_separation_px, deterministic=True)

    def star_grid_top_nms_candidates_centroid(self, index, threshold, grid_cols, grid_rows, candidates_per_cell,
                                              max_output_candidates, min_separation_px, centroid_radius,
                                              centroid_global_mean_background=False, deterministic=False):
        """star_grid_top_nms_candidates[_deterministic]_centroid
        (native_bindings.cpp:12065-12122): the grid catalogue with window
        centroids of the kept positions."""
        return self._grid_candidates(index, threshold, grid_cols, grid_rows, candidates_per_cell,
                                     max_output_candidates, min_separation_px, deterministic, centroid_radius,
                                     centroid_global_mean_background)

    def star_grid_top_nms_candidates_deterministic_centroid(self, index, threshold, grid_cols, grid_rows,
                                                            candidates_per_cell, max_output_candidates,
                                                            min_separation_px, centroid_radius,
                                                            centroid_global_mean_background=False):
        return self.star_grid_top_nms_candidates_centroid(index, threshold, grid_cols, grid_rows, candidates_per_cell,
                                                          max_output_candidates, min_separation_px, centroid_radius,
                                                          centroid_global_mean_background, deterministic=True)

    def star_grid_top_nms_candidates_batch(self, indices, threshold, grid_cols, grid_rows, candidates_per_cell,
                                           max_output_candidates, min_separation_px, deterministic=False):
        """star_grid_top_nms_candidates_batch (native_bindings.cpp:12123-12644):
        the per-frame catalogues of `indices`, each item tagged with its frame."""
        out = []
        for i in indices:
            item = self.star_grid_top_nms_candidates(i, threshold, grid_cols, grid_rows, candidates_per_cell,
                                                     max_output_candidates, min_separation_px, deterministic)
            item["frame_index"] = int(i)
            out.append(item)
        return out

    def refine_matrix_translation_candidates_batch_to_reference(self, reference_index, moving_indices, matrices,
                                                                search_radius_px=1.0, coarse_step_px=0.25,
                                                                fine_radius_px=0.25, fine_step_px=0.0625,
                                                                coarse_sample_stride=4, final_sample_stride=1):
        """refine_matrix_translation_candidates_batch_to_reference
        (native_bindings.cpp:10186-10416): one seed matrix per moving frame,
        each searched as refine_matrix_translation_candidates_to_reference;
        the coarse candidates of all frames share one launch, as do the fine
        ones."""
        moving = [int(i) for i in list(moving_indices)]
        # parse_matrix_stack (native_bindings.cpp:1365-1388): one (3, 3) matrix
        # is a stack of one; an empty stack is an error even without frames
        mats = np.asarray(matrices, dtype=np.float32)
        if mats.shape == (3, 3):
            mats = mats.reshape(1, 3, 3)
        if mats.ndim != 3 or mats.shape[1:] != (3, 3):
            raise ValueError("matrices must have shape (N, 3, 3) or (3, 3)")
        if mats.shape[0] <= 0:
            raise ValueError("matrices must contain at least one matrix")
        if moving and mats.shape[0] != len(moving):
            raise ValueError("batch refine requires one seed matrix per moving frame")
        prm = N.AlignSearchParams(float(search_radius_px), float(coarse_step_px), float(fine_radius_px),
                                  float(fine_step_px), int(coarse_sample_stride), int(final_sample_stride))
        n = len(moving)
        out = (N.AlignCandidate * max(n, 1))()
        idx = np.ascontiguousarray(np.asarray(moving, dtype=np.int64).reshape(n))
        mats = np.ascontiguousarray(mats)
        N.call("gh_refine_matrix_translation_candidates_batch", self._h, int(reference_index),
               idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), n,
               mats.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(prm), out)
        results = []
        for k, i in enumerate(moving):
            c = out[k]
            results.append({
                "reference_index": int(reference_index), "moving_index": i,
                "matrix": np.asarray(list(c.matrix), np.float32).reshape(3, 3).astype(float).tolist(),
                "dx_correction": float(c.dx), "dy_correction": float(c.dy),
                "metrics": {"valid_pixels": int(c.valid_pixels), "sampled_pixels": int(c.sampled_pixels),
                            "sample_stride": int(c.sample_stride), "rms": float(c.rms),
                            "mean_abs_diff": float(c.mean_abs_diff), "ncc": float(c.ncc),
                            "model": "matrix_alignment_metrics_cuda_candidate_grid"},
                "coarse_candidates": int(c.coarse_candidates), "fine_candidates": int(c.fine_candidates),
                "model": "resident_cuda_matrix_metric_translation_batch_refine_grid"})
        return results

    def estimate_translation_from_stars_to_reference(self, reference_index, moving_index, threshold,
                                                     max_candidates=64, tolerance_px=1.0, max_abs_dx=-1.0,
                                                     max_abs_dy=None, prior_dx=0.0, prior_dy=0.0,
                                                     prior_radius_px=-1.0, grid_cols=0, grid_rows=0):
        """estimate_translation_from_stars_to_reference
        (native_bindings.cpp:12645-12880): reference - moving offset voted by
        the pair offsets of the two frames' star catalogues, refined by the
        mutual nearest matches under it."""
        prm = N.StarTranslationParams(float(threshold), int(max_candidates), float(tolerance_px), float(max_abs_dx),
                                      -1.0 if max_abs_dy is None else float(max_abs_dy), float(prior_dx),
                                      float(prior_dy), float(prior_radius_px), int(grid_cols), int(grid_rows))
        res = N.StarTranslationResult()
        N.call("gh_estimate_translation_from_stars", self._h, int(reference_index), int(moving_index),
               ctypes.byref(prm), ctypes.byref(res))
        grid = int(grid_cols) > 0 or int(grid_rows) > 0
        rn, mn = int(res.reference_count), int(res.moving_count)
        return {"dx": float(res.dx), "dy": float(res.dy), "inliers": int(res.inliers),
                "refined_dx": float(res.refined_dx), "refined_dy": float(res.refined_dy),
                "mutual_inliers": int(res.mutual_inliers), "rms_px": float(res.rms_px),
                "candidate_count": rn * mn, "reference_count": rn, "moving_count": mn,
                "reference_total_count": int(res.reference_total_count),
                "moving_total_count": int(res.moving_total_count), "threshold": float(threshold),
                "max_candidates": int(max_candidates),
                "catalog_capacity": int(grid_cols) * int(grid_rows) if grid else int(max_candidates),
                "candidate_selection": "grid_brightest_per_cell" if grid else "top_flux",
                "grid_cols": int(grid_cols), "grid_rows": int(grid_rows), "tolerance_px": float(tolerance_px),
                "max_abs_dx": float(res.max_abs_dx), "max_abs_dy": float(res.max_abs_dy),
                "prior_dx": float(prior_dx), "prior_dy": float(prior_dy),
                "prior_radius_px": float(res.prior_radius_px), "reference_index": int(reference_index),
                "moving_index": int(moving_index), "model": "resident_star_catalog_pair_offset_translation"}

    def refine_matrix_translation_candidates_to_reference(self, reference_index, moving_index, matrices,
                                                          search_radius_px=1.0, coarse_step_px=0.25,
                                                          fine_radius_px=0.25, fine_step_px=0.0625,
                                                          coarse_sample_stride=4, final_sample_stride=1):
        """refine_matrix_translation_candidates_to_reference
        (native_bindings.cpp:10079-10184): coarse then fine grids of
        translation corrections for each seed matrix, scored by the bilinear
        alignment metrics (rms, mean |diff|, ncc) against the reference."""
        m = np.asarray(matrices, dtype=np.float32)
        if m.shape == (3, 3):
            m = m[None]
        if m.ndim != 3 or m.shape[1:] != (3, 3):
            raise ValueError("matrices must have shape (N, 3, 3) or (3, 3)")
        if m.shape[0] <= 0:
            raise ValueError("matrices must contain at least one matrix")
        m = np.ascontiguousarray(m)
        prm = N.AlignSearchParams(float(search_radius_px), float(coarse_step_px), float(fine_radius_px),
                                  float(fine_step_px), int(coarse_sample_stride), int(final_sample_stride))
        seeds = (N.AlignCandidate * m.shape[0])()
        selected = ctypes.c_int64(-1)
        N.call("gh_refine_matrix_translation_candidates", self._h, int(reference_index), int(moving_index),
               m.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(m.shape[0]), ctypes.byref(prm), seeds,
               ctypes.byref(selected))
        model = "matrix_alignment_metrics_cuda_candidate_grid"

        def metrics(c):
            return {"valid_pixels": int(c.valid_pixels), "sampled_pixels": int(c.sampled_pixels),
                    "sample_stride": int(c.sample_stride), "rms": float(c.rms),
                    "mean_abs_diff": float(c.mean_abs_diff), "ncc": float(c.ncc), "model": model}

        def matrix(c):
            return np.asarray(list(c.matrix), np.float32).reshape(3, 3).astype(float).tolist()

        seed_results = [{"seed_index": k, "matrix": matrix(c), "dx_correction": float(c.dx),
                         "dy_correction": float(c.dy), "metrics": metrics(c),
                         "coarse_candidates": int(c.coarse_candidates), "fine_candidates": int(c.fine_candidates)}
                        for k, c in enumerate(seeds)]
        best = seeds[selected.value]
        return {"matrix": matrix(best), "dx_correction": float(best.dx), "dy_correction": float(best.dy),
                "metrics": metrics(best), "selected_index": int(selected.value), "seed_count": int(m.shape[0]),
                "seed_results": seed_results, "coarse_candidates_per_seed": int(seeds[0].coarse_candidates),
                "search_radius_px": float(search_radius_px), "coarse_step_px": float(coarse_step_px),
                "fine_radius_px": float(fine_radius_px), "fine_step_px": float(fine_step_px),
                "coarse_sample_stride": int(coarse_sample_stride), "final_sample_stride": int(final_sample_stride),
                "reference_index": int(reference_index), "moving_index": int(moving_index),
                "model": "resident_cuda_matrix_metric_translation_multi_seed_refine_grid"}

    def _tile_local(self, target_mask, tile_extents, tile_multipliers, weights, rejection, low_sigma, high_sigma,
                    winsorize):
        f = self.frame_count
        if self.loaded_count != f:
            raise RuntimeError("all resident frames must be loaded before integration")
        target = np.ascontiguousarray(np.asarray(target_mask, dtype=np.uint8).reshape((f,)))
        extents = np.asarray(tile_extents, dtype=np.int32)
        if extents.ndim != 2 or extents.shape[1] != 4:
            raise ValueError("tile_extents must have shape (tile_count, 4)")
        extents = np.ascontiguousarray(extents)
        multipliers = np.ascontiguousarray(np.asarray(tile_multipliers, dtype=np.float32))
        if multipliers.ndim != 1 or multipliers.shape[0] != extents.shape[0]:
            raise ValueError("tile_multipliers must have shape (tile_count,)")
        w = None
        if weights is not None:
            w = np.ascontiguousarray(np.asarray(weights, dtype=np.float32).reshape((f,)))
        prm = N.ClipParams(N.GH_CLIP_STACK, rejection, int(bool(winsorize)), float(low_sigma), float(high_sigma), -1.0)
        shape = (self.height, self.width)
        maps = [np.empty(shape, np.float32) for _ in range(2 if rejection == N.GH_REJECT_NONE else 5)]
        ptrs = [ctypes.c_void_p(a.ctypes.data) for a in maps] + [None] * (5 - len(maps))
        t0 = time.perf_counter()
        N.call("gh_integrate_tile_local", self._h,
               None if w is None else w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
               target.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
               extents.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
               multipliers.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(extents.shape[0]), ctypes.byref(prm),
               *ptrs, None)
        tiles = int(extents.shape[0])
        timing = {"schema_version": 1, "frame_count": f, "target_frame_count": int(np.count_nonzero(target)),
                  "tile_count": tiles, "total_s": time.perf_counter() - t0, "weights_bytes": f * 4,
                  "target_mask_bytes": f, "tile_extents_bytes": tiles * 16, "tile_multipliers_bytes": tiles * 4,
                  "output_bytes": self.pixels_per_frame * 4 * len(maps), "modifies_resident_stack": False}
        return maps, timing

    def integrate_tile_local_mean(self, target_mask, tile_extents, tile_multipliers, weights=None):
        """integrate_tile_local_mean (native_bindings.cpp:12998-13225) ->
        (master, weight_map, timing): the weighted mean of the stack where a
        target frame's weight is multiplied by the multiplier of the tile that
        holds the pixel (1 outside every tile)."""
        maps, timing = self._tile_local(target_mask, tile_extents, tile_multipliers, weights, N.GH_REJECT_NONE,
                                        3.0, 3.0, False)
        timing.update(timing_model="native_resident_tile_local_weighted_mean_one_sync", rejection="none")
        return maps[0], maps[1], timing

    def integrate_tile_local_sigma_clip(self, target_mask, tile_extents, tile_multipliers, weights=None,
                                        low_sigma=3.0, high_sigma=3.0, winsorize=True):
        """integrate_tile_local_sigma_clip (native_bindings.cpp:14337-14630)
        -> (master, weight, coverage, low, high, timing): integrate_sigma_clip
        with the tile-local weights of integrate_tile_local_mean; a frame whose
        weight at a pixel is not positive does not enter its statistics."""
        if self.loaded_count != self.frame_count:
            raise RuntimeError("all resident frames must be loaded before integration")
        if not (float(low_sigma) > 0.0 and float(high_sigma) > 0.0):
            raise ValueError("sigma thresholds must be positive")
        maps, timing = self._tile_local(target_mask, tile_extents, tile_multipliers, weights,
                                        N.GH_REJECT_SIGMA_CLIP, low_sigma, high_sigma, winsorize)
        timing.update(timing_model="native_resident_tile_local_sigma_clip_one_sync",
                      rejection="winsorized_sigma" if winsorize else "sigma_clip", low_sigma=float(low_sigma),
                      high_sigma=float(high_sigma), winsorize=bool(winsorize))
        return (*maps, timing)

    def integrate_matrix_warped_mean(self, matrices, weights=None, interpolation="bilinear", clamping_threshold=-1.0,
                                     download_mode="full"):
        """integrate_matrix_warped_mean (native_bindings.cpp:14632-14838) ->
        (master, weight|None, coverage|None, geometric|None, timing): the
        weighted mean of every frame resampled through its matrix on the fly;
        the resident stack is not modified."""
        master, m, call_s = self._clip(weights, matrices, interpolation, clamping_threshold, N.GH_REJECT_NONE, 3.0,
                                       3.0, False, download_mode, geometric=True)
        timing = {"schema_version": 1, "timing_model": "native_fused_matrix_warp_weighted_mean_one_sync",
                  "interpolation": interpolation, "clamping_threshold": float(clamping_threshold),
                  "rejection": "none", "frame_count": self.frame_count, "total_s": call_s,
                  "download_mode": download_mode, "diagnostic_maps_downloaded": download_mode == "full",
                  "weight_map_downloaded": download_mode != "master_only", "avoids_stack_scatter": True,
                  "modifies_resident_stack": False, "inverse_batch_bytes": self.frame_count * 9 * 8,
                  "weights_bytes": self.frame_count * 4,
                  "output_bytes": self.pixels_per_frame * 4 * (1 + (download_mode != "master_only")
                                                               + 2 * (download_mode == "full"))}
        return master, m["weight"], m["coverage"], m["geometric"], timing

    def integrate_matrix_warped_sigma_clip(self, matrices, weights=None, interpolation="bilinear",
                                           clamping_threshold=-1.0, low_sigma=3.0, high_sigma=3.0, winsorize=True,
                                           download_mode="full"):
        """integrate_matrix_warped_sigma_clip (native_bindings.cpp:14840-15115)
        -> (master, weight|None, coverage|None, low|None, high|None,
        geometric|None, timing)."""
        if not (float(low_sigma) > 0.0 and float(high_sigma) > 0.0):
            raise ValueError("sigma thresholds must be positive")
        master, m, call_s = self._clip(weights, matrices, interpolation, clamping_threshold, N.GH_REJECT_SIGMA_CLIP,
                                       low_sigma, high_sigma, winsorize, download_mode, geometric=True)
        timing = {"schema_version": 1, "timing_model": "native_fused_matrix_warp_sigma_clip_one_sync",
                  "interpolation": interpolation, "clamping_threshold": float(clamping_threshold),
                  "rejection": "winsorized_sigma" if winsorize else "sigma_clip", "winsorize": bool(winsorize),
                  "low_sigma": float(low_sigma), "high_sigma": float(high_sigma), "frame_count": self.frame_count,
                  "total_s": call_s, "download_mode": download_mode,
                  "diagnostic_maps_downloaded": download_mode == "full",
                  "weight_map_downloaded": download_mode != "master_only", "avoids_stack_scatter": True,
                  "modifies_resident_stack": False, "inverse_batch_bytes": self.frame_count * 9 * 8,
                  "weights_bytes": self.frame_count * 4,
                  "output_bytes": self.pixels_per_frame * 4 * (1 + (download_mode != "master_only")
                                                               + 4 * (download_mode == "full"))}
        return master, m["weight"], m["coverage"], m["low"], m["high"], m["geometric"], timing

    def integrate_hardened_winsorized_sigma(self, weights=None, low_sigma=3.0, high_sigma=3.0, min_samples=3,
                                            max_reject_fraction=0.5, count_map_dtype="float32", profile=False,
                                            download_mode="full"):
        """integrate_hardened_winsorized_sigma (native_bindings.cpp:13395,
        binding :21174-21184) -> (master, weight|None, coverage|None, low|None,
        high|None[, profile])."""
        result = self._winsorized(weights, low_sigma, high_sigma, min_samples, max_reject_fraction, count_map_dtype,
                                  download_mode)
        if profile:
            return (*result[0], result[1])
        return result[0]

    def integrate_hardened_winsorized_sigma_timed(self, weights=None, low_sigma=3.0, high_sigma=3.0, min_samples=3,
                                                  max_reject_fraction=0.5, count_map_dtype="float32",
                                                  download_mode="full"):
        maps, prof = self._winsorized(weights, low_sigma, high_sigma, min_samples, max_reject_fraction,
                                      count_map_dtype, download_mode)
        return (*maps, prof)

    def _winsorized(self, weights, low_sigma, high_sigma, min_samples, max_reject_fraction, count_map_dtype,
                    download_mode, statistics_variant=-1, tile_size=0, require_positive_scale=0):
        if count_map_dtype not in ("float32", "uint16"):
            raise ValueError("count_map_dtype must be 'float32' or 'uint16'")
        modes = {
            "full": {"master", "weight", "coverage", "low", "high"},
            "master_weight": {"master", "weight"},
            "master_only": {"master"},
        }
        if download_mode not in modes:
            raise ValueError("download_mode must be 'full', 'master_weight' or 'master_only'")
        if not (float(low_sigma) > 0 and float(high_sigma) > 0):
            raise ValueError("low_sigma and high_sigma must be positive")
        if int(min_samples) < 1:
            raise ValueError("min_samples must be at least 1")
        if not 0.0 <= float(max_reject_fraction) <= 1.0:
            raise ValueError("max_reject_fraction must be between 0 and 1")
        cdt = N.GH_COUNT_U16 if count_map_dtype == "uint16" else N.GH_COUNT_F32
        params = N.IntegrateParams(N.GH_REJECT_WINSORIZED_SIGMA, float(low_sigma), float(high_sigma),
                                   int(min_samples), float(max_reject_fraction), cdt, int(statistics_variant),
                                   int(require_positive_scale), int(tile_size))
        want = modes[download_mode]
        start = time.perf_counter()
        out = self._integrate(weights, params, want, cdt)
        total = time.perf_counter() - start
        maps = tuple(out.get(name) for name in ("master", "weight", "coverage", "low", "high"))
        active = self._frame_count if weights is None else int(
            np.count_nonzero(np.isfinite(np.asarray(weights, dtype=np.float32)) & (np.asarray(weights, dtype=np.float32) > 0))
        )
        prof = {
            "native_method": "ResidentCalibratedStack.integrate_hardened_winsorized_sigma",
            "resident_winsorized_mode": "hardened_cpu_parity",
            "frame_count": self._frame_count,
            "active_frame_count": active,
            "pixel_count": self.pixels_per_frame,
            "count_map_dtype_requested": count_map_dtype,
            "count_map_dtype": count_map_dtype,
            "download_mode": download_mode,
            "native_profile": {
                "schema_version": 1,
                "percentile_strategy": "register_sorting_network_order_statistics",
                "winsorized_accumulation_order": "fp64",
                "kernel": "integrate_winsorized_kernel",
            },
            "total_s": total,
        }
        return maps, prof

    def integrate_stack_engine(self, low_sigma=3.0, high_sigma=3.0, min_samples=3, max_reject_fraction=0.5,
                               tile_size=512):
        """CPUStackEngine semantics (master calibration): per-tile statistics
        arithmetic and the positive-scale rejection rule
        (src/glass/engine/stack_engine.py:39-120, :578-593)."""
        maps, _ = self._winsorized(None, low_sigma, high_sigma, min_samples, max_reject_fraction, "float32", "full",
                                   statistics_variant=-1, tile_size=int(tile_size), require_positive_scale=1)
        return maps

    def integrate_stack_device(self, master_out, low_sigma=3.0, high_sigma=3.0, min_samples=3,
                               max_reject_fraction=0.5, tile_size=512, stream=None) -> None:
        """CPUStackEngine semantics (see integrate_stack_engine) into a GPU tensor."""
        params = N.IntegrateParams(N.GH_REJECT_WINSORIZED_SIGMA, float(low_sigma), float(high_sigma),
                                   int(min_samples), float(max_reject_fraction), N.GH_COUNT_F32, -1, 1,
                                   int(tile_size))
        self._integrate(None, params, set(), N.GH_COUNT_F32, device_outputs={"master": master_out}, stream=stream)

    def warp_frames(self, indices, matrices, interpolation="lanczos3", semantics="cpu", fill=math.nan,
                    clamping_threshold=-1.0, track_coverage=True, stream=None) -> dict:
        """Batched matrix warp on a caller stream; semantics "cpu" (engine/warp.py)
        or "resident" (non-finite taps skipped, optional clamp)."""
        interp = N.GH_INTERP_LANCZOS3 if interpolation == "lanczos3" else N.GH_INTERP_BILINEAR
        return self._warp(indices, matrices, interp, fill, clamping_threshold, track_coverage,
                          semantics == "resident", stream)

    def integrate_to_device(self, outputs: dict, weights=None, rejection="winsorized_sigma", low_sigma=3.0,
                            high_sigma=3.0, min_samples=3, max_reject_fraction=0.5, count_map_dtype="uint16",
                            dq_geometric=True, dq_active_frames=0, stream=None) -> None:
        """Integrate into caller-owned GPU tensors (keys master, weight,
        coverage, low, high, dq); nothing is copied to the host."""
        cdt = N.GH_COUNT_U16 if count_map_dtype == "uint16" else N.GH_COUNT_F32
        rej = N.GH_REJECT_WINSORIZED_SIGMA if rejection == "winsorized_sigma" else N.GH_REJECT_NONE
        params = N.IntegrateParams(rej, float(low_sigma), float(high_sigma), int(min_samples),
                                   float(max_reject_fraction), cdt, -1, 0, 0)
        self._integrate(weights, params, set(), cdt, dq="dq" in outputs, dq_geometric=dq_geometric,
                        dq_active_frames=dq_active_frames, device_outputs=outputs, stream=stream)

    def synchronize(self) -> None:
        N.call("gh_synchronize", self._h)
