"""Host-side staging of raw light frames for ResidentCalibratedStack: the
asynchronous / multistream calibration entry points and the FITS path
completion queue of the reference (cpp/src/native_bindings.cpp:5979-8778,
the completion queue :2148-2915).

Every frame still goes through gh_calibrate_frames (the library's upload ring
overlaps its copy with calibration); what this module adds is the host side:
a ring of pinned buffers, reader threads that fill them from FITS files, and
completion order.  Reads finish out of order; each finished read is calibrated
at once, and its pinned buffer returns to the pool when the calibration that
reads it has run.
"""

from __future__ import annotations

import ctypes
import os
import threading
import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from typing import Any, Sequence

import numpy as np


def _read_raw(path, offset: int, nbytes: int, out) -> int:
    """Read `nbytes` undecoded data bytes at `offset` of `path` into `out`."""
    view = memoryview(out).cast("B")
    if view.nbytes < nbytes:
        raise ValueError("raw buffer is smaller than the frame")
    fd = os.open(path, os.O_RDONLY)
    try:
        done = 0
        while done < nbytes:
            got = os.preadv(fd, [view[done:nbytes]], int(offset) + done)
            if got <= 0:
                raise ValueError(f"{path}: FITS image data is truncated")
            done += got
    finally:
        os.close(fd)
    return nbytes


class FitsCompletionQueue:
    """Reader threads fill `buffers` pinned host buffers with the raw data of
    `paths`; `calibrate` takes the reads in completion order."""

    def __init__(self, stack, indices, paths, data_offsets, byte_counts, queue_buffer_count: int, worker_count: int):
        import torch

        self.stack = stack
        self.indices = [int(i) for i in indices]
        self.paths = [str(p) for p in paths]
        n = len(self.indices)
        if not (len(self.paths) == len(data_offsets) == len(byte_counts) == n):
            raise ValueError("indices, paths, data_offsets and byte_counts must have the same length")
        self.frame_bytes = stack.pixels_per_frame * 2
        for b in byte_counts:
            if int(b) != self.frame_bytes:
                raise ValueError("byte_counts must equal height * width * 2 for BITPIX=16 lights")
        self.offsets = [int(o) for o in data_offsets]
        nbuf = max(1, min(int(queue_buffer_count), n or 1))
        self.buffers = [torch.empty(self.frame_bytes, dtype=torch.uint8, pin_memory=True) for _ in range(nbuf)]
        self.free = list(range(nbuf))
        self.cond = threading.Condition()
        self.cancelled = False
        self.pool = ThreadPoolExecutor(max_workers=max(1, int(worker_count)))
        self.futures = {}
        self.next = 0
        self.read_s = 0.0
        self.t0 = time.perf_counter()
        self._submit()

    def _read(self, k: int, slot: int):
        t = time.perf_counter()
        _read_raw(self.paths[k], self.offsets[k], self.frame_bytes, self.buffers[slot].numpy())
        return k, slot, time.perf_counter() - t

    def _submit(self):
        with self.cond:
            while not self.cancelled and self.free and self.next < len(self.indices):
                slot = self.free.pop(0)
                self.futures[self.pool.submit(self._read, self.next, slot)] = self.next
                self.next += 1

    def cancel(self) -> None:
        with self.cond:
            self.cancelled = True
        for fut in list(self.futures):
            fut.cancel()
        self.pool.shutdown(wait=True)

    def calibrate(self, light_exposures, dark_exposures, policy) -> dict[str, Any]:
        import torch

        from . import _native as N

        lights = list(light_exposures)
        darks = (list(dark_exposures) if isinstance(dark_exposures, (list, tuple, np.ndarray))
                 else [dark_exposures] * len(self.indices))
        if len(lights) != len(self.indices) or len(darks) != len(self.indices):
            raise ValueError("light_exposures and dark_exposures must match the queued frames")
        stream = torch.cuda.Stream()
        pending: list[tuple[Any, int]] = []  # (event, slot) of calibrations in flight
        order = []
        try:
            while self.futures:
                done, _ = wait(list(self.futures), return_when=FIRST_COMPLETED)
                for fut in done:
                    self.futures.pop(fut)
                    k, slot, dt = fut.result()
                    self.read_s += dt
                    self.stack._calibrate([self.indices[k]], ctypes.c_void_p(self.buffers[slot].data_ptr()),
                                          self.frame_bytes, N.GH_FMT_FITS_U16BE_BZERO, [lights[k]], [darks[k]],
                                          policy, False, stream.cuda_stream)
                    ev = torch.cuda.Event()
                    ev.record(stream)
                    pending.append((ev, slot))
                    order.append(self.indices[k])
                # return buffers whose calibration has run, then read more
                still = []
                for ev, slot in pending:
                    if ev.query():
                        with self.cond:
                            self.free.append(slot)
                    else:
                        still.append((ev, slot))
                pending = still
                if not self.free and pending and self.next < len(self.indices):
                    ev, slot = pending.pop(0)
                    ev.synchronize()
                    with self.cond:
                        self.free.append(slot)
                self._submit()
            stream.synchronize()
        finally:
            self.pool.shutdown(wait=True)
        return {"schema_version": 1, "frame_count": len(order), "completion_order": order,
                "queue_buffer_count": len(self.buffers), "file_read_s": self.read_s,
                "bytes_read": len(order) * self.frame_bytes, "total_s": time.perf_counter() - self.t0}


def calibrate_host_frames(stack, indices: Sequence[int], lights, light_exposures, dark_exposures, policy,
                          raw_format: str) -> tuple[list, Any]:
    """Queue one host buffer per frame on a fresh stream (uploads overlap
    calibration).  Returns (keep-alive buffers, stream)."""
    import torch

    from . import _native as N
    from .resident import _buffer

    fmt = {"float32": N.GH_FMT_F32, "fits_u16be_bzero": N.GH_FMT_FITS_U16BE_BZERO}[raw_format]
    dtype, nbytes = ((np.float32, stack.pixels_per_frame * 4) if fmt == N.GH_FMT_F32
                     else (np.uint8, stack.pixels_per_frame * 2))
    shape = (stack.height, stack.width) if fmt == N.GH_FMT_F32 else (nbytes,)
    darks = (list(dark_exposures) if isinstance(dark_exposures, (list, tuple, np.ndarray))
             else [dark_exposures] * len(indices))
    stream = torch.cuda.Stream()
    keep = []
    for k, index in enumerate(indices):
        ptr, ref = _buffer(lights[k], dtype, shape, "light")
        stack._calibrate([int(index)], ptr, nbytes, fmt, [light_exposures[k]], [darks[k]], policy, False,
                         stream.cuda_stream)
        keep.append(ref)
    return keep, stream
