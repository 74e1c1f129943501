"""GPU parity of host-raw calibration through the upload ring (pinned and
pageable host buffers, more frames than ring slots, caller streams) and of
FITS files read into the pinned ring, against the device-raw path and the
oracle (bit-exact)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from glass_amd import ResidentCalibratedStack
from glass_amd import fits as F
from glass_amd.synthetic import encode_fits_u16be_bzero

pytestmark = pytest.mark.gpu

POLICY = {"master_dark_includes_bias": True, "dark_scaling_enabled": True, "flat_floor": 1e-6, "pedestal": 0.0}


def _masters(h, w, seed=1):
    rng = np.random.default_rng(seed)
    return (rng.uniform(90, 110, (h, w)).astype(np.float32), rng.uniform(100, 140, (h, w)).astype(np.float32),
            rng.uniform(0.5, 1.5, (h, w)).astype(np.float32))


def _expected(adu, masters, light_exp, dark_exp):
    bias, dark, flat = masters
    pol = {k: POLICY[k] for k in ("master_dark_includes_bias", "dark_scaling_enabled", "flat_floor", "pedestal")}
    return oracle.calibrate_light(adu, bias, dark, flat, light_exp, dark_exp, **pol)


@pytest.mark.parametrize("pinned", [True, False])
def test_host_raw_ring_matches_device_and_oracle(gpu, pinned):
    import torch

    f, h, w = 7, 48, 64  # more frames than the three ring slots
    rng = np.random.default_rng(4)
    adu = rng.integers(0, 65536, (f, h, w), dtype=np.uint16)
    raw = np.concatenate([np.frombuffer(bytes(encode_fits_u16be_bzero(a)), np.uint8) for a in adu])
    masters = _masters(h, w)
    exps = [30.0 + k for k in range(f)]
    host = torch.from_numpy(raw.copy()).pin_memory() if pinned else raw
    with ResidentCalibratedStack(f, h, w) as s, ResidentCalibratedStack(f, h, w) as d:
        for st in (s, d):
            st.set_calibration_masters(*masters)
        order = [3, 0, 6, 1, 5, 2, 4]
        s.calibrate_frames_host(order, host, "fits_u16be_bzero", exps, 60.0, POLICY)
        d.calibrate_frames_device(order, torch.from_numpy(raw).cuda(), "fits_u16be_bzero", exps, 60.0, POLICY)
        got = s.download_frames_tile(list(range(f)), 0, 0, w, h)
        dev = d.download_frames_tile(list(range(f)), 0, 0, w, h)
    assert np.array_equal(got, dev)
    for k, frame in enumerate(order):
        assert np.array_equal(got[frame], _expected(adu[k], masters, exps[k], 60.0))


def test_host_raw_on_a_caller_stream_then_warp(gpu):
    import torch

    f, h, w = 5, 40, 56
    rng = np.random.default_rng(8)
    adu = rng.integers(1000, 60000, (f, h, w), dtype=np.uint16)
    raw = torch.from_numpy(np.ascontiguousarray(adu)).view(torch.uint8).reshape(-1).pin_memory()
    masters = _masters(h, w, 2)
    stream = torch.cuda.Stream()
    with ResidentCalibratedStack(f, h, w) as s:
        s.set_calibration_masters(*masters)
        s.calibrate_frames_host(list(range(f)), raw, "uint16", [60.0] * f, 60.0, POLICY, stream=stream.cuda_stream)
        shift = np.array([[1.0, 0.0, 2.0], [0.0, 1.0, -1.0], [0.0, 0.0, 1.0]])
        s.warp_frames([1, 3], np.stack([shift, shift]), "bilinear", "cpu", stream=stream.cuda_stream)
        stream.synchronize()
        got = s.download_frames_tile(list(range(f)), 0, 0, w, h)
    for k in range(f):
        exp = _expected(adu[k], masters, 60.0, 60.0)
        if k in (1, 3):
            exp = oracle.warp_matrix(exp, shift, fill=np.nan, interpolation="bilinear")[0]
        assert np.array_equal(got[k], exp, equal_nan=True), k


def test_calibrate_fits_files_through_the_pinned_ring(gpu, tmp_path):
    f, h, w = 5, 36, 52
    rng = np.random.default_rng(12)
    adu = rng.integers(0, 65536, (f, h, w), dtype=np.uint16)
    paths = []
    for k in range(f):
        p = tmp_path / f"light_{k:03d}.fits"
        F.write_u16be_bzero(p, adu[k], {"EXPTIME": 45.0, "IMAGETYP": "light"})
        paths.append(p)
    masters = _masters(h, w, 3)
    released = []
    with ResidentCalibratedStack(f, h, w) as s:
        s.set_calibration_masters(*masters)
        prof = s.calibrate_fits_files([4, 2, 0, 1, 3], paths, [45.0] * f, 90.0, POLICY, ring_frames=2)
        got = s.download_frames_tile([4, 2, 0, 1, 3], 0, 0, w, h)
        # the reference's multistream entry over in-memory raw bytes
        raws = [np.frombuffer(bytes(encode_fits_u16be_bzero(a)), np.uint8) for a in adu]
        s.calibrate_frames_fits_u16be_bzero_host_async_multistream_callback_release_timed(
            list(range(f)), raws, [45.0] * f, [90.0] * f, 2, 2, released.append, POLICY)
        again = s.download_frames_tile(list(range(f)), 0, 0, w, h)
    assert prof["frame_count"] == f and prof["bytes_read"] == f * h * w * 2 and prof["ring_frames"] == 2
    assert released == list(range(f))
    for k in range(f):
        exp = _expected(adu[k], masters, 45.0, 90.0)
        assert np.array_equal(got[k], exp)
        assert np.array_equal(again[k], exp)
    with ResidentCalibratedStack(1, h + 1, w) as s:
        with pytest.raises(ValueError, match="shape_mismatch"):
            s.calibrate_fits_files([0], paths[:1], [45.0], 90.0, POLICY)


def test_async_and_multistream_calibration_entries(gpu):
    f, h, w = 6, 30, 44
    rng = np.random.default_rng(21)
    adu = rng.integers(0, 65536, (f, h, w), dtype=np.uint16)
    lights = [a.astype(np.float32) for a in adu]
    masters = _masters(h, w, 7)
    exps = [20.0 + k for k in range(f)]
    expected = [_expected(adu[k], masters, exps[k], 60.0) for k in range(f)]
    released = []
    with ResidentCalibratedStack(f, h, w) as s:
        s.set_calibration_masters(*masters)
        s.calibrate_frame_pinned_async(0, lights[0], exps[0], 60.0, POLICY)
        s.calibrate_frame_host_async_timed(1, lights[1], exps[1], 60.0, POLICY)
        done = s.finish_pending_calibration_timed()
        assert done["pending"] is True and done["frame_count"] == 2
        assert s.finish_pending_calibration_timed()["pending"] is False
        s.calibrate_frames_host_async_multistream_timed([2, 3], lights[2:4], exps[2:4], [60.0, 60.0], 2, POLICY)
        s.calibrate_frames_host_async_multistream_h2d_release_timed([4], lights[4:5], exps[4:5], [60.0], 2, POLICY)
        s.finish_pending_calibration_timed()
        s.calibrate_frames_host_async_multistream_callback_release_timed([5], lights[5:6], exps[5:6], [60.0], 2, 1,
                                                                         released.append, POLICY)
        got = s.download_frames_tile(list(range(f)), 0, 0, w, h)
    assert released == [0]
    for k in range(f):
        assert np.array_equal(got[k], expected[k]), k


@pytest.mark.parametrize("buffers,workers", [(2, 3), (8, 1)])
def test_fits_completion_queue(gpu, tmp_path, buffers, workers):
    f, h, w = 7, 26, 34
    rng = np.random.default_rng(buffers * 10 + workers)
    adu = rng.integers(0, 65536, (f, h, w), dtype=np.uint16)
    paths, offsets = [], []
    for k in range(f):
        p = tmp_path / f"l{k}.fits"
        F.write_u16be_bzero(p, adu[k])
        spec = F.simple_fits_image_spec(p)
        paths.append(str(p))
        offsets.append(spec.data_offset)
    masters = _masters(h, w, 8)
    order = [6, 2, 4, 0, 1, 5, 3]
    with ResidentCalibratedStack(f, h, w) as s:
        s.set_calibration_masters(*masters)
        prof = s.calibrate_frames_fits_u16be_bzero_paths_completion_queue_timed(
            order, [paths[i] for i in order], [offsets[i] for i in order], [h * w * 2] * f, [30.0] * f, 60.0, 2,
            buffers, workers, POLICY)
        got = s.download_frames_tile(list(range(f)), 0, 0, w, h)
        assert sorted(prof["completion_order"]) == list(range(f)) and prof["bytes_read"] == f * h * w * 2
        # begin / cancel leaves the stack usable; paths form of the multistream entry
        s.begin_fits_u16be_bzero_paths_completion_queue_read_timed([0], paths[:1], offsets[:1], [h * w * 2], 1, 2, 1)
        s.cancel_fits_u16be_bzero_paths_completion_queue_read()
        s.calibrate_frames_fits_u16be_bzero_paths_multistream_timed([1, 0], [paths[0], paths[1]], offsets[:2],
                                                                   [h * w * 2] * 2, [30.0] * 2, [60.0] * 2, 2, 2,
                                                                   POLICY)
        swapped = s.download_frames_tile([0, 1], 0, 0, w, h)
    for k in range(f):
        # frame index `order[j]` holds file `order[j]`
        assert np.array_equal(got[k], _expected(adu[k], masters, 30.0, 60.0)), k
    assert np.array_equal(swapped[1], _expected(adu[0], masters, 30.0, 60.0))
    assert np.array_equal(swapped[0], _expected(adu[1], masters, 30.0, 60.0))
