"""gh_integrate_noalloc (no device allocation, library-owned scratch) against
gh_integrate on the same resident stacks, through the C ABI: bit-identical
maps on consecutive calls, the second and third on smaller inputs, with the
per-tile statistics path, u16 counts, unit and non-unit weights, the
>256-frame LDS kernel, and the argument errors of the no-allocation form."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from glass_amd import ResidentCalibratedStack
from glass_amd import _native as N

pytestmark = pytest.mark.gpu


def _stack_of(frames):
    s = ResidentCalibratedStack(frames.shape[0], frames.shape[1], frames.shape[2])
    for i, f in enumerate(frames):
        s.upload_calibrated_frame(i, f)
    return s


def _frames(seed, f, h, w):
    rng = np.random.default_rng(seed)
    x = rng.normal(1000.0, 4.0, size=(f, h, w)).astype(np.float32)
    hot = rng.random(x.shape) < 0.02
    x[hot] += rng.uniform(50, 500, size=int(hot.sum())).astype(np.float32)
    x[rng.random(x.shape) < 0.01] = np.nan
    return x


def _outputs(torch, npix, count_u16, dq_geometric, active):
    cdt = torch.int16 if count_u16 else torch.float32
    t = {"master": torch.full((npix,), -7.0, device="cuda"), "weight": torch.full((npix,), -7.0, device="cuda"),
         "coverage": torch.full((npix,), 3, dtype=cdt, device="cuda"),
         "low": torch.full((npix,), 3, dtype=cdt, device="cuda"),
         "high": torch.full((npix,), 3, dtype=cdt, device="cuda"),
         "dq": torch.full((npix,), 5, dtype=torch.int32, device="cuda")}
    o = N.IntegrateOutputs()
    for k, v in t.items():
        setattr(o, k, v.data_ptr())
    o.dq_use_geometric = int(dq_geometric)
    o.dq_active_frames = int(active)
    return t, o


def _run(entry, stack, weights, params, count_u16, dq_geometric, active):
    import torch

    npix = stack.height * stack.width
    t, o = _outputs(torch, npix, count_u16, dq_geometric, active)
    wptr = None
    if weights is not None:
        w = np.ascontiguousarray(weights, np.float32)
        wptr = w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    N.call(entry, stack._h, wptr, ctypes.byref(params), ctypes.byref(o), None)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy().copy() for k, v in t.items()}


def _same(a, b):
    for k in a:
        assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), k


def _params(count_u16=False, tile_size=0, positive_scale=0, rejection=N.GH_REJECT_WINSORIZED_SIGMA):
    return N.IntegrateParams(rejection, 3.0, 3.0, 3, 0.5, N.GH_COUNT_U16 if count_u16 else N.GH_COUNT_F32, -1,
                             positive_scale, tile_size)


def test_noalloc_matches_gh_integrate_on_consecutive_calls(gpu):
    import torch  # noqa: F401

    # 1: 40 frames 33 x 47, weights with a zero and a non-unit value, DQ with
    #    the geometric coverage of a warped frame
    frames = _frames(1, 40, 33, 47)
    weights = np.ones(40, np.float32)
    weights[5] = 0.0
    weights[7] = 2.5
    print("noalloc case 1: 40 x 33 x 47, weighted, geometric DQ", flush=True)
    with _stack_of(frames) as s:
        s.reset_warp_coverage()
        for _ in range(39):
            s.accumulate_full_warp_coverage_frame()
        m = np.array([[1.0, 0.0, 0.75], [0.0, 1.0, -0.5], [0.0, 0.0, 1.0]])
        s.apply_matrix_bilinear_frames([39], m[None])
        ref = _run("gh_integrate", s, weights, _params(), False, True, 39)
        got = _run("gh_integrate_noalloc", s, weights, _params(), False, True, 39)
        _same(ref, got)
        assert np.all(np.isfinite(got["master"])) and got["high"].max() > 0

    # 2: smaller, unit weights, u16 counts, per-tile statistics (StackEngine)
    frames = _frames(2, 7, 19, 23)
    print("noalloc case 2: 7 x 19 x 23, unit weights, u16 counts, tiles of 8", flush=True)
    with _stack_of(frames) as s:
        prm = _params(count_u16=True, tile_size=8, positive_scale=1)
        ref = _run("gh_integrate", s, None, prm, True, False, 7)
        got = _run("gh_integrate_noalloc", s, None, prm, True, False, 7)
        _same(ref, got)
        # the rejection-free mean through the same scratch
        prm = _params(rejection=N.GH_REJECT_NONE)
        _same(_run("gh_integrate", s, None, prm, False, False, 7),
              _run("gh_integrate_noalloc", s, None, prm, False, False, 7))

    # 3: more frames than the register kernels hold (LDS sort), small frames
    frames = _frames(3, 300, 5, 9)
    print("noalloc case 3: 300 x 5 x 9 (LDS path)", flush=True)
    with _stack_of(frames) as s:
        ref = _run("gh_integrate", s, None, _params(), False, False, 300)
        got = _run("gh_integrate_noalloc", s, None, _params(), False, False, 300)
        _same(ref, got)


def test_noalloc_argument_errors(gpu):
    frames = _frames(4, 3, 130, 130)
    with _stack_of(frames) as s:
        npix = 130 * 130
        host = np.zeros(npix, np.float32)
        o = N.IntegrateOutputs()
        o.master = host.ctypes.data
        with pytest.raises(ValueError, match="device outputs"):
            N.call("gh_integrate_noalloc", s._h, None, ctypes.byref(_params()), ctypes.byref(o), None)
        import torch

        dev = torch.zeros(npix, device="cuda")
        o = N.IntegrateOutputs()
        o.master = dev.data_ptr()
        # tiles of 1 px: 16,900 statistics tiles, beyond the 16,384 of the scratch
        with pytest.raises(ValueError, match="tile grid"):
            N.call("gh_integrate_noalloc", s._h, None, ctypes.byref(_params(tile_size=1, positive_scale=1)),
                   ctypes.byref(o), None)
