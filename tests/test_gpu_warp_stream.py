"""GPU parity of the two affine Lanczos3 warp kernels through the C ABI.

The library picks the streamed kernel (register window fed by buffer loads,
per-strip f32 fractions) for near-identity maps and the LDS-tiled kernel
otherwise; GLASS_HIP_WARP=stream / tiled forces one.  Both are held to the
oracle's FP64 warp (values ~100: 2e-4 absolute, the f32 weight tables against
FP64 weights), with identical non-finite patterns and coverage maps, and to
each other (they differ only in f32 rounding)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from glass_amd import ResidentCalibratedStack

pytestmark = pytest.mark.gpu


def _rot(deg, tx, ty, scale=1.0):
    a = np.deg2rad(deg)
    return np.array([[np.cos(a) * scale, -np.sin(a) * scale, tx], [np.sin(a) * scale, np.cos(a) * scale, ty],
                     [0, 0, 1]], np.float64)


def _warp(frames, mats, kernel, monkeypatch, fill=np.nan, resident=False):
    monkeypatch.setenv("GLASS_HIP_WARP", kernel)
    frames = np.asarray(frames, np.float32)
    n, h, w = frames.shape
    with ResidentCalibratedStack(n, h, w) as s:
        for i, f in enumerate(frames):
            s.upload_calibrated_frame(i, f)
        s.reset_warp_coverage()
        if resident:
            s.apply_matrix_lanczos3_frames(list(range(n)), np.stack(mats), fill)
        else:
            s.apply_matrix_frames_cpu_semantics(list(range(n)), np.stack(mats), "lanczos3", fill=fill)
        got = s.download_frames_tile(list(range(n)), 0, 0, w, h)
        geo = s.warp_coverage_map()
    return got, geo


def _check_against_oracle(frames, mats, got, geo, fill=np.nan, atol=2e-4):
    h, w = frames.shape[1:]
    cov_sum = np.zeros((h, w), np.float32)
    for i in range(len(frames)):
        exp, cov = oracle.warp_matrix(frames[i], mats[i], fill=fill, interpolation="lanczos3")
        cov_sum += cov
        fin = np.isfinite(exp)
        assert np.array_equal(np.isfinite(got[i]), fin), i
        nf = ~fin
        assert np.array_equal(got[i][nf], exp[nf].astype(np.float32), equal_nan=True), i
        err = np.max(np.abs(got[i][fin] - exp[fin]), initial=0.0)
        assert err <= atol, (i, err)
    assert np.array_equal(geo, cov_sum)


# Near-identity maps of registration drift: rotations of both signs (the x
# fraction wraps within strips, 7-column windows), none (6 columns), a scale
# that drifts the y fraction, integer and half-pixel shifts, and shifts that
# put the footprint edge through strips and waves.
NEAR_IDENTITY = [_rot(0.8, 3.3, -2.6), _rot(-1.0, -0.4, 1.25), _rot(0.0, 2.0, 1.0), _rot(0.0, 7.5, -3.5),
                 _rot(0.3, -10.25, 6.7, 1.0003), _rot(-0.05, 0.5, 0.5, 0.9997), _rot(1.0, -60.0, 45.0),
                 _rot(-0.7, 33.0, -20.0)]


@pytest.mark.parametrize("kernel", ["stream", "tiled"])
def test_near_identity_maps_match_oracle(gpu, monkeypatch, kernel):
    rng = np.random.default_rng(31)
    h, w = 203, 333  # partial tiles in both directions, lanes past the last column
    frames = rng.normal(100, 10, size=(len(NEAR_IDENTITY), h, w)).astype(np.float32)
    got, geo = _warp(frames, NEAR_IDENTITY, kernel, monkeypatch)
    _check_against_oracle(frames, NEAR_IDENTITY, got, geo)
    # the whole-pixel shift (2, 1) copies its source exactly: output (y, x)
    # takes source (y - 1, x - 2)
    _, cov = oracle.warp_matrix(frames[2], NEAR_IDENTITY[2], fill=np.nan, interpolation="lanczos3")
    shifted = np.full((h, w), np.nan, np.float32)
    shifted[1:, 2:] = frames[2][:-1, :-2]
    assert np.array_equal(got[2][cov > 0], shifted[cov > 0])


def test_stream_and_tiled_kernels_agree(gpu, monkeypatch):
    rng = np.random.default_rng(32)
    h, w = 256, 384
    frames = rng.normal(100, 10, size=(len(NEAR_IDENTITY), h, w)).astype(np.float32)
    a, geo_a = _warp(frames, NEAR_IDENTITY, "stream", monkeypatch)
    b, geo_b = _warp(frames, NEAR_IDENTITY, "tiled", monkeypatch)
    assert np.array_equal(geo_a, geo_b)
    assert np.array_equal(np.isfinite(a), np.isfinite(b))
    fin = np.isfinite(a)
    assert np.max(np.abs(a[fin] - b[fin])) <= 1e-4


@pytest.mark.parametrize("kernel", ["stream", "tiled"])
def test_general_affine_maps_match_oracle(gpu, monkeypatch, kernel):
    # Maps far from the identity: every strip of the streamed kernel takes the
    # exact row-by-row form (rows skipped or repeated, fractions wrapping many
    # times); 90 degrees, a zoom, 40 degrees, a shear.
    rng = np.random.default_rng(33)
    h, w = 140, 150
    frames = rng.normal(100, 10, size=(4, h, w)).astype(np.float32)
    shear = np.array([[1.0, 0.35, -20.0], [0.0, 1.0, 2.0], [0, 0, 1.0]])
    mats = [_rot(90.0, 145.0, -2.0), _rot(0.5, -30.0, -20.0, 1.4), _rot(40.0, 60.0, -50.0), shear]
    got, geo = _warp(frames, mats, kernel, monkeypatch)
    _check_against_oracle(frames, mats, got, geo)


@pytest.mark.parametrize("kernel", ["stream", "tiled"])
def test_nonfinite_sources_and_resident_semantics(gpu, monkeypatch, kernel):
    # Frames holding NaN / inf take the per-row paths (CPU semantics: table
    # rows, then non-finite results per tap); resident semantics skip
    # non-finite taps.  A finite frame in the same launch stays on the
    # streamed rows.
    rng = np.random.default_rng(34)
    h, w = 130, 210
    frames = rng.normal(100, 10, size=(3, h, w)).astype(np.float32)
    frames[0, 60:62, :] = np.nan
    frames[0, 20, 40] = np.inf
    ys, xs = rng.integers(5, h - 5, 30), rng.integers(5, w - 5, 30)
    frames[1, ys, xs] = np.nan
    mats = [_rot(0.5, 1.7, -2.3), _rot(-0.2, 3.0, -2.0), _rot(0.9, -4.0, 2.5)]
    got, geo = _warp(frames, mats, kernel, monkeypatch)
    _check_against_oracle(frames, mats, got, geo)
    got_r, geo_r = _warp(frames, mats, kernel, monkeypatch, resident=True)
    cov_sum = np.zeros((h, w), np.float32)
    for i in range(3):
        exp, cov = oracle.warp_matrix_resident(frames[i], mats[i], np.nan, "lanczos3")
        cov_sum += cov
        ok = cov > 0
        assert np.array_equal(np.isfinite(got_r[i]), ok), i
        assert np.allclose(got_r[i][ok], exp[ok], rtol=0, atol=1e-3), (i, np.max(np.abs(got_r[i][ok] - exp[ok])))
    assert np.array_equal(geo_r, cov_sum)


def test_stream_kernel_workload_geometry(gpu, monkeypatch):
    # The stacking workload's drift (up to 60 px and 1 degree) on a frame of
    # star-like peaks over a sky level, whose sharp gradients make a wrong tap
    # or weight visible; compared on every pixel with the FP64 oracle warp.
    rng = np.random.default_rng(35)
    h, w = 640, 1024
    yy, xx = np.indices((h, w), dtype=np.float64)
    img = np.full((h, w), 0.02)
    for _ in range(300):
        cx, cy, s = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(0.8, 1.7)
        img += rng.uniform(0.05, 0.9) * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
    img += rng.normal(0, 0.01, size=(h, w))
    frames = np.stack([img] * 4).astype(np.float32)
    mats = [_rot(rng.uniform(-1, 1), rng.uniform(-60, 60), rng.uniform(-60, 60)) for _ in range(4)]
    got, geo = _warp(frames, mats, "stream", monkeypatch)
    _check_against_oracle(frames, mats, got, geo, atol=2e-6)
