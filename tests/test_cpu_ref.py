"""oracle/cpu_ref.c (the C/OpenMP restatement timed as bench.py's CPU
baseline) against the numpy oracle on the same seeded inputs (CPU only)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import cpu_ref


@pytest.fixture(scope="module")
def ref():
    cpu_ref.build(force=False)
    return cpu_ref


def test_calibrate_matches_oracle_bit_exact(ref):
    rng = np.random.default_rng(11)
    h, w, n = 17, 23, 3
    adu = rng.integers(0, 65536, size=(n, h, w), dtype=np.uint16)
    raw = oracle.calibration.encode_fits_u16be_bzero(adu)
    bias = rng.normal(500, 5, (h, w)).astype(np.float32)
    dark = rng.normal(40, 2, (h, w)).astype(np.float32)
    flat = rng.uniform(0.5, 1.2, (h, w)).astype(np.float32)
    flat[0, :3] = [0.0, -1.0, 1e-9]   # below the floor
    for includes_bias, pedestal in ((True, 0.0), (False, 100.0)):
        scale = oracle.calibration.dark_scale_factor(30.0, 60.0)
        got = ref.calibrate_u16be(raw, h * w, bias, dark, flat, np.float32(scale), includes_bias, 1e-6, pedestal)
        for k in range(n):
            exp = oracle.calibrate_light(adu[k].astype(np.float32), bias, dark, flat, 30.0, 60.0,
                                         master_dark_includes_bias=includes_bias, flat_floor=1e-6, pedestal=pedestal)
            assert np.array_equal(got[k].reshape(h, w), exp), k


@pytest.mark.parametrize("interpolation", ["bilinear", "lanczos3"])
def test_warp_matches_oracle(ref, interpolation):
    rng = np.random.default_rng(12)
    img = rng.normal(100, 10, (31, 37)).astype(np.float32)
    a = np.deg2rad(0.8)
    m = np.array([[np.cos(a), -np.sin(a), 2.37], [np.sin(a), np.cos(a), -1.61], [0, 0, 1.0]])
    got, cov = ref.warp(img, np.linalg.inv(m), interpolation, fill=np.nan)
    exp, ecov = oracle.warp_matrix(img, m, fill=np.nan, interpolation=interpolation)
    assert np.array_equal(cov, ecov)
    ok = ecov > 0
    assert np.array_equal(np.isnan(got), np.isnan(exp))
    if interpolation == "bilinear":
        assert np.array_equal(got[ok], exp[ok])
    else:   # FP64 weights and sums in both; summation order may differ in the last bit
        assert np.allclose(got[ok], exp[ok], rtol=1e-6, atol=0)


@pytest.mark.parametrize("frames", [5, 24])
def test_integrate_matches_oracle(ref, frames):
    rng = np.random.default_rng(13 + frames)
    stack = rng.normal(1000, 4, (frames, 9, 11)).astype(np.float32)
    hot = rng.random(stack.shape) < 0.03
    stack[hot] += rng.uniform(50, 400, int(hot.sum())).astype(np.float32)
    stack[rng.random(stack.shape) < 0.05] = np.nan   # per-pixel valid-count arithmetic
    weights = rng.uniform(0.5, 2.0, frames).astype(np.float32)
    weights[1] = 0.0
    got = ref.integrate_winsorized(stack, weights)
    exp = oracle.weighted_integrate_stack(stack, weights=weights, rejection="winsorized_sigma")
    for name, g, e in zip(("master", "weight", "coverage", "low", "high"), got, exp):
        if name == "master":
            assert np.allclose(g, e, rtol=2e-7, atol=0), name
        else:
            assert np.array_equal(g, e), name


def test_warp_map64_matches_oracle(ref):
    from oracle.warp import warp_lanczos3_source_map

    rng = np.random.default_rng(14)
    img = rng.normal(100, 10, (29, 35)).astype(np.float32)
    b = np.array([0.9998, -0.017, 1.3, 0.017, 0.9998, -0.8])
    got = ref.warp_map64(img, b, 0.0)
    exp = warp_lanczos3_source_map(img, b, 0.0, np.float64)
    assert np.array_equal(got == 0.0, exp == 0.0)
    assert np.allclose(got, exp, rtol=1e-14, atol=0)
