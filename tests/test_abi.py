"""C-ABI boundary checks that need no GPU: the library builds/loads, exports
every entry point include/glass_hip.h declares, and the host layer validates
arguments the way the reference bindings do."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import glass_amd
from glass_amd import _native as N
from glass_amd.pipeline import variance_aware_weights

HEADER = Path(__file__).resolve().parents[1] / "include" / "glass_hip.h"


def _declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^int\s+(gh_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    if not N.LIBRARY_PATH.exists():
        from glass_amd.build import build

        build(verbose=False)
    return N.load_library()


def test_library_exports_every_declared_symbol(lib):
    declared = _declared_functions()
    assert len(declared) >= 20
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, missing
    assert sorted(N.EXPORTED_SYMBOLS) == declared


def test_version_and_error_channel(lib):
    assert lib.gh_version() == 1
    # a NULL handle pointer is a value error, reported through gh_last_error
    status = lib.gh_stack_create(1, 1, 1, None)
    assert status == N.GH_ERR_VALUE
    assert "NULL" in N.last_error()


def test_invalid_dimensions_raise_value_error(lib):
    with pytest.raises(ValueError, match="non-empty"):
        N.call("gh_stack_create", 0, 4, 4, ctypes.byref(ctypes.c_void_p()))


def test_host_inverse_and_dark_scale():
    with pytest.raises(ValueError, match="singular"):
        glass_amd.invert_affine(np.zeros((3, 3)))
    m = np.array([[1.0, 0.1, 3.0], [-0.1, 1.0, -2.0], [0.0, 0.0, 1.0]])
    assert np.allclose(glass_amd.invert_affine(m) @ m, np.eye(3))
    # src/glass/cpu/calibration.py:63-66 semantics
    assert glass_amd.dark_scale(30.0, 60.0, None, True) == 0.5
    assert glass_amd.dark_scale(30.0, None, None, True) == 1.0
    assert glass_amd.dark_scale(30.0, 0, None, True) == 1.0
    assert glass_amd.dark_scale(30.0, 60.0, {"dark_scaling_enabled": False}, True) == 1.0
    assert glass_amd.dark_scale(30.0, 60.0, None, False) == 1.0
    assert glass_amd.dark_scale(1.0, 3.0, None, True) == float(np.float32(1.0 / 3.0))


def test_variance_aware_weights_match_oracle():
    import oracle

    rms = [4.0, 2.0, 1.0, 0.5]
    ours = variance_aware_weights(rms)
    ref = oracle.quality_weights([{"background_rms": r} for r in rms], "variance_aware")
    assert np.allclose(ours, np.asarray(ref, np.float32))
    masked = variance_aware_weights(rms, [True, False, True, True])
    assert masked[1] == 0.0 and masked[0] == ours[0]


def test_simple_snr_weights_rule():
    """resident_cuda.py:1948-1955: closed forms of the rule, and the driver's
    vectorised form equals the oracle's per-frame function."""
    import oracle
    from glass_amd.pipeline import simple_snr_weights

    cases = [
        ({"valid_pixels": 0, "mean": 5.0, "std": 1.0}, 0.0),       # no pixels -> weight 0
        ({"valid_pixels": 10, "mean": 5.0, "std": 0.0}, 1.0e6),    # std <= eps -> 1/eps
        ({"valid_pixels": 10, "mean": 5.0, "std": 1.0e-6}, 1.0e6),
        ({"valid_pixels": 10, "mean": -6.0, "std": 2.0}, 3.0),     # |mean| / std
        ({"valid_pixels": 10, "mean": 1.0e-12, "std": 1.0}, 1.0e-6),  # floored at eps
    ]
    for stats, expected in cases:
        assert oracle.simple_snr_weight(stats) == pytest.approx(expected, rel=1e-12)
    got = simple_snr_weights([c[0] for c in cases])
    assert np.array_equal(got, np.asarray([oracle.simple_snr_weight(c[0]) for c in cases], np.float32))
    masked = simple_snr_weights([c[0] for c in cases], [True, True, False, True, True])
    assert masked[2] == 0.0 and masked[3] == np.float32(3.0)
