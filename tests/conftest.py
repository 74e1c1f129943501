from __future__ import annotations

import json
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libglass_hip.so")


@pytest.fixture(scope="session")
def golden_314():
    data = np.load(GOLDEN / "s2_gate_314.npz", allow_pickle=False)
    out = {k: data[k] for k in data.files}
    out["policy"] = json.loads(bytes(out.pop("policy_json")).decode())
    return out


@pytest.fixture(scope="session")
def gpu():
    """Skip-free GPU gate: tests marked gpu must have a working device."""
    import glass_amd

    if not glass_amd.cuda_available():
        pytest.fail("gpu test selected but no HIP device / libglass_hip.so is available")
    return glass_amd
