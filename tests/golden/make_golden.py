"""Write tests/golden/s2_gate_314.npz from the reference's own fixture files.

The reference ships a complete CPU-path run on a 24x24 synthetic dataset:
inputs under runs/checkpoints/s2_gate_314_synthetic_data/{bias,dark,flat,light}
and the outputs of ``glass audit --backend cpu --tile-size 8``
(runs/checkpoints/s2_gate_314_cpu_run/run_command.txt) under
runs/checkpoints/s2_gate_314_cpu_run/calib_cache/{masters,calibrated} and
.../integration.  This script only converts those data files (inputs and
expected outputs) into one .npz; it imports nothing from the reference.

Run in a container that has the reference mounted:
    python tests/golden/make_golden.py /root/reference
"""

from __future__ import annotations

import json
import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle.fits import read_fits  # noqa: E402


def main(reference_root: str) -> None:
    base = Path(reference_root) / "runs" / "checkpoints"
    data_dir = base / "s2_gate_314_synthetic_data"
    run_dir = base / "s2_gate_314_cpu_run"
    plan = json.loads((run_dir / "processing_plan.json").read_text())
    out: dict[str, np.ndarray] = {}
    exposures: dict[str, float] = {}
    for kind in ("bias", "dark", "flat", "light"):
        frames = []
        for index in range(3):
            data, header = read_fits(data_dir / kind / f"{kind}_{index:03d}.fits")
            frames.append(data.astype(np.float32))
            exposures[f"{kind}_{index}"] = float(header.get("EXPTIME", 0.0))
        out[f"raw_{kind}"] = np.stack(frames)
    masters = run_dir / "calib_cache" / "masters"
    for kind in ("bias", "dark", "flat"):
        (path,) = sorted(masters.glob(f"master_{kind}_*.fits"))
        out[f"master_{kind}"] = read_fits(path)[0].astype(np.float32)
    calibrated = []
    for frame_id in ("F000010", "F000011", "F000012"):
        calibrated.append(read_fits(run_dir / "calib_cache" / "calibrated" / f"calibrated_{frame_id}.fits")[0])
    out["calibrated_light"] = np.stack(calibrated).astype(np.float32)
    out["registered_F000010"] = read_fits(run_dir / "registered_cache" / "registered_F000010.fits")[0]
    out["coverage_F000010"] = read_fits(run_dir / "coverage_cache" / "coverage_F000010.fits")[0]
    for name in ("master_H", "weight_map_H", "coverage_map_H", "low_rejection_H", "high_rejection_H"):
        out[f"integration_{name}"] = read_fits(run_dir / "integration" / f"{name}.fits")[0]
    out["light_exposure_s"] = np.asarray([exposures[f"light_{i}"] for i in range(3)], dtype=np.float64)
    out["dark_exposure_s"] = np.asarray([exposures[f"dark_{i}"] for i in range(3)], dtype=np.float64)
    policy = plan["calibration_plan"]["calibration_policy"]
    out["policy_json"] = np.frombuffer(json.dumps(policy, sort_keys=True).encode(), dtype=np.uint8)
    out["tile_size"] = np.asarray(8, dtype=np.int64)
    np.savez_compressed(HERE / "s2_gate_314.npz", **out)
    # Two of the reference's FITS files, byte for byte, for the FITS reader
    # and writer tests (a raw light and a calibrated light).
    shutil.copyfile(data_dir / "light" / "light_000.fits", HERE / "s2_gate_314_light_000.fits")
    shutil.copyfile(run_dir / "calib_cache" / "calibrated" / "calibrated_F000010.fits",
                    HERE / "s2_gate_314_calibrated_F000010.fits")
    print("wrote", HERE / "s2_gate_314.npz", sorted(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
