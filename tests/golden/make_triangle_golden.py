"""Write tests/golden/triangle_golden.npz: input/output vectors of the
reference's FP64 CPU bridge for local triangle-asterism descriptors,
``glass.cpu.registration._local_triangle_descriptors``
(src/glass/cpu/registration.py:236-303).

The reference checks its GPU descriptors against this function
(tests/test_gpu_registration_search.py:87-108: indices equal, descriptors and
areas allclose); these vectors let the oracle and the HIP path be held to the
same check without the reference at run time.  This script imports the
reference module in a container where the reference is mounted, runs it on the
reference test's literal catalogue and on seeded catalogues, and stores inputs
and outputs (data only).

Usage (from the repository root):
    python tests/golden/make_triangle_golden.py /root/reference
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

# tests/test_gpu_registration_search.py:74-84
LITERAL = [(10.0, 10.0), (20.0, 11.0), (13.0, 24.0), (32.0, 31.0), (6.0, 38.0), (38.0, 7.0), (27.0, 19.0),
           (16.0, 34.0)]


def cases(rng):
    """(name, float32 xy, max_stars, neighbors, max_descriptors)."""
    out = [("literal", np.asarray(LITERAL, np.float32), 8, 5, 64)]
    out.append(("field64", rng.uniform(20, 2000, (64, 2)).astype(np.float32), 64, 5, 1200))
    out.append(("field80_n8_cut", rng.uniform(0, 1500, (90, 2)).astype(np.float32), 80, 8, 300))
    out.append(("field30_n3", rng.uniform(0, 400, (30, 2)).astype(np.float32), 80, 3, 1200))
    out.append(("field24_n12", rng.uniform(0, 600, (24, 2)).astype(np.float32), 80, 12, 1200))
    # close pairs (area gate) and a collinear run (degenerate triangles)
    pts = rng.uniform(0, 800, (36, 2))
    pts[5] = pts[4] + (0.6, 0.3)
    pts[11] = pts[10] + (0.0, 0.9)
    pts[20:26] = np.stack([np.linspace(100, 400, 6), np.linspace(50, 350, 6)], axis=1)
    out.append(("close_and_collinear", pts.astype(np.float32), 80, 6, 1200))
    out.append(("more_stars_than_cap", rng.uniform(0, 1000, (50, 2)).astype(np.float32), 20, 5, 1200))
    out.append(("neighbors_above_count", rng.uniform(0, 300, (5, 2)).astype(np.float32), 80, 9, 1200))
    return out


def main(reference_root: str) -> None:
    sys.path.insert(0, str(Path(reference_root) / "src"))
    from glass.cpu.registration import _local_triangle_descriptors

    rng = np.random.default_rng(20261204)
    out: dict[str, np.ndarray] = {}
    names = []
    for name, xy, max_stars, neighbors, max_descriptors in cases(rng):
        points = xy.astype(np.float64)
        got = _local_triangle_descriptors(points, max_points=max_stars, neighbors=neighbors,
                                          max_descriptors=max_descriptors)
        key = f"tri/{name}"
        out[f"{key}/xy"] = xy
        out[f"{key}/args"] = np.asarray([max_stars, neighbors, max_descriptors], np.int64)
        out[f"{key}/descriptors"] = np.asarray([item[0] for item in got], np.float64).reshape(-1, 2)
        out[f"{key}/indices"] = np.asarray([item[1] for item in got], np.int64).reshape(-1, 3)
        out[f"{key}/areas"] = np.asarray([item[2] for item in got], np.float64).reshape(-1)
        names.append(name)
    out["names"] = np.asarray(names)
    np.savez_compressed(HERE / "triangle_golden.npz", **out)
    print(f"wrote {HERE / 'triangle_golden.npz'}: {len(names)} catalogues")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
