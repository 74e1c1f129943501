"""GPU parity of the triangle-asterism / star-pair similarity registration
(glass_amd.triangles over libglass_hip.so) against the oracle restatement
(oracle/triangles.py) on the same inputs.

Bar: descriptors, indices, areas, inlier counts, candidate indices, RMS values,
matrices, refit statistics and the top-k list bit-exact; the rotation angle
(atan2f, not correctly rounded on either side) within 4 float32 ulp.  Plus the
reference tests' own assertions (tests/test_gpu_registration_search.py:87-310,
:982-1120) and the golden vectors of the reference's CPU bridge.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from glass_amd import triangles as G
from oracle import triangles as T

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden" / "triangle_golden.npz"
LITERAL = np.asarray([(10.0, 10.0), (20.0, 11.0), (13.0, 24.0), (32.0, 31.0), (6.0, 38.0), (38.0, 7.0),
                      (27.0, 19.0), (16.0, 34.0)], np.float64)


def similarity(angle_deg, scale, t):
    a = np.deg2rad(angle_deg)
    return scale * np.asarray([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]]), np.asarray(t, np.float64)


def assert_descriptors_equal(got, exp):
    for k in ("count", "raw_count", "neighbors", "max_stars"):
        assert got[k] == exp[k], k
    assert np.array_equal(got["indices"], exp["indices"])
    assert np.array_equal(got["descriptors"].view(np.uint32), exp["descriptors"].view(np.uint32))
    assert np.array_equal(got["areas"].view(np.uint32), exp["areas"].view(np.uint32))


def ulp_close(a, b, n=4):
    a, b = np.float32(a), np.float32(b)
    return a == b or abs(float(a) - float(b)) <= n * float(np.spacing(np.maximum(abs(a), abs(b))))


def assert_fit_equal(got, exp):
    assert got["inliers"] == exp["inliers"]
    assert got["best_candidate_index"] == exp["best_candidate_index"]
    assert got["candidate_count"] == exp["candidate_count"]
    assert got["scored_candidates"] == exp["scored_candidates"]
    assert got["status"] == exp["status"]
    g = np.asarray(got["matrix"], np.float32)
    assert np.array_equal(g.view(np.uint32), np.asarray(exp["matrix"], np.float32).view(np.uint32)), (g, exp["matrix"])
    assert np.float32(got["scale"]) == np.float32(exp["scale"])
    assert ulp_close(got["rotation_rad"], exp["rotation_rad"])
    assert (np.isnan(got["rms_px"]) and np.isnan(exp["rms_px"])) or np.float32(got["rms_px"]) == exp["rms_px"]


def field(seed, n, span=2000.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(20, span, (n, 2))


def moved(ref, angle, scale, t, seed, jitter=0.05, drop=4, extra=6, span=2000.0):
    """Moving catalogue of a reference under a similarity (moving ->
    reference = (lin, t)), jittered, with stars dropped and spurious ones added,
    shuffled."""
    rng = np.random.default_rng(seed)
    lin, tt = similarity(angle, scale, t)
    mov = (ref - tt) @ np.linalg.inv(lin).T + rng.normal(0, jitter, ref.shape)
    keep = np.ones(len(ref), bool)
    keep[rng.choice(len(ref), drop, replace=False)] = False
    mov = np.vstack([mov[keep], rng.uniform(20, span, (extra, 2))])
    return mov[rng.permutation(len(mov))]


# ------------------------------------------------------------ descriptors


def test_descriptors_golden_catalogues(gpu):
    z = np.load(GOLDEN)
    for name in z["names"]:
        xy = z[f"tri/{name}/xy"]
        ms, nb, md = (int(v) for v in z[f"tri/{name}/args"])
        got = G.triangle_asterism_descriptors_f32(xy[:, 0], xy[:, 1], ms, nb, md)
        assert_descriptors_equal(got, T.triangle_descriptors(xy[:, 0], xy[:, 1], ms, nb, md))
    # the reference's own check against its CPU bridge (literal catalogue)
    got = G.triangle_asterism_descriptors_f32(LITERAL[:, 0].astype(np.float32), LITERAL[:, 1].astype(np.float32),
                                              max_stars=8, neighbors=5, max_descriptors=64)
    assert got["count"] == len(z["tri/literal/indices"]) and got["neighbors"] == 5 and got["raw_count"] == 80
    assert np.array_equal(got["indices"], z["tri/literal/indices"])
    assert np.allclose(got["descriptors"], z["tri/literal/descriptors"].astype(np.float32))
    assert np.allclose(got["areas"], z["tri/literal/areas"].astype(np.float32))


@pytest.mark.parametrize("seed,n,max_stars,nb,max_desc", [
    (1, 64, 64, 5, 1200), (2, 200, 80, 5, 1200), (3, 300, 300, 16, 5000), (4, 40, 80, 9, 100),
    (5, 3, 80, 5, 10), (6, 70, 70, 4, 2), (7, 130, 130, 7, 1200)])
def test_descriptors_random_catalogues(gpu, seed, n, max_stars, nb, max_desc):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(0, 1500, (n, 2)).astype(np.float32)
    if n > 10:
        xy[3] = np.nan
        xy[7, 1] = np.inf
        xy[9] = xy[8] + np.float32(0.5)  # area gate
    got = G.triangle_asterism_descriptors_f32(xy[:, 0], xy[:, 1], max_stars, nb, max_desc)
    assert_descriptors_equal(got, T.triangle_descriptors(xy[:, 0], xy[:, 1], max_stars, nb, max_desc))


def test_descriptors_exact_ties(gpu):
    """Integer grid: many equal neighbour distances and equal areas."""
    rng = np.random.default_rng(9)
    for nb in (3, 5, 8, 16):
        xy = rng.integers(0, 30, (90, 2)).astype(np.float32)
        got = G.triangle_asterism_descriptors_f32(xy[:, 0], xy[:, 1], 90, nb, 3000)
        assert_descriptors_equal(got, T.triangle_descriptors(xy[:, 0], xy[:, 1], 90, nb, 3000))


def test_descriptor_edges(gpu):
    x = np.asarray([0, 10, 0], np.float32)
    y = np.asarray([0, 0, 10], np.float32)
    for kw in ({}, {"max_descriptors": 0}, {"max_stars": 2}, {"neighbors": 40}):
        assert_descriptors_equal(G.triangle_asterism_descriptors_f32(x, y, **kw), T.triangle_descriptors(x, y, **kw))
    empty = G.triangle_asterism_descriptors_f32(np.zeros(0), np.zeros(0))
    assert empty["count"] == 0 and empty["descriptors"].shape == (0, 2) and empty["indices"].shape == (0, 3)
    with pytest.raises(ValueError, match="max_stars"):
        G.triangle_asterism_descriptors_f32(x, y, max_stars=-1)
    with pytest.raises(ValueError, match="max_descriptors"):
        G.triangle_asterism_descriptors_f32(x, y, max_descriptors=-1)
    with pytest.raises(ValueError, match="same shape"):
        G.triangle_asterism_descriptors_f32(x, y[:2])


def test_descriptor_batch_matches_single_catalogs(gpu):
    """tests/test_gpu_registration_search.py:111-154 plus catalogues of
    unequal sizes and an empty one."""
    reference = LITERAL
    catalogs = [reference, reference + np.asarray([1.25, -0.75]), reference[:4] + np.asarray([-2.0, 1.5]),
                field(3, 150), np.zeros((0, 2)), field(4, 2), field(5, 64)]
    batch = G.triangle_asterism_descriptors_batch_f32([c[:, 0] for c in catalogs], [c[:, 1] for c in catalogs],
                                                      max_stars=80, neighbors=5, max_descriptors=300)
    assert [b["batch_index"] for b in batch] == list(range(len(catalogs)))
    assert all(b["batch_count"] == len(catalogs) for b in batch)
    for b, c in zip(batch, catalogs):
        single = G.triangle_asterism_descriptors_f32(c[:, 0], c[:, 1], 80, 5, 300)
        assert_descriptors_equal(b, single)
        assert_descriptors_equal(b, T.triangle_descriptors(c[:, 0], c[:, 1], 80, 5, 300))
    assert G.triangle_asterism_descriptors_batch_f32([], []) == []


# ------------------------------------------------------------ triangle similarity


def descriptors(xy, max_stars=80, nb=5, max_desc=1200):
    xy = np.asarray(xy, np.float32)
    return T.triangle_descriptors(xy[:, 0], xy[:, 1], max_stars, nb, max_desc)


def both_triangle_fits(ref, mov, tol, radius, max_stars=80, nb=5, max_desc=1200):
    ref, mov = np.asarray(ref, np.float32), np.asarray(mov, np.float32)
    rd, md = descriptors(ref, max_stars, nb, max_desc), descriptors(mov, max_stars, nb, max_desc)
    args = (ref[:, 0], ref[:, 1], mov[:, 0], mov[:, 1], rd["descriptors"], rd["indices"], md["descriptors"],
            md["indices"])
    got = G.estimate_similarity_from_triangle_descriptors_f32(*args, tolerance_px=tol, descriptor_radius=radius)
    exp = T.similarity_from_triangles(*args, tolerance_px=tol, descriptor_radius=radius)
    return got, exp, rd, md


def test_reference_triangle_similarity_literal(gpu):
    """tests/test_gpu_registration_search.py:157-208."""
    lin, t = similarity(6.0, 1.015, (3.5, -2.25))
    mov = (LITERAL - t) @ np.linalg.inv(lin).T
    got, exp, rd, md = both_triangle_fits(LITERAL, mov, 0.1, 0.01, 8, 5, 64)
    assert_fit_equal(got, exp)
    assert got["status"] == "ok" and got["inliers"] == len(LITERAL)
    assert got["candidate_count"] == rd["count"] * md["count"] * 2
    assert got["rms_px"] < 1.0e-3
    m = np.asarray(got["matrix"], np.float64)
    assert np.allclose(m[:2, :2], lin, atol=1e-4) and np.allclose(m[:2, 2], t, atol=1e-4)


@pytest.mark.parametrize("seed,n,angle,scale,t,tol,radius", [
    (11, 64, 0.4, 1.0002, (12.5, -7.25), 2.0, 0.1),     # the resident default shape
    (12, 64, -3.0, 0.995, (-40.0, 25.0), 1.0, 0.05),
    (13, 80, 15.0, 1.03, (5.0, 5.0), 3.0, 0.1),
    (14, 30, 0.0, 1.0, (0.75, -0.5), 0.5, 0.02),
    (15, 120, 1.0, 1.0, (100.0, -60.0), 2.0, 0.1),
])
def test_triangle_similarity_random_fields(gpu, seed, n, angle, scale, t, tol, radius):
    ref = field(seed, n)
    mov = moved(ref, angle, scale, t, seed + 100)
    got, exp, _, _ = both_triangle_fits(ref, mov, tol, radius)
    assert_fit_equal(got, exp)
    assert got["status"] == "ok"
    lin, tt = similarity(angle, scale, t)
    m = np.asarray(got["matrix"], np.float64)
    assert np.allclose(m[:2, :2], lin, atol=2e-3) and np.allclose(m[:2, 2], tt, atol=2.0)


def test_triangle_similarity_nonfinite_and_failures(gpu):
    ref = field(21, 50)
    mov = moved(ref, 2.0, 1.0, (3.0, 4.0), 22)
    ref = ref.astype(np.float32)
    mov = mov.astype(np.float32)
    ref[5] = np.nan
    mov[[2, 9]] = np.inf
    got, exp, rd, md = both_triangle_fits(ref, mov, 2.0, 0.1)
    assert_fit_equal(got, exp)
    # descriptors pointing at non-finite or out-of-range stars are skipped
    rdi = rd["indices"].copy()
    rdi[0] = [0, 1, 999]
    args = (ref[:, 0], ref[:, 1], mov[:, 0], mov[:, 1], rd["descriptors"], rdi, md["descriptors"], md["indices"])
    assert_fit_equal(G.estimate_similarity_from_triangle_descriptors_f32(*args, 2.0, 0.1),
                     T.similarity_from_triangles(*args, 2.0, 0.1))
    # nothing within the descriptor radius: failed, identity
    far = np.full_like(md["descriptors"], 50.0)
    args = (ref[:, 0], ref[:, 1], mov[:, 0], mov[:, 1], rd["descriptors"], rd["indices"], far, md["indices"])
    got = G.estimate_similarity_from_triangle_descriptors_f32(*args, 2.0, 0.1)
    assert_fit_equal(got, T.similarity_from_triangles(*args, 2.0, 0.1))
    assert got["status"] == "failed" and got["best_candidate_index"] == -1
    with pytest.raises(ValueError, match="at least one descriptor"):
        G.estimate_similarity_from_triangle_descriptors_f32(ref[:2, 0], ref[:2, 1], *args[2:])
    with pytest.raises(ValueError, match="tolerance_px"):
        G.estimate_similarity_from_triangle_descriptors_f32(*args, -1.0, 0.1)
    with pytest.raises(ValueError, match="row counts"):
        G.estimate_similarity_from_triangle_descriptors_f32(*args[:5], rd["indices"][:-1], *args[6:])


def test_triangle_similarity_large_catalogues_global_path(gpu):
    """Catalogues too large for the LDS staging (1700 moving stars): the
    scorer's global-memory variant."""
    ref = field(31, 260, span=4000.0)
    mov = moved(ref, 0.7, 1.001, (9.0, -4.0), 32, drop=10, extra=1450, span=4000.0)
    ref32, mov32 = ref.astype(np.float32), mov.astype(np.float32)
    rd = descriptors(ref32, 80, 5, 400)
    md = descriptors(mov32, 80, 5, 400)
    args = (ref32[:, 0], ref32[:, 1], mov32[:, 0], mov32[:, 1], rd["descriptors"], rd["indices"], md["descriptors"],
            md["indices"])
    got = G.estimate_similarity_from_triangle_descriptors_f32(*args, 2.0, 0.004)
    assert_fit_equal(got, T.similarity_from_triangles(*args, 2.0, 0.004))


def test_triangle_similarity_batch_matches_single_fits(gpu):
    """tests/test_gpu_registration_search.py:211-310 plus random frames."""
    reference = LITERAL
    rd = descriptors(reference, 8, 5, 64)
    moving = []
    for angle, scale, t in [(4.0, 1.01, (2.5, -1.25)), (-3.0, 0.992, (-1.75, 3.0))]:
        lin, tt = similarity(angle, scale, t)
        moving.append((reference - tt) @ np.linalg.inv(lin).T)
    mds = [descriptors(m, 8, 5, 64) for m in moving]
    batch = G.estimate_similarity_from_triangle_descriptors_batch_f32(
        reference[:, 0], reference[:, 1], rd["descriptors"], rd["indices"], [m[:, 0] for m in moving],
        [m[:, 1] for m in moving], [d["descriptors"] for d in mds], [d["indices"] for d in mds], 0.1, 0.01)
    assert [b["batch_index"] for b in batch] == [0, 1] and [b["batch_count"] for b in batch] == [2, 2]
    for b, m, d in zip(batch, moving, mds):
        single = G.estimate_similarity_from_triangle_descriptors_f32(
            reference[:, 0], reference[:, 1], m[:, 0], m[:, 1], rd["descriptors"], rd["indices"], d["descriptors"],
            d["indices"], 0.1, 0.01)
        assert b["status"] == "ok" and b["reference_device_reuse"] is True
        assert_fit_equal(b, single)
    ref = field(41, 64)
    rd = descriptors(ref)
    frames = [moved(ref, a, 1.0, (a * 3, -a), 50 + i) for i, a in enumerate((0.2, -0.5, 1.5, 0.0, 2.5))]
    mds = [descriptors(f) for f in frames]
    batch = G.estimate_similarity_from_triangle_descriptors_batch_f32(
        ref[:, 0], ref[:, 1], rd["descriptors"], rd["indices"], [f[:, 0] for f in frames], [f[:, 1] for f in frames],
        [d["descriptors"] for d in mds], [d["indices"] for d in mds], 2.0, 0.1)
    for b, f, d in zip(batch, frames, mds):
        f32, r32 = f.astype(np.float32), ref.astype(np.float32)
        exp = T.similarity_from_triangles(r32[:, 0], r32[:, 1], f32[:, 0], f32[:, 1], rd["descriptors"],
                                          rd["indices"], d["descriptors"], d["indices"], 2.0, 0.1)
        assert_fit_equal(b, exp)


# ------------------------------------------------------------ catalogue similarity

MOVING8 = [(10.0, 11.0), (25.0, 14.0), (18.0, 34.0), (41.0, 37.0), (53.0, 21.0), (63.0, 45.0), (31.0, 58.0),
           (72.0, 18.0)]


def catalog_case(points, angle, scale, t, jitter, extra_m, extra_r):
    lin = similarity(angle, scale, t)[0].astype(np.float32)
    moving = np.asarray(points, np.float32)
    reference = moving @ lin.T + np.asarray(t, np.float32) + jitter
    return (np.vstack([reference, np.asarray(extra_r, np.float32)]).astype(np.float32),
            np.vstack([moving, np.asarray(extra_m, np.float32)]).astype(np.float32), lin, np.asarray(t, np.float32))


def assert_catalog_equal(got, exp):
    assert_fit_equal(got, exp)
    for k in ("refined_inliers", "refit_status", "refit_status_code", "top_k", "prior_dx", "prior_dy",
              "max_abs_rotation_rad"):
        assert got[k] == exp[k], k
    assert (np.isnan(got["refit_rms_px"]) and np.isnan(exp["refit_rms_px"])) or \
        np.float32(got["refit_rms_px"]) == exp["refit_rms_px"]
    assert len(got["top_candidates"]) == len(exp["top_candidates"])
    for g, e in zip(got["top_candidates"], exp["top_candidates"]):
        assert g["candidate_index"] == e["candidate_index"] and g["inliers"] == e["inliers"]
        assert np.float32(g["rms_px"]) == e["rms_px"]
        assert np.array_equal(np.asarray(g["matrix"], np.float32), e["matrix"])


def both_catalog_fits(ref, mov, **kw):
    args = (ref[:, 0], ref[:, 1], mov[:, 0], mov[:, 1])
    return G.estimate_similarity_from_catalogs_f32(*args, **kw), T.similarity_from_catalogs(*args, **kw)


def test_reference_catalog_pair_candidates(gpu):
    """tests/test_gpu_registration_search.py:982-1048."""
    ref, mov, lin, t = catalog_case(MOVING8, -4.0, 0.985, (-1.75, 4.5), 0.0, [[90.0, 90.0], [5.0, 70.0]],
                                    [[4.0, 4.0], [80.0, 8.0]])
    got, exp = both_catalog_fits(ref, mov, tolerance_px=0.15, min_pair_distance=3.0, prior_dx=float(t[0]),
                                 prior_dy=float(t[1]), prior_radius_px=2.0, min_scale=0.95, max_scale=1.05,
                                 max_abs_rotation_rad=0.2, top_k=4)
    assert_catalog_equal(got, exp)
    assert got["status"] == "ok" and got["model"] == "catalog_pair_similarity_hip"
    assert got["inliers"] >= 8 and got["refined_inliers"] >= 8 and got["refit_status"] in {"ok", "rejected"}
    assert got["top_k"] == 4 and len(got["top_candidates"]) == 4
    assert got["top_candidates"][0]["inliers"] >= got["top_candidates"][-1]["inliers"]
    assert got["candidate_count"] == len(ref) * (len(ref) - 1) * len(mov) * (len(mov) - 1)
    assert got["prior_radius_px"] == 2.0
    assert abs(got["min_scale"] - 0.95) < 1e-6 and abs(got["max_scale"] - 1.05) < 1e-6
    assert abs(got["max_abs_rotation_rad"] - 0.2) < 1e-6
    assert abs(got["scale"] - 0.985) < 1e-4 and abs(got["rotation_rad"] - np.deg2rad(-4.0)) < 1e-4
    assert got["rms_px"] < 1e-3
    m = np.asarray(got["matrix"], np.float32)
    assert np.allclose(m[:2, :2], lin, atol=1e-4) and np.allclose(m[:2, 2], t, atol=1e-3)


def test_reference_catalog_refit(gpu):
    """tests/test_gpu_registration_search.py:1051-1118."""
    moving = MOVING8 + [(84.0, 52.0), (48.0, 74.0)]
    jitter = np.asarray([(0.04, -0.02), (-0.03, 0.02), (0.01, 0.03), (-0.02, -0.04), (0.03, 0.01), (-0.01, -0.03),
                         (0.02, 0.04), (-0.04, 0.00), (0.01, -0.01), (0.00, 0.02)], np.float32)
    ref, mov, lin, t = catalog_case(moving, 2.5, 1.01, (3.5, -2.25), jitter, [[2.0, 88.0], [95.0, 9.0]],
                                    [[88.0, 4.0], [5.0, 92.0]])
    got, exp = both_catalog_fits(ref, mov, tolerance_px=0.25, min_pair_distance=5.0, prior_dx=float(t[0]),
                                 prior_dy=float(t[1]), prior_radius_px=2.0, min_scale=0.98, max_scale=1.04,
                                 max_abs_rotation_rad=0.1)
    assert_catalog_equal(got, exp)
    assert got["status"] == "ok" and got["refit_status"] == "ok" and got["refined_inliers"] >= 10
    assert got["rms_px"] < 0.08
    m = np.asarray(got["matrix"], np.float32)
    assert np.allclose(m[:2, :2], lin, atol=0.003) and np.allclose(m[:2, 2], t, atol=0.12)


@pytest.mark.parametrize("seed,nr,nm,kw", [
    (61, 20, 22, dict(tolerance_px=1.0, top_k=7)),
    (62, 40, 36, dict(tolerance_px=2.0, prior_dx=6.0, prior_dy=-3.0, prior_radius_px=5.0, top_k=12)),
    (63, 12, 14, dict(tolerance_px=3.0, min_scale=0.9, max_scale=1.1, max_abs_rotation_rad=0.5, top_k=1)),
    (64, 20, 20, dict(tolerance_px=0.5, min_pair_distance=50.0, top_k=0)),
])
def test_catalog_similarity_random(gpu, seed, nr, nm, kw):
    rng = np.random.default_rng(seed)
    ref = rng.uniform(0, 600, (nr, 2))
    mov = moved(ref, 1.5, 1.002, (6.0, -3.0), seed + 7, jitter=0.1, drop=min(4, nr - 4), extra=nm - nr + 4,
                span=600.0)[:nm]
    ref, mov = ref.astype(np.float32), mov.astype(np.float32)
    mov[1] = np.nan
    got, exp = both_catalog_fits(ref, mov, **kw)
    assert_catalog_equal(got, exp)


def test_catalog_similarity_multi_block_refit(gpu):
    """More than 256 moving stars: the refit's tree sums span two blocks."""
    rng = np.random.default_rng(71)
    ref = np.asarray([[100.0, 100.0], [400.0, 120.0], [260.0, 380.0], [90.0, 330.0]], np.float32)
    lin, t = similarity(0.5, 1.0, (4.0, -2.0))
    mov = rng.uniform(0, 500, (270, 2))
    mov[[3, 100, 200, 265]] = (ref - t) @ np.linalg.inv(lin).T
    mov = mov.astype(np.float32)
    got, exp = both_catalog_fits(ref, mov, tolerance_px=1.0, min_pair_distance=20.0, prior_dx=4.0, prior_dy=-2.0,
                                 prior_radius_px=3.0, top_k=5)
    assert_catalog_equal(got, exp)
    assert got["inliers"] >= 4


def test_catalog_similarity_failures_and_errors(gpu):
    x = np.asarray([0, 10, 5], np.float32)
    got, exp = both_catalog_fits(np.c_[x, x], np.c_[x, x], min_pair_distance=100.0, top_k=3)
    assert_catalog_equal(got, exp)
    assert got["status"] == "failed" and got["refit_status"] == "failed" and got["top_candidates"] == []
    with pytest.raises(ValueError, match="at least two stars"):
        G.estimate_similarity_from_catalogs_f32(x[:1], x[:1], x, x)
    with pytest.raises(ValueError, match="scale constraints"):
        G.estimate_similarity_from_catalogs_f32(x, x, x, x, min_scale=2.0, max_scale=1.0)
    with pytest.raises(ValueError, match="top_k"):
        G.estimate_similarity_from_catalogs_f32(x, x, x, x, top_k=-1)
    # the binding's positional conventions (glass_cuda.py:1753-1767)
    a = G.estimate_similarity_from_catalogs_f32(x, x + 1, x + 2, x + 1, 2.0, 2.0, 0.0, 0.0, -1.0, 0.0,
                                                float(np.finfo(np.float32).max), -1.0, 0)
    b = G.estimate_similarity_from_catalogs_f32(x, x + 1, x + 2, x + 1)
    assert a["matrix"] == b["matrix"] and a["best_candidate_index"] == b["best_candidate_index"]


def test_repeat_calls_identical(gpu):
    ref = field(81, 64)
    mov = moved(ref, 0.3, 1.0, (1.0, 2.0), 82)
    g1, _, _, _ = both_triangle_fits(ref, mov, 2.0, 0.1)
    g2, _, _, _ = both_triangle_fits(ref, mov, 2.0, 0.1)
    assert g1 == g2


def test_best_seed_parameters_on_repeated_calls(gpu):
    """Regression: more scoring blocks than reduction threads (2048 > 256) and
    a different append order on every call; the best seed's parameters must
    come with its index every time."""
    ref = field(13, 80)
    mov = moved(ref, 15.0, 1.03, (5.0, 5.0), 113)
    ref32, mov32 = ref.astype(np.float32), mov.astype(np.float32)
    rd, md = descriptors(ref32), descriptors(mov32)
    args = (ref32[:, 0], ref32[:, 1], mov32[:, 0], mov32[:, 1], rd["descriptors"], rd["indices"], md["descriptors"],
            md["indices"])
    exp = T.similarity_from_triangles(*args, 3.0, 0.1)
    assert exp["scored_candidates"] > 8192  # several candidates per scoring wave
    for _ in range(12):
        assert_fit_equal(G.estimate_similarity_from_triangle_descriptors_f32(*args, 3.0, 0.1), exp)


def test_resident_triangle_route_recovers_frames(gpu):
    """Catalogue -> descriptors -> similarity on resident frames (the resident
    default route, resident_cuda.py:10820-11420) on lights with known affine
    maps; every fit equals the oracle's on the same catalogues."""
    from glass_amd import ResidentCalibratedStack
    from glass_amd.synthetic import make_set

    data = make_set(seed=5, n_lights=6, height=512, width=640, n_cal=1, n_stars=220, fwhm=3.0, max_shift=20.0,
                    max_rot_deg=0.5, max_scale=5e-4, sky=0.02, noise=0.002)
    with ResidentCalibratedStack(6, 512, 640) as stack:
        for i in range(6):
            stack.upload_calibrated_frame(i, data.light[i])
        st = stack.frame_global_stats(0, exact_median=True)
        thr = float(st["median"] + 8.0 * 1.4826 * st["mad"])
        route = G.register_resident_frames(stack, 0, range(1, 6), thr, max_stars=64, tolerance_px=1.0)
    corners = np.array([[0, 0, 1], [639, 0, 1], [0, 511, 1], [639, 511, 1]], np.float64).T
    assert sorted(route["fits"]) == [1, 2, 3, 4, 5]
    rc = route["catalogs"][0]
    rd = T.triangle_descriptors(rc["x"], rc["y"], 64, 5, 1200)
    for i, fit in route["fits"].items():
        assert fit["status"] == "ok" and fit["inliers"] >= 20
        err = np.abs((np.asarray(fit["matrix"], np.float64) @ corners)[:2] - (data.matrices[i] @ corners)[:2]).max()
        assert err < 0.25, (i, err)
        mc = route["catalogs"][i]
        md = T.triangle_descriptors(mc["x"], mc["y"], 64, 5, 1200)
        assert_descriptors_equal(route["descriptors"][i] | {"max_stars": 64}, md | {"max_stars": 64})
        exp = T.similarity_from_triangles(rc["x"], rc["y"], mc["x"], mc["y"], rd["descriptors"], rd["indices"],
                                          md["descriptors"], md["indices"], 1.0, 0.1)
        assert_fit_equal(fit, exp)
