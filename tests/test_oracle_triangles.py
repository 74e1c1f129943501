"""Pinning the triangle-asterism similarity oracle (oracle/triangles.py).

* descriptors: golden vectors of the reference's FP64 CPU bridge
  (tests/golden/triangle_golden.npz, written by make_triangle_golden.py), held
  to the reference's own check (tests/test_gpu_registration_search.py:87-108:
  indices equal, descriptors and areas close), plus a loop restatement of the
  kernel's insertion-based neighbour search on catalogues full of exact ties;
* fits: the three-point FP64 fit and the catalogue refit against the SVD
  Procrustes fit of src/glass/cpu/registration.py:140-162 (oracle
  fit_similarity, itself pinned bit-exact by registration_golden.npz) and
  exact recovery of generated transforms;
* scoring: a per-star loop restatement of the mutual-match rule on small
  catalogues with ties, and the reference tests' literal cases
  (tests/test_gpu_registration_search.py:157-210, :982-1120);
* reduction order and top-k: a hand-worked tree sum where order matters, and
  the top-k rule against its declarative form on tie-free lists.
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

from oracle import triangles as T
from oracle.registration import fit_similarity

GOLDEN = Path(__file__).parent / "golden" / "triangle_golden.npz"
LITERAL = np.asarray([(10.0, 10.0), (20.0, 11.0), (13.0, 24.0), (32.0, 31.0), (6.0, 38.0), (38.0, 7.0),
                      (27.0, 19.0), (16.0, 34.0)], np.float64)


def similarity(angle_deg, scale, t):
    a = np.deg2rad(angle_deg)
    lin = scale * np.asarray([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]], np.float64)
    return lin, np.asarray(t, np.float64)


def golden_cases():
    z = np.load(GOLDEN)
    return [(str(n), z) for n in z["names"]]


def assert_descriptors_match_golden(got, z, name):
    """The reference's check, with rows whose FP64 areas tie (equal to float32
    rounding) compared as sets: the float32 route may order them otherwise."""
    k = f"tri/{name}"
    xy = z[f"{k}/xy"]
    gi, gd, ga = z[f"{k}/indices"], z[f"{k}/descriptors"], z[f"{k}/areas"]
    assert got["count"] == len(gi)
    # float32 cancellation in the area of a thin triangle: a few ulp of the
    # coordinate products
    atol = 8 * np.finfo(np.float32).eps * float(np.max(np.abs(xy))) ** 2
    i = 0
    while i < len(gi):
        j = i + 1
        while j < len(gi) and abs(ga[j] - ga[i]) <= atol:
            j += 1
        if j - i == 1:
            assert np.array_equal(got["indices"][i], gi[i]), (name, i)
            np.testing.assert_allclose(got["descriptors"][i], gd[i], rtol=1e-5)
        else:
            want = {tuple(r) for r in gi[i:j]}
            have = {tuple(r) for r in got["indices"][i:j]}
            assert want == have, (name, i, j)
        np.testing.assert_allclose(got["areas"][i:j], ga[i:j], rtol=0, atol=atol)
        i = j


@pytest.mark.parametrize("name,z", golden_cases(), ids=[n for n, _ in golden_cases()])
def test_descriptors_match_reference_cpu_bridge(name, z):
    k = f"tri/{name}"
    xy = z[f"{k}/xy"]
    max_stars, neighbors, max_descriptors = (int(v) for v in z[f"{k}/args"])
    got = T.triangle_descriptors(xy[:, 0], xy[:, 1], max_stars, neighbors, max_descriptors)
    assert_descriptors_match_golden(got, z, name)


def test_literal_catalogue_counts():
    """tests/test_gpu_registration_search.py:87-108."""
    got = T.triangle_descriptors(LITERAL[:, 0], LITERAL[:, 1], max_stars=8, neighbors=5, max_descriptors=64)
    assert got["neighbors"] == 5 and got["raw_count"] == 80
    z = np.load(GOLDEN)
    assert got["count"] == len(z["tri/literal/indices"])


def test_descriptor_edge_rules():
    x = np.asarray([0.0, 10.0, 0.0], np.float32)
    y = np.asarray([0.0, 0.0, 10.0], np.float32)
    got = T.triangle_descriptors(x, y, max_stars=80, neighbors=5, max_descriptors=10)
    # neighbour count clamps to the catalogue, at least 3: one combination
    assert got["neighbors"] == 3 and got["raw_count"] == 3 and got["count"] == 1
    # right isosceles: longest side 0-1/0-2 ... hypotenuse 1-2, third vertex 0
    assert tuple(got["indices"][0])[2] == 0
    np.testing.assert_allclose(got["descriptors"][0], [np.sqrt(2.0), 1.0], rtol=1e-7)
    assert got["areas"][0] == 50.0
    assert T.triangle_descriptors(x[:2], y[:2])["count"] == 0
    assert T.triangle_descriptors(x, y, max_descriptors=0)["count"] == 0
    assert T.triangle_descriptors(x, y, max_stars=2)["raw_count"] == 0
    with pytest.raises(ValueError, match="max_stars"):
        T.triangle_descriptors(x, y, max_stars=-1)
    with pytest.raises(ValueError, match="max_descriptors"):
        T.triangle_descriptors(x, y, max_descriptors=-1)
    # area below 2 pixels and a non-finite anchor give no descriptors
    tiny = T.triangle_descriptors(np.asarray([0, 1, 0], np.float32), np.asarray([0, 0, 1], np.float32))
    assert tiny["count"] == 0
    nan = T.triangle_descriptors(np.asarray([0, 10, np.nan, 0], np.float32), np.asarray([0, 0, 5, 10], np.float32))
    assert nan["count"] == 1 and 2 not in nan["indices"][0]


def _insertion_neighbors(x, y, anchor, nb):
    """registration_kernels.cu:815-847 as a per-candidate insertion loop."""
    best_d = [float(np.finfo(np.float32).max)] * nb
    best_i = [-1] * nb
    if not (math.isfinite(x[anchor]) and math.isfinite(y[anchor])):
        return best_i
    for c in range(len(x)):
        if not (math.isfinite(x[c]) and math.isfinite(y[c])):
            continue
        dx = np.float32(x[c]) - np.float32(x[anchor])
        dy = np.float32(y[c]) - np.float32(y[anchor])
        d = float(dx * dx + dy * dy)
        at = next((p for p in range(nb) if d < best_d[p] or (d == best_d[p] and c < best_i[p])), -1)
        if at < 0:
            continue
        best_d[at + 1 :] = best_d[at:-1]
        best_i[at + 1 :] = best_i[at:-1]
        best_d[at], best_i[at] = d, c
    return best_i


def test_neighbor_search_ties_and_gaps():
    rng = np.random.default_rng(7)
    for trial in range(20):
        n = int(rng.integers(3, 25))
        x = rng.integers(0, 6, n).astype(np.float32)  # integer grid: many exact distance ties
        y = rng.integers(0, 6, n).astype(np.float32)
        if trial % 3 == 0:
            x[rng.integers(0, n)] = np.nan
        nb = int(rng.integers(3, 17))
        for anchor in range(n):
            want = _insertion_neighbors(x, y, anchor, nb)
            assert list(T.nearest_neighbors(x, y, anchor, nb)) == want


def _loop_scores(params, rx, ry, mx, my, tol):
    """registration_kernels.cu:1366-1435 as plain per-star loops (float32)."""
    f = np.float32
    a, b, tx, ty = (f(v) for v in params)
    tol2 = f(tol) * f(tol)
    inl, total = 0, f(0.0)
    fin = lambda u, v: math.isfinite(u) and math.isfinite(v)  # noqa: E731
    for m in range(len(mx)):
        if not fin(mx[m], my[m]):
            continue
        X = a * mx[m] - b * my[m] + tx
        Y = b * mx[m] + a * my[m] + ty
        bd, br = tol2, -1
        for r in range(len(rx)):
            if not fin(rx[r], ry[r]):
                continue
            dx, dy = X - rx[r], Y - ry[r]
            d = dx * dx + dy * dy
            if d <= bd:
                bd, br = d, r
        if br < 0:
            continue
        rd, rm = tol2, -1
        for o in range(len(mx)):
            if not fin(mx[o], my[o]):
                continue
            ox = a * mx[o] - b * my[o] + tx
            oy = b * mx[o] + a * my[o] + ty
            dx, dy = ox - rx[br], oy - ry[br]
            d = dx * dx + dy * dy
            if d <= rd:
                rd, rm = d, o
        if rm == m:
            inl += 1
            total = f(total + bd)
    return inl, (f(np.sqrt(total / f(inl))) if inl else f(np.nan))


def test_mutual_scores_match_loop_restatement():
    rng = np.random.default_rng(11)
    for trial in range(12):
        nr, nm = int(rng.integers(3, 14)), int(rng.integers(3, 14))
        rx = rng.integers(0, 8, nr).astype(np.float32)  # duplicates and exact distance ties
        ry = rng.integers(0, 8, nr).astype(np.float32)
        mx = rng.integers(0, 8, nm).astype(np.float32)
        my = rng.integers(0, 8, nm).astype(np.float32)
        if trial % 4 == 1:
            rx[0] = np.inf
            my[nm - 1] = np.nan
        params = np.asarray([[1, 0, 0, 0], [0, 1, 1, -1], [0.5, 0.25, 2, 1], [1, 0, 0.5, 0.5]], np.float32)
        inl, rms = T.mutual_scores(params, rx, ry, mx, my, 1.5)
        for k in range(len(params)):
            want_inl, want_rms = _loop_scores(params[k], rx, ry, mx, my, 1.5)
            assert inl[k] == want_inl, (trial, k)
            assert (np.isnan(rms[k]) and np.isnan(want_rms)) or rms[k] == want_rms, (trial, k)


def test_three_point_fit_matches_procrustes_and_recovers_transform():
    rng = np.random.default_rng(3)
    mx, my, rx, ry, want = [], [], [], [], []
    for _ in range(200):
        lin, t = similarity(rng.uniform(-30, 30), rng.uniform(0.8, 1.25), rng.uniform(-50, 50, 2))
        m = rng.uniform(0, 2000, (3, 2)).astype(np.float32).astype(np.float64)
        r = (m @ lin.T + t + rng.normal(0, 0.3, (3, 2))).astype(np.float32).astype(np.float64)
        mx.append(m[:, 0]), my.append(m[:, 1]), rx.append(r[:, 0]), ry.append(r[:, 1])
        want.append(fit_similarity(m, r))
    params, ok = T.three_point_fit((np.asarray(mx), np.asarray(my)), (np.asarray(rx), np.asarray(ry)))
    assert ok.all()
    for p, w in zip(params, want):
        np.testing.assert_allclose(p[:2], [w[0, 0], w[1, 0]], rtol=2e-6, atol=1e-7)
        np.testing.assert_allclose(p[2:], [w[0, 2], w[1, 2]], rtol=2e-6, atol=2e-4)
    # exact recovery without noise (dyadic transform, exact float32 points)
    m = np.asarray([[10.0, 20.0], [300.0, 40.0], [120.0, 500.0]])
    r = m @ np.asarray([[0.75, -0.5], [0.5, 0.75]]).T + [16.0, -8.0]
    p, ok = T.three_point_fit((m[None, :, 0], m[None, :, 1]), (r[None, :, 0], r[None, :, 1]))
    assert ok[0] and list(p[0]) == [0.75, 0.5, 16.0, -8.0]
    # collinear-identical moving points: degenerate
    p, ok = T.three_point_fit((np.ones((1, 3)), np.ones((1, 3))), (r[None, :, 0], r[None, :, 1]))
    assert not ok[0]


def _catalog_descriptors(xy):
    return T.triangle_descriptors(xy[:, 0], xy[:, 1], max_stars=8, neighbors=5, max_descriptors=64)


def test_reference_triangle_similarity_literal():
    """tests/test_gpu_registration_search.py:157-208."""
    lin, t = similarity(6.0, 1.015, (3.5, -2.25))
    ref = LITERAL
    mov = (ref - t) @ np.linalg.inv(lin).T
    rd, md = _catalog_descriptors(ref.astype(np.float32)), _catalog_descriptors(mov.astype(np.float32))
    got = T.similarity_from_triangles(ref[:, 0], ref[:, 1], mov[:, 0], mov[:, 1], rd["descriptors"], rd["indices"],
                                      md["descriptors"], md["indices"], tolerance_px=0.1, descriptor_radius=0.01)
    assert got["status"] == "ok"
    assert got["inliers"] == len(ref)
    assert got["candidate_count"] == rd["count"] * md["count"] * 2
    assert got["rms_px"] < 1.0e-3
    np.testing.assert_allclose(got["matrix"][:2, :2], lin, atol=1e-4)
    np.testing.assert_allclose(got["matrix"][:2, 2], t, atol=1e-4)
    # the best candidate's own parameters, and its place in the index order
    pos = int(np.flatnonzero(got["candidate_indices"] == got["best_candidate_index"])[0])
    assert got["candidate_inliers"][pos] == got["inliers"]
    others = np.flatnonzero(got["candidate_inliers"] == got["inliers"])
    assert np.all(got["candidate_rms"][others] >= got["rms_px"])


def test_triangle_similarity_errors_and_failure():
    x = np.asarray([0, 10, 0], np.float32)
    d = np.asarray([[1.2, 1.1]], np.float32)
    i = np.asarray([[0, 1, 2]], np.int32)
    with pytest.raises(ValueError, match="tolerance_px"):
        T.similarity_from_triangles(x, x, x, x, d, i, d, i, tolerance_px=-1)
    with pytest.raises(ValueError, match="descriptor_radius"):
        T.similarity_from_triangles(x, x, x, x, d, i, d, i, descriptor_radius=-1)
    with pytest.raises(ValueError, match="row counts"):
        T.similarity_from_triangles(x, x, x, x, d, i[:0], d, i)
    with pytest.raises(ValueError, match="at least one descriptor"):
        T.similarity_from_triangles(x[:2], x[:2], x, x, d, i, d, i)
    far = np.asarray([[9.0, 9.0]], np.float32)
    got = T.similarity_from_triangles(x, x + 1, x, x + 1, d, i, far, i)
    assert got["status"] == "failed" and got["best_candidate_index"] == -1 and np.isnan(got["rms_px"])
    assert np.array_equal(got["matrix"], np.eye(3, dtype=np.float32))


def _catalog_case(points, angle_deg, scale, t, jitter, extra_m, extra_r):
    lin, tt = similarity(angle_deg, scale, t)
    lin32 = lin.astype(np.float32)
    moving = np.asarray(points, np.float32)
    reference = moving @ lin32.T + np.asarray(t, np.float32) + jitter
    return (np.vstack([reference, np.asarray(extra_r, np.float32)]),
            np.vstack([moving, np.asarray(extra_m, np.float32)]), lin32, np.asarray(t, np.float32))


MOVING8 = [(10.0, 11.0), (25.0, 14.0), (18.0, 34.0), (41.0, 37.0), (53.0, 21.0), (63.0, 45.0), (31.0, 58.0),
           (72.0, 18.0)]


def test_reference_catalog_similarity_pair_candidates():
    """tests/test_gpu_registration_search.py:982-1048."""
    ref, mov, lin, t = _catalog_case(MOVING8, -4.0, 0.985, (-1.75, 4.5), 0.0, [[90.0, 90.0], [5.0, 70.0]],
                                     [[4.0, 4.0], [80.0, 8.0]])
    got = T.similarity_from_catalogs(ref[:, 0], ref[:, 1], mov[:, 0], mov[:, 1], tolerance_px=0.15,
                                     min_pair_distance=3.0, prior_dx=float(t[0]), prior_dy=float(t[1]),
                                     prior_radius_px=2.0, min_scale=0.95, max_scale=1.05, max_abs_rotation_rad=0.2,
                                     top_k=4)
    assert got["status"] == "ok"
    assert got["inliers"] >= 8 and got["refined_inliers"] >= 8
    assert got["refit_status"] in {"ok", "rejected"}
    assert got["top_k"] == 4 and len(got["top_candidates"]) == 4
    assert got["top_candidates"][0]["inliers"] >= got["top_candidates"][-1]["inliers"]
    assert got["candidate_count"] == len(ref) * (len(ref) - 1) * len(mov) * (len(mov) - 1)
    assert got["prior_radius_px"] == 2.0
    assert abs(got["scale"] - 0.985) < 1e-4
    assert abs(got["rotation_rad"] - np.deg2rad(-4.0)) < 1e-4
    assert got["rms_px"] < 1e-3
    np.testing.assert_allclose(got["matrix"][:2, :2], lin, atol=1e-4)
    np.testing.assert_allclose(got["matrix"][:2, 2], t, atol=1e-3)


def test_reference_catalog_similarity_refit():
    """tests/test_gpu_registration_search.py:1051-1118; the refit equals the
    SVD Procrustes fit over the seed's matches."""
    moving = MOVING8 + [(84.0, 52.0), (48.0, 74.0)]
    jitter = np.asarray([(0.04, -0.02), (-0.03, 0.02), (0.01, 0.03), (-0.02, -0.04), (0.03, 0.01), (-0.01, -0.03),
                         (0.02, 0.04), (-0.04, 0.00), (0.01, -0.01), (0.00, 0.02)], np.float32)
    ref, mov, lin, t = _catalog_case(moving, 2.5, 1.01, (3.5, -2.25), jitter, [[2.0, 88.0], [95.0, 9.0]],
                                     [[88.0, 4.0], [5.0, 92.0]])
    got = T.similarity_from_catalogs(ref[:, 0], ref[:, 1], mov[:, 0], mov[:, 1], tolerance_px=0.25,
                                     min_pair_distance=5.0, prior_dx=float(t[0]), prior_dy=float(t[1]),
                                     prior_radius_px=2.0, min_scale=0.98, max_scale=1.04, max_abs_rotation_rad=0.1)
    assert got["status"] == "ok" and got["refit_status"] == "ok"
    assert got["refined_inliers"] >= 10
    assert got["rms_px"] < 0.08
    np.testing.assert_allclose(got["matrix"][:2, :2], lin, atol=0.003)
    np.testing.assert_allclose(got["matrix"][:2, 2], t, atol=0.12)
    # independent fit over the same matches
    best = int(np.flatnonzero(got["candidate_indices"] == got["best_candidate_index"])[0])
    match = T.seed_matches(got["candidate_params"][best], ref[:, 0], ref[:, 1], mov[:, 0], mov[:, 1], 0.25)
    keep = match >= 0
    assert keep.sum() == got["refined_inliers"]
    m = np.stack([mov[keep, 0], mov[keep, 1]], 1).astype(np.float64)
    r = np.stack([ref[match[keep], 0], ref[match[keep], 1]], 1).astype(np.float64)
    want = fit_similarity(m, r)
    np.testing.assert_allclose(got["matrix"], want.astype(np.float32), rtol=1e-6, atol=1e-5)
    res = (m @ want[:2, :2].T + want[:2, 2]) - r
    assert abs(float(got["rms_px"]) - math.sqrt(float(np.mean(np.sum(res * res, 1))))) < 1e-6


def test_catalog_argument_rules():
    x = np.asarray([0, 10, 5], np.float32)
    with pytest.raises(ValueError, match="at least two stars"):
        T.similarity_from_catalogs(x[:1], x[:1], x, x)
    with pytest.raises(ValueError, match="tolerance_px"):
        T.similarity_from_catalogs(x, x, x, x, tolerance_px=-1)
    with pytest.raises(ValueError, match="min_pair_distance"):
        T.similarity_from_catalogs(x, x, x, x, min_pair_distance=-1)
    with pytest.raises(ValueError, match="scale constraints"):
        T.similarity_from_catalogs(x, x, x, x, min_scale=2.0, max_scale=1.0)
    with pytest.raises(ValueError, match="top_k"):
        T.similarity_from_catalogs(x, x, x, x, top_k=-1)
    got = T.similarity_from_catalogs(x, x + 3, x, x + 3, prior_dx=5.0, prior_radius_px=-1.0,
                                     max_abs_rotation_rad=-3.0)
    assert got["prior_dx"] == 0.0 and got["max_abs_rotation_rad"] == -1.0
    # every pair closer than min_pair_distance: no candidate, both fits fail
    got = T.similarity_from_catalogs(x, x, x, x, min_pair_distance=100.0)
    assert got["status"] == "failed" and got["refit_status"] == "failed" and got["refit_status_code"] == 1
    assert got["best_candidate_index"] == -1 and got["scored_candidates"] == 0


def test_tree_sum_order():
    v = np.zeros(256)
    v[0], v[1], v[129] = 1.0e16, 1.0, 1.0
    # the tree adds 1 + 1 before meeting 1e16 (exact 1e16 + 2); a running sum
    # would lose both ones to round-to-even
    assert T._tree_sum(v[:, None])[0] == 1.0e16 + 2.0
    assert (1.0e16 + 1.0) + 1.0 == 1.0e16
    ints = np.arange(700, dtype=np.float64)
    assert T._tree_sum(ints[:, None])[0] == float(ints.sum())


def _declarative_top(inl, rms, idx, k):
    order = sorted(range(len(idx)), key=lambda q: (-inl[q], rms[q], idx[q]))
    order = [q for q in order if inl[q] >= 0 and np.isfinite(rms[q])]
    per = max(1, (k + 3) // 4)
    chosen = []
    for s in sorted({int(inl[q]) for q in order}, reverse=True):
        for q in [q for q in order if inl[q] == s][:per]:
            if len(chosen) < k:
                chosen.append(q)
    for q in order:
        if len(chosen) < k and q not in chosen:
            chosen.append(q)
    return sorted(chosen, key=lambda q: (-inl[q], rms[q], idx[q]))


def test_top_candidates_declarative_form():
    rng = np.random.default_rng(5)
    for trial in range(40):
        n = int(rng.integers(1, 60))
        inl = rng.integers(0, 9, n)
        rms = rng.permutation(n).astype(np.float32) / 7 + np.float32(0.01)  # distinct: tie-free
        rms[inl == 0] = np.nan
        idx = np.sort(rng.choice(10_000, n, replace=False))
        k = int(rng.integers(1, 12))
        got = T.top_candidates(idx, inl, rms, None, k, 8, 8)
        assert got == _declarative_top(inl, rms, idx, k), trial
