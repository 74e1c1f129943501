"""Triangle-asterism similarity registration oracle (test infrastructure; see
oracle/__init__.py).

Restates, in float32 / float64 numpy with the operation order of the
reference's GPU route (no fused multiply-adds), the three stages of the
resident default registration ``similarity_cuda_triangle``:

* ``triangle_descriptors`` -- src/glass_cuda.py:1835-1901 (argument rules),
  cpp/src/native_bindings.cpp:17044-17253 (neighbour count, raw slot count,
  host de-duplication and area ordering) and
  cpp/cuda/registration_kernels.cu:755-927 (per-slot neighbour search,
  combination order, area / side-length gates, vertex order, descriptor).
* ``similarity_from_triangles`` -- src/glass_cuda.py:2050-2140,
  native_bindings.cpp:17597-17848 and registration_kernels.cu:1237-1436
  (descriptor-radius gate, two vertex orders per descriptor pair, FP64
  three-point similarity fit, mutual-nearest-neighbour scoring) with the
  best-candidate rule of registration_kernels.cu:1438-1528.
* ``similarity_from_catalogs`` -- src/glass_cuda.py:1726-1824,
  native_bindings.cpp:18254-18678 (argument rules, top-k candidate selection)
  and registration_kernels.cu:1060-1236, :1438-1821, :2159-2286 (ordered
  two-star pair seeds with scale / rotation / prior gates, the same scoring
  and best rule, the FP64 inlier refit with its 256-wide tree sums, refit
  acceptance).

One choice is this implementation's own: the reference keeps, among raw
descriptor slots naming the same three stars, whichever one its unstable
``std::sort`` leaves first; here the lowest slot (anchor-major, combination
order) is kept.  Such duplicates carry the same descriptor; their area can
differ in the last bit (another base vertex).

Pinning (tests/test_oracle_triangles.py): the reference tests' literal
catalogues and assertions (tests/test_gpu_registration_search.py:87-310,
:934-1120); golden vectors of the reference's FP64 CPU bridge
``_local_triangle_descriptors`` (src/glass/cpu/registration.py:236-303,
tests/golden/triangle_golden.npz, indices exact, descriptors and areas to
float32 rounding); the three-point and refit fits against an independent SVD
Procrustes fit (registration.py:140-162) and exact recovery of generated
transforms; brute-force pure-Python restatements of the neighbour search and
the mutual-match scoring on small catalogues; the top-k rule against its
declarative form on tie-free candidate lists.
"""

from __future__ import annotations

from itertools import combinations

import numpy as np

F32 = np.float32
F32_MAX = np.float32(np.finfo(np.float32).max)
REFIT_BLOCK = 256


def _f32(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1))


# --------------------------------------------------------------------------
# descriptors


def neighbor_setup(n: int, max_stars: int, neighbors: int):
    """(count, neighbour count, combinations per anchor, raw slot count):
    native_bindings.cpp:17062-17068."""
    count = min(int(n), int(max_stars))
    nb = min(16, max(3, min(count, int(neighbors))))
    combos = nb * (nb - 1) * (nb - 2) // 6
    raw = count * combos if count >= 3 else 0
    return count, nb, combos, raw


def nearest_neighbors(x, y, anchor: int, nb: int) -> np.ndarray:
    """Indices of the nb nearest finite stars of one anchor (itself included),
    float32 squared distances (candidate - anchor), ties to the lower index,
    -1 padding; registration_kernels.cu:815-847.  Distances at or above
    FLT_MAX never enter the list."""
    out = np.full(nb, -1, np.int64)
    ax, ay = x[anchor], y[anchor]
    if not (np.isfinite(ax) and np.isfinite(ay)):
        return out
    ok = np.isfinite(x) & np.isfinite(y)
    dx = x - ax
    dy = y - ay
    d2 = dx * dx + dy * dy
    cand = np.nonzero(ok & (d2 < F32_MAX))[0]
    order = cand[np.argsort(d2[cand], kind="stable")][:nb]
    out[: len(order)] = order
    return out


def _slot(x, y, i0, i1, i2):
    """One raw descriptor slot (registration_kernels.cu:853-926): None when a
    gate rejects it, else ((d0, d1), (first, second, third), area)."""
    if i0 < 0 or i1 < 0 or i2 < 0 or i0 == i1 or i1 == i2 or i0 == i2:
        return None
    x0, y0, x1, y1, x2, y2 = x[i0], y[i0], x[i1], y[i1], x[i2], y[i2]
    area = np.abs(((x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0)) * F32(0.5))
    if area < F32(2.0):
        return None

    def dist(l, r):
        dx = x[l] - x[r]
        dy = y[l] - y[r]
        return np.sqrt(dx * dx + dy * dy)

    d01, d12, d20 = dist(i0, i1), dist(i1, i2), dist(i2, i0)
    s, m, l = d01, d12, d20
    if s > m:
        s, m = m, s
    if m > l:
        m, l = l, m
    if s > m:
        s, m = m, s
    if s <= F32(1.0e-6) or m <= F32(1.0e-6):
        return None
    first, second, third, tf, ts = i0, i1, i2, d20, d12
    if d12 >= d01 and d12 >= d20:
        first, second, third, tf, ts = i1, i2, i0, d01, d20
    elif d20 >= d01 and d20 >= d12:
        first, second, third, tf, ts = i2, i0, i1, d12, d01
    if tf < ts:
        first, second = second, first
    return (l / m, m / s), (first, second, third), area


def raw_descriptor_slots(x, y, max_stars: int = 80, neighbors: int = 5):
    """Every raw slot in slot order (anchor-major, then lexicographic
    combinations of neighbour positions): list of None / slot tuples."""
    x, y = _f32(x), _f32(y)
    count, nb, _combos, raw = neighbor_setup(len(x), max_stars, neighbors)
    if raw == 0:
        return []
    x, y = x[:count], y[:count]
    combo_list = list(combinations(range(nb), 3))
    slots = []
    for anchor in range(count):
        nbrs = nearest_neighbors(x, y, anchor, nb)
        for ca, cb, cc in combo_list:
            slots.append(_slot(x, y, int(nbrs[ca]), int(nbrs[cb]), int(nbrs[cc])))
    return slots


def triangle_descriptors(x, y, max_stars: int = 80, neighbors: int = 5, max_descriptors: int = 1200) -> dict:
    """triangle_asterism_descriptors_f32 (glass_cuda.py:1835; bindings
    :17044-17253)."""
    if max_stars < 0:
        raise ValueError("max_stars must be non-negative")
    if max_descriptors < 0:
        raise ValueError("max_descriptors must be non-negative")
    x, y = _f32(x), _f32(y)
    if x.shape != y.shape:
        raise ValueError("catalog coordinate arrays must have the same shape")
    count, nb, _combos, raw = neighbor_setup(len(x), max_stars, neighbors)
    result = {"count": 0, "raw_count": int(raw), "max_stars": int(max_stars), "neighbors": int(nb),
              "descriptors": np.zeros((0, 2), np.float32), "indices": np.zeros((0, 3), np.int32),
              "areas": np.zeros(0, np.float32)}
    if count < 3 or max_descriptors == 0 or raw == 0:
        return result
    kept = {}
    for slot in raw_descriptor_slots(x, y, max_stars, neighbors):
        if slot is None:
            continue
        key = tuple(sorted(slot[1]))
        if key not in kept:
            kept[key] = slot
    ordered = sorted(kept.items(), key=lambda kv: (-float(kv[1][2]), kv[0]))[:max_descriptors]
    n = len(ordered)
    result["count"] = n
    result["descriptors"] = np.asarray([[s[0][0], s[0][1]] for _, s in ordered], np.float32).reshape(n, 2)
    result["indices"] = np.asarray([s[1] for _, s in ordered], np.int32).reshape(n, 3)
    result["areas"] = np.asarray([s[2] for _, s in ordered], np.float32).reshape(n)
    return result


# --------------------------------------------------------------------------
# scoring shared by both similarity estimators


def transformed(params, mx, my):
    """(K, M) float32 moving coordinates under K similarity parameter rows
    (a, b, tx, ty): x' = a*x - b*y + tx, y' = b*x + a*y + ty in that order."""
    a, b, tx, ty = (params[:, i : i + 1].astype(np.float32) for i in range(4))
    with np.errstate(invalid="ignore", over="ignore"):
        return (a * mx - b * my) + tx, (b * mx + a * my) + ty


def _last_min(d2, tol2, valid=None):
    """Per row: (last column attaining the minimum among valid d2 <= tol2,
    that minimum); -1 / tol2 when no column qualifies.  The '<=' running
    minimum of registration_kernels.cu:1384-1395."""
    q = d2 <= tol2
    if valid is not None:
        q &= valid
    best = np.where(q, d2, np.float32(np.inf)).min(axis=1)
    hit = q & (d2 == best[:, None])
    last = d2.shape[1] - 1 - np.argmax(hit[:, ::-1], axis=1)
    idx = np.where(hit.any(axis=1), last, -1)
    return idx, np.where(idx >= 0, best, tol2).astype(np.float32)


def mutual_scores(params, rx, ry, mx, my, tolerance_px: float, chunk: int = 4096):
    """Inliers and float32 RMS of each parameter row: a moving star counts
    when its nearest reference star within tolerance has it as its own nearest
    transformed moving star; the RMS sums the inliers' squared distances in
    moving-index order in float32 (registration_kernels.cu:1366-1435)."""
    params = np.asarray(params, np.float32).reshape(-1, 4)
    rx, ry, mx, my = _f32(rx), _f32(ry), _f32(mx), _f32(my)
    tol = F32(tolerance_px)
    tol2 = tol * tol
    rok = np.isfinite(rx) & np.isfinite(ry)
    mok = np.flatnonzero(np.isfinite(mx) & np.isfinite(my))
    rxv, ryv = rx[rok], ry[rok]
    rmap = np.flatnonzero(rok)
    inliers = np.zeros(len(params), np.int64)
    rms = np.full(len(params), np.nan, np.float32)
    for s in range(0, len(params), chunk):
        p = params[s : s + chunk]
        k = len(p)
        tx, ty = transformed(p, mx, my)
        txv, tyv = tx[:, mok], ty[:, mok]
        total = np.zeros(k, np.float32)
        count = np.zeros(k, np.int64)
        for pos, m in enumerate(mok):
            if len(rxv) == 0:
                break
            dx = tx[:, m : m + 1] - rxv[None, :]
            dy = ty[:, m : m + 1] - ryv[None, :]
            fidx, fbest = _last_min(dx * dx + dy * dy, tol2)
            matched = fidx >= 0
            bref = rmap[np.where(matched, fidx, 0)]
            ex = txv - rx[bref][:, None]
            ey = tyv - ry[bref][:, None]
            ridx, _ = _last_min(ex * ex + ey * ey, tol2)
            inl = matched & (ridx == pos)
            total = np.where(inl, total + fbest, total).astype(np.float32)
            count += inl
        inliers[s : s + k] = count
        with np.errstate(invalid="ignore", divide="ignore"):
            r = np.sqrt(total / count.astype(np.float32)).astype(np.float32)
        rms[s : s + k] = np.where(count > 0, r, np.float32(np.nan))
    return inliers, rms


def best_candidate(indices, inliers, rms):
    """Position of the best scored candidate: most inliers, then lowest RMS,
    then lowest candidate index, among rows with inliers >= 0 and finite RMS
    (registration_kernels.cu:1438-1528).  -1 when none qualifies."""
    indices = np.asarray(indices, np.int64)
    ok = (np.asarray(inliers) >= 0) & np.isfinite(rms)
    if not ok.any():
        return -1
    pos = np.flatnonzero(ok)
    order = np.lexsort((indices[pos], rms[pos], -np.asarray(inliers)[pos]))
    return int(pos[order[0]])


def _fit_result(a, b, tx, ty):
    a, b, tx, ty = F32(a), F32(b), F32(tx), F32(ty)
    matrix = np.asarray([[a, -b, tx], [b, a, ty], [0, 0, 1]], np.float32)
    return matrix, np.sqrt(a * a + b * b), np.arctan2(b, a)


# --------------------------------------------------------------------------
# similarity from triangle descriptors


def three_point_fit(mxyz, rxyz):
    """FP64 least-squares similarity (a, b, tx, ty) of three moving points onto
    three reference points, float32 outputs; registration_kernels.cu:1320-1355.
    Arrays of shape (K, 3) (float32 inputs).  Returns (params (K, 4) float32,
    ok (K,) bool)."""
    mx, my = (np.asarray(v, np.float32).astype(np.float64) for v in mxyz)
    rx, ry = (np.asarray(v, np.float32).astype(np.float64) for v in rxyz)
    n = 3.0
    smx = (mx[:, 0] + mx[:, 1]) + mx[:, 2]
    smy = (my[:, 0] + my[:, 1]) + my[:, 2]
    srx = (rx[:, 0] + rx[:, 1]) + rx[:, 2]
    sry = (ry[:, 0] + ry[:, 1]) + ry[:, 2]
    mmx, mmy, mrx, mry = smx / n, smy / n, srx / n, sry / n
    power = mx[:, 0] * mx[:, 0]
    same = mx[:, 0] * rx[:, 0]
    cross = mx[:, 0] * ry[:, 0]
    for j in range(3):
        if j:
            power = power + mx[:, j] * mx[:, j]
            same = same + mx[:, j] * rx[:, j]
            cross = cross + mx[:, j] * ry[:, j]
        power = power + my[:, j] * my[:, j]
        same = same + my[:, j] * ry[:, j]
        cross = cross - my[:, j] * rx[:, j]
    den = power - n * (mmx * mmx + mmy * mmy)
    na = same - n * (mmx * mrx + mmy * mry)
    nbv = cross - n * (mmx * mry - mmy * mrx)
    with np.errstate(all="ignore"):
        a = (na / den).astype(np.float32)
        b = (nbv / den).astype(np.float32)
        ad, bd = a.astype(np.float64), b.astype(np.float64)
        tx = (mrx - (ad * mmx - bd * mmy)).astype(np.float32)
        ty = (mry - (bd * mmx + ad * mmy)).astype(np.float32)
    ok = (den > 1.0e-20) & np.isfinite(a) & np.isfinite(b) & np.isfinite(tx) & np.isfinite(ty)
    return np.stack([a, b, tx, ty], axis=1), ok


def triangle_candidates(rx, ry, mx, my, rdesc, ridx, mdesc, midx, descriptor_radius: float):
    """Candidates that pass the gates, in candidate-index order: (indices
    (K,) int64 = (rd * Md + md) * 2 + orientation, params (K, 4) float32);
    registration_kernels.cu:1258-1362."""
    rdesc = np.asarray(rdesc, np.float32).reshape(-1, 2)
    mdesc = np.asarray(mdesc, np.float32).reshape(-1, 2)
    ridx = np.asarray(ridx, np.int64).reshape(-1, 3)
    midx = np.asarray(midx, np.int64).reshape(-1, 3)
    nr, nm = len(rx), len(mx)
    md_n = len(mdesc)
    r = F32(descriptor_radius)
    ddx = rdesc[:, None, 0] - mdesc[None, :, 0]
    ddy = rdesc[:, None, 1] - mdesc[None, :, 1]
    fin = (np.isfinite(rdesc).all(axis=1)[:, None] & np.isfinite(mdesc).all(axis=1)[None, :])
    with np.errstate(over="ignore", invalid="ignore"):
        near = fin & ~((ddx * ddx + ddy * ddy) > r * r)
    rd, md = np.nonzero(near)
    cand_idx, rows_r, rows_m = [], [], []
    for o in (0, 1):
        tri = ridx[rd].copy()
        if o:
            tri[:, [0, 1]] = tri[:, [1, 0]]
        cand_idx.append((rd * md_n + md) * 2 + o)
        rows_r.append(tri)
        rows_m.append(midx[md])
    idx = np.concatenate(cand_idx)
    rt = np.concatenate(rows_r)
    mt = np.concatenate(rows_m)
    inb = ((rt >= 0) & (rt < nr)).all(axis=1) & ((mt >= 0) & (mt < nm)).all(axis=1)
    idx, rt, mt = idx[inb], rt[inb], mt[inb]
    rxs, rys, mxs, mys = rx[rt], ry[rt], mx[mt], my[mt]
    fin = (np.isfinite(rxs).all(axis=1) & np.isfinite(rys).all(axis=1) & np.isfinite(mxs).all(axis=1)
           & np.isfinite(mys).all(axis=1))
    idx, rxs, rys, mxs, mys = idx[fin], rxs[fin], rys[fin], mxs[fin], mys[fin]
    params, ok = three_point_fit((mxs, mys), (rxs, rys))
    order = np.argsort(idx[ok], kind="stable")
    return idx[ok][order], params[ok][order]


def similarity_from_triangles(rx, ry, mx, my, rdesc, ridx, mdesc, midx, tolerance_px: float = 2.0,
                              descriptor_radius: float = 0.1) -> dict:
    """estimate_similarity_from_triangle_descriptors_f32 (glass_cuda.py:2050;
    bindings :17597-17848)."""
    if tolerance_px < 0.0:
        raise ValueError("tolerance_px must be non-negative")
    if descriptor_radius < 0.0:
        raise ValueError("descriptor_radius must be non-negative")
    rx, ry, mx, my = _f32(rx), _f32(ry), _f32(mx), _f32(my)
    rdesc = np.asarray(rdesc, np.float32).reshape(-1, 2)
    mdesc = np.asarray(mdesc, np.float32).reshape(-1, 2)
    ridx = np.asarray(ridx, np.int32).reshape(-1, 3)
    midx = np.asarray(midx, np.int32).reshape(-1, 3)
    if rx.shape != ry.shape or mx.shape != my.shape:
        raise ValueError("catalog coordinate arrays must have matching x/y shapes")
    if len(rdesc) != len(ridx) or len(mdesc) != len(midx):
        raise ValueError("descriptor and index row counts must match")
    if len(rx) < 3 or len(mx) < 3 or len(rdesc) == 0 or len(mdesc) == 0:
        raise ValueError("triangle descriptor similarity requires at least one descriptor and three stars")
    cand, params = triangle_candidates(rx, ry, mx, my, rdesc, ridx, mdesc, midx, descriptor_radius)
    inl, rms = mutual_scores(params, rx, ry, mx, my, tolerance_px)
    best = best_candidate(cand, inl, rms)
    out = {"candidate_count": int(len(rdesc) * len(mdesc) * 2), "reference_count": int(len(rx)),
           "moving_count": int(len(mx)), "reference_descriptor_count": int(len(rdesc)),
           "moving_descriptor_count": int(len(mdesc)), "tolerance_px": float(F32(tolerance_px)),
           "descriptor_radius": float(F32(descriptor_radius)), "scored_candidates": int(len(cand)),
           "candidate_indices": cand, "candidate_params": params, "candidate_inliers": inl, "candidate_rms": rms}
    if best < 0:
        out.update(matrix=np.eye(3, dtype=np.float32), scale=F32(1.0), rotation_rad=F32(0.0), rms_px=F32(np.nan),
                   inliers=0, best_candidate_index=-1, status="failed")
        return out
    matrix, scale, rot = _fit_result(*params[best])
    out.update(matrix=matrix, scale=scale, rotation_rad=rot, rms_px=rms[best], inliers=int(inl[best]),
               best_candidate_index=int(cand[best]), status="ok")
    return out


# --------------------------------------------------------------------------
# similarity from catalogues


def catalog_arguments(tolerance_px=2.0, min_pair_distance=2.0, prior_dx=None, prior_dy=None, prior_radius_px=None,
                      min_scale=None, max_scale=None, max_abs_rotation_rad=None, top_k=0) -> dict:
    """Python defaults (glass_cuda.py:1755-1767) and binding rules
    (native_bindings.cpp:18283-18302) as float32 values."""
    p = {
        "tolerance_px": F32(tolerance_px),
        "min_pair_distance": F32(min_pair_distance),
        "prior_dx": F32(0.0 if prior_dx is None else prior_dx),
        "prior_dy": F32(0.0 if prior_dy is None else prior_dy),
        "prior_radius_px": F32(-1.0 if prior_radius_px is None else prior_radius_px),
        "min_scale": F32(0.0 if min_scale is None else min_scale),
        "max_scale": F32(np.finfo(np.float32).max if max_scale is None else max_scale),
        "max_abs_rotation_rad": F32(-1.0 if max_abs_rotation_rad is None else max_abs_rotation_rad),
        "top_k": int(top_k),
    }
    if p["tolerance_px"] < 0:
        raise ValueError("tolerance_px must be non-negative")
    if p["min_pair_distance"] < 0:
        raise ValueError("min_pair_distance must be non-negative")
    if p["prior_radius_px"] < 0:
        p["prior_dx"] = F32(0.0)
        p["prior_dy"] = F32(0.0)
    if p["min_scale"] < 0 or p["max_scale"] < p["min_scale"]:
        raise ValueError("scale constraints must satisfy 0 <= min_scale <= max_scale")
    if p["max_abs_rotation_rad"] < 0:
        p["max_abs_rotation_rad"] = F32(-1.0)
    if p["top_k"] < 0:
        raise ValueError("top_k must be non-negative")
    return p


def ordered_pairs(n: int):
    """(first, second) of ordered pair index p in [0, n (n-1)):
    registration_kernels.cu:746-753."""
    p = np.arange(n * (n - 1), dtype=np.int64)
    first = p // (n - 1)
    off = p % (n - 1)
    return first, np.where(off >= first, off + 1, off)


def pair_candidates(rx, ry, mx, my, p: dict):
    """Pair-seed candidates passing the gates, in candidate-index order:
    (indices (K,) int64 = ref_pair * M (M - 1) + moving_pair, params (K, 4)
    float32); registration_kernels.cu:1060-1143."""
    nr, nm = len(rx), len(mx)
    ra, rb = ordered_pairs(nr)
    ma, mb = ordered_pairs(nm)
    npm = len(ma)
    mpd2 = p["min_pair_distance"] * p["min_pair_distance"]
    out_idx, out_par = [], []
    mvx, mvy = mx[mb] - mx[ma], my[mb] - my[ma]
    mfin = np.isfinite(mx[ma]) & np.isfinite(my[ma]) & np.isfinite(mx[mb]) & np.isfinite(my[mb])
    md2 = mvx * mvx + mvy * mvy
    mok = mfin & ~(md2 <= mpd2)
    ma_, mvx_, mvy_, md2_ = ma[mok], mvx[mok], mvy[mok], md2[mok]
    mpos = np.flatnonzero(mok)
    max_, may_ = mx[ma_], my[ma_]
    for rp in range(len(ra)):
        i, j = ra[rp], rb[rp]
        if not (np.isfinite(rx[i]) and np.isfinite(ry[i]) and np.isfinite(rx[j]) and np.isfinite(ry[j])):
            continue
        rvx, rvy = rx[j] - rx[i], ry[j] - ry[i]
        if rvx * rvx + rvy * rvy <= mpd2:
            continue
        with np.errstate(all="ignore"):
            a = (mvx_ * rvx + mvy_ * rvy) / md2_
            b = (mvx_ * rvy - mvy_ * rvx) / md2_
            tx = rx[i] - (a * max_ - b * may_)
            ty = ry[i] - (b * max_ + a * may_)
            scale = np.sqrt(a * a + b * b)
            rot = np.arctan2(b, a)
        ok = ~((scale < p["min_scale"]) | (scale > p["max_scale"]))
        if p["max_abs_rotation_rad"] >= 0:
            ok &= ~(np.abs(rot) > p["max_abs_rotation_rad"])
        if p["prior_radius_px"] >= 0:
            ex = tx - p["prior_dx"]
            ey = ty - p["prior_dy"]
            with np.errstate(all="ignore"):
                ok &= ~((ex * ex + ey * ey) > p["prior_radius_px"] * p["prior_radius_px"])
        if ok.any():
            out_idx.append(rp * npm + mpos[ok])
            out_par.append(np.stack([a[ok], b[ok], tx[ok], ty[ok]], axis=1).astype(np.float32))
    if not out_idx:
        return np.zeros(0, np.int64), np.zeros((0, 4), np.float32)
    return np.concatenate(out_idx), np.concatenate(out_par)


def _tree_sum(values: np.ndarray) -> np.ndarray:
    """Sum over axis 0 in the order of a 256-thread shared-memory tree per
    block of 256 consecutive entries (stride 128, 64, ..., 1), then the block
    partials added one after another from 0.0 (registration_kernels.cu:
    1641-1658, :1691-1702)."""
    n = values.shape[0]
    blocks = max(1, -(-n // REFIT_BLOCK))
    pad = np.zeros((blocks * REFIT_BLOCK,) + values.shape[1:], values.dtype)
    pad[:n] = values
    total = np.zeros(values.shape[1:], values.dtype)
    for blk in range(blocks):
        v = pad[blk * REFIT_BLOCK : (blk + 1) * REFIT_BLOCK].copy()
        stride = REFIT_BLOCK // 2
        while stride > 0:
            v[:stride] = v[:stride] + v[stride : 2 * stride]
            stride //= 2
        total = total + v[0]
    return total


def seed_matches(params, rx, ry, mx, my, tolerance_px):
    """Nearest reference star within tolerance of every moving star under one
    seed's parameters (one-sided, '<=' last minimum); -1 when none.  The
    per-star search of registration_kernels.cu:1598-1624."""
    p = np.asarray(params, np.float32).reshape(1, 4)
    tol = F32(tolerance_px)
    tol2 = tol * tol
    tx, ty = transformed(p, mx, my)
    rok = np.isfinite(rx) & np.isfinite(ry)
    dx = tx[0][:, None] - rx[None, :]
    dy = ty[0][:, None] - ry[None, :]
    idx, _ = _last_min(dx * dx + dy * dy, tol2, rok[None, :])
    idx[~(np.isfinite(mx) & np.isfinite(my))] = -1
    return idx


def refit(params, rx, ry, mx, my, tolerance_px):
    """The FP64 refit of the best seed over its one-sided matches
    (registration_kernels.cu:1530-1820): (matrix f32 (3,3), scale, rotation,
    refit_inliers, status, rms) with status 0 ok, 1 fewer than two matches,
    2 degenerate."""
    match = seed_matches(params, rx, ry, mx, my, tolerance_px)
    fin = np.isfinite(params).all()
    if not fin:
        match[:] = -1
    m = match >= 0
    mxd = np.where(m, mx.astype(np.float64), 0.0)
    myd = np.where(m, my.astype(np.float64), 0.0)
    rxd = np.where(m, rx[np.where(m, match, 0)].astype(np.float64), 0.0)
    ryd = np.where(m, ry[np.where(m, match, 0)].astype(np.float64), 0.0)
    cols = np.stack([mxd, myd, rxd, ryd, mxd * mxd + myd * myd, mxd * rxd + myd * ryd, mxd * ryd - myd * rxd],
                    axis=1)
    sums = _tree_sum(cols)
    count = int(m.sum())
    eye = np.eye(3, dtype=np.float32)
    if count < 2:
        return eye, F32(1.0), F32(0.0), count, 1, F32(np.nan)
    n = float(count)
    mmx, mmy, mrx, mry = sums[0] / n, sums[1] / n, sums[2] / n, sums[3] / n
    den = sums[4] - n * (mmx * mmx + mmy * mmy)
    if den <= 1.0e-20:
        return eye, F32(1.0), F32(0.0), count, 2, F32(np.nan)
    a = (sums[5] - n * (mmx * mrx + mmy * mry)) / den
    b = (sums[6] - n * (mmx * mry - mmy * mrx)) / den
    tx = mrx - (a * mmx - b * mmy)
    ty = mry - (b * mmx + a * mmy)
    matrix = np.asarray([[a, -b, tx], [b, a, ty], [0, 0, 1]], np.float64).astype(np.float32)
    scale = F32(np.sqrt(a * a + b * b))
    rot = F32(np.arctan2(b, a))
    md = matrix.astype(np.float64)
    ex = (md[0, 0] * mxd + md[0, 1] * myd + md[0, 2]) - rxd
    ey = (md[1, 0] * mxd + md[1, 1] * myd + md[1, 2]) - ryd
    res = _tree_sum(np.where(m, ex * ex + ey * ey, 0.0)[:, None])[0]
    rms = F32(np.sqrt(res / n))
    return matrix, scale, rot, count, 0, rms


def top_candidates(indices, inliers, rms, params, top_k: int, reference_count: int, moving_count: int):
    """Positions of the top-k candidate list (native_bindings.cpp:18424-18530):
    per inlier count the ceil(k/4) lowest-RMS candidates, from the highest
    count down, then the overall best k in arrival order, finally ordered by
    inliers desc, RMS asc, index asc.  ``indices`` ascending."""
    if top_k <= 0:
        return []
    per_limit = max(1, (top_k + 3) // 4)
    buckets_n = max(reference_count, moving_count) + 1
    buckets = [[] for _ in range(buckets_n)]
    overall = []

    def insert(selected, pos, limit):
        if len(selected) < limit:
            selected.append(pos)
            return
        worst = 0
        for q in range(1, len(selected)):
            cur, w = selected[q], selected[worst]
            if inliers[cur] < inliers[w] or (inliers[cur] == inliers[w] and rms[cur] > rms[w]):
                worst = q
        w = selected[worst]
        if inliers[pos] > inliers[w] or (inliers[pos] == inliers[w] and rms[pos] < rms[w]):
            selected[worst] = pos

    for pos in range(len(indices)):
        if inliers[pos] < 0 or not np.isfinite(rms[pos]):
            continue
        insert(overall, pos, top_k)
        if inliers[pos] < buckets_n:
            insert(buckets[inliers[pos]], pos, per_limit)
    chosen = []
    for score in range(buckets_n - 1, -1, -1):
        if len(chosen) >= top_k:
            break
        for pos in sorted(buckets[score], key=lambda q: (float(rms[q]), int(indices[q]))):
            if len(chosen) >= top_k:
                break
            if pos not in chosen:
                chosen.append(pos)
    for pos in overall:
        if len(chosen) >= top_k:
            break
        if pos not in chosen:
            chosen.append(pos)
    chosen.sort(key=lambda q: (-int(inliers[q]), float(rms[q]), int(indices[q])))
    return chosen


def similarity_from_catalogs(rx, ry, mx, my, tolerance_px=2.0, min_pair_distance=2.0, prior_dx=None, prior_dy=None,
                             prior_radius_px=None, min_scale=None, max_scale=None, max_abs_rotation_rad=None,
                             top_k=0) -> dict:
    """estimate_similarity_from_catalogs_f32 (glass_cuda.py:1726-1824;
    bindings :18254-18678)."""
    rx, ry, mx, my = _f32(rx), _f32(ry), _f32(mx), _f32(my)
    if rx.shape != ry.shape or mx.shape != my.shape:
        raise ValueError("catalog coordinate arrays must have matching x/y shapes")
    if len(rx) < 2 or len(mx) < 2:
        raise ValueError("similarity catalog estimation requires at least two stars per catalog")
    p = catalog_arguments(tolerance_px, min_pair_distance, prior_dx, prior_dy, prior_radius_px, min_scale,
                          max_scale, max_abs_rotation_rad, top_k)
    nr, nm = len(rx), len(mx)
    cand, params = pair_candidates(rx, ry, mx, my, p)
    inl, rms = mutual_scores(params, rx, ry, mx, my, p["tolerance_px"])
    best = best_candidate(cand, inl, rms)
    out = {"candidate_count": int(nr * (nr - 1) * nm * (nm - 1)), "reference_count": nr, "moving_count": nm,
           "scored_candidates": int(len(cand)), "candidate_indices": cand, "candidate_params": params,
           "candidate_inliers": inl, "candidate_rms": rms}
    out.update({k: (float(v) if k != "top_k" else v) for k, v in p.items()})
    if best < 0:
        matrix, scale, rot, rms_px, inliers, best_idx = np.eye(3, dtype=np.float32), F32(1.0), F32(0.0), F32(
            np.nan), 0, -1
        refit_inl, status, refit_rms = 0, 1, F32(np.nan)
    else:
        matrix, scale, rot = _fit_result(*params[best])
        rms_px, inliers, best_idx = rms[best], int(inl[best]), int(cand[best])
        r_matrix, r_scale, r_rot, refit_inl, status, refit_rms = refit(params[best], rx, ry, mx, my,
                                                                       p["tolerance_px"])
        if status == 0:
            if refit_inl < inliers or not np.isfinite(refit_rms) or not np.isfinite(rms_px) or refit_rms > rms_px:
                status = 3
            else:
                matrix, scale, rot, rms_px = r_matrix, r_scale, r_rot, refit_rms
    tops = []
    for pos in top_candidates(cand, inl, rms, params, p["top_k"], nr, nm):
        m_, s_, r_ = _fit_result(*params[pos])
        tops.append({"candidate_index": int(cand[pos]), "inliers": int(inl[pos]), "rms_px": rms[pos],
                     "matrix": m_, "scale": s_, "rotation_rad": r_})
    out.update(matrix=matrix, scale=scale, rotation_rad=rot, rms_px=rms_px, inliers=inliers,
               refined_inliers=int(refit_inl), refit_status_code=int(status),
               refit_status="ok" if status == 0 else ("rejected" if status == 3 else "failed"),
               refit_rms_px=refit_rms, best_candidate_index=best_idx, top_candidates=tops,
               status="ok" if inliers > 0 else "failed")
    return out
