"""The workload configurations at their real sizes on the GPU, each against
the CPU oracle under the end-to-end tolerance contract (tests/contract.py):

* 32 lights of 2048 x 2048 from u16 raw with 16 + 16 + 16 calibration frames,
  400 stars, drift +-20 px / +-0.5 deg / scale 1 +- 5e-4, cosmic rays, two
  satellite trails, Lanczos3, variance-aware weights -- run end to end by the
  drop-in driver; masters compared in full, the maps on three row bands
  (both warp edges and the middle) against the oracle's calibrate -> FP64
  warp -> integrate of those rows;
* 200 lights of 9600 x 6422 generated in HBM, 20 + 20 + 20 calibration
  frames, 3000 stars, drift +-60 px / +-1 deg, 7 zero-weight frames -- the
  timed path of the benchmark; three blocks (top-left warp edge, centre,
  bottom-right edge) against the oracle on the blocks' source crops;
* the batched form at its frame sizes (1024 x 1024, 2048 x 1536,
  4096 x 3072), mixed interpolation, masked (NaN trail) and zero-weight
  frames -- one block per job against the oracle;
* the refined registration at 4096 x 4096: parameter gradient against FP64
  central differences (h = 1e-4 px) of the full-frame CPU warp, Gauss-Newton
  recovery of known maps, and the stack of the refined frames against the
  oracle's stack (same matrices: the contract; true matrices: reported RMS).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from blocks import is_identity, oracle_maps, warped_block
from contract import check_contract
from glass_amd import ResidentCalibratedStack
from glass_amd import pipeline as P
from glass_amd.synthetic import make_set, to_adu

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _compare_block(got, exp, samples, cov, weights, dq_got=None, dq_exp=None, all_valid=False):
    report = check_contract(got, exp, samples, cov > 0.5, weights, all_valid=all_valid)
    if dq_got is not None:
        differ = dq_got.astype(np.int64) != dq_exp.astype(np.int64)
        if report.count_mismatches:
            exempt = np.zeros(differ.shape, bool)
            for p in report.exempt_pixels:
                exempt[p] = True
            differ &= ~exempt
        assert not differ.any(), np.argwhere(differ)[:10]
    return report


# --------------------------------------------------------------- 32 x 2048^2
@pytest.fixture(scope="module")
def config1(gpu):
    _say("config 1: generating 32 x 2048 x 2048 lights + 16 x 3 calibration frames")
    data = make_set(seed=11, n_lights=32, height=2048, width=2048, n_cal=16, n_stars=400, fwhm=3.0,
                    max_shift=20.0, max_rot_deg=0.5, max_scale=5e-4, cosmic_density=1e-4, trails=2)
    adu = to_adu(data.light)
    bias, dark, flat = (to_adu(x).astype(np.float32) for x in (data.bias, data.dark, data.flat))
    pol = {"tile_size": 512}
    mb = P.build_master_bias(bias, pol)
    md = P.build_master_dark(dark, mb, pol)
    mf, _ = P.build_master_flat(flat, mb, pol)
    _say("config 1: drop-in driver run")
    res = P.run_resident_stack(adu, [data.light_exposure_s] * 32, {"bias": mb, "dark": md, "flat": mf},
                               data.matrices, dark_exposure_s=data.dark_exposure_s, interpolation="lanczos3",
                               weighting="variance_aware", policy=pol)
    cal = np.stack([oracle.calibrate_light(adu[i], mb, md, mf, data.light_exposure_s, data.dark_exposure_s)
                    for i in range(32)])
    return {"data": data, "bias": bias, "dark": dark, "flat": flat, "pol": pol, "masters": (mb, md, mf),
            "res": res, "cal": cal}


def test_config1_masters_full_size(config1):
    mb, md, mf = config1["masters"]
    pol = config1["pol"]
    assert np.array_equal(mb, oracle.build_master_bias(config1["bias"], pol))
    assert np.array_equal(md, oracle.build_master_dark(config1["dark"], mb, pol))
    assert np.array_equal(mf, oracle.build_master_flat(config1["flat"], mb, pol)[0])


def test_config1_weights(config1):
    rms = [oracle.frame_global_stats(c)["std"] for c in config1["cal"]]
    exp = oracle.quality_weights([{"background_rms": r} for r in rms], "variance_aware")
    assert np.allclose(config1["res"].weights, np.asarray(exp, np.float32), rtol=1e-6, atol=0)


@pytest.mark.parametrize("y0", [0, 1008, 2016])
def test_config1_band_end_to_end(config1, y0):
    data, res, cal = config1["data"], config1["res"], config1["cal"]
    block = (0, y0, 2048, 32)
    _say(f"config 1: oracle band rows {y0}..{y0 + 32}")

    def fetch(f, box):
        bx0, by0, bx1, by1 = box
        return cal[f][by0:by1, bx0:bx1]

    samples, cov = warped_block(fetch, data.matrices, block, (2048, 2048))
    exp, dq, geo = oracle_maps(samples, cov, res.weights)
    sl = (slice(y0, y0 + 32), slice(0, 2048))
    got = (res.master[sl], res.weight[sl], res.coverage[sl], res.low[sl], res.high[sl])
    assert np.array_equal(res.geometric[sl], geo)
    report = _compare_block(got, exp, samples, cov, res.weights, res.dq[sl], dq)
    _say(f"config 1 band {y0}: {report}")
    assert exp[3].sum() + exp[4].sum() > 0          # cosmic rays / trails are rejected
    if y0 in (0, 2016):
        assert np.any(cov < 0.5)                   # a warp-edge band


# ------------------------------------------------------------ 200 x 9600 x 6422
C2_H, C2_W, C2_F = 6422, 9600, 200
C2_BLOCKS = [(0, 0, 512, 32), (4544, 3195, 512, 32), (9088, 6390, 512, 32)]


@pytest.fixture(scope="module")
def config2(gpu):
    import torch

    from glass_amd.synthetic_gpu import make_calibration_raw, make_light_raw

    h, w, f = C2_H, C2_W, C2_F
    _say("config 2: generating 200 x 9600 x 6422 lights and 20 x 3 calibration frames in HBM")
    bias_raw = make_calibration_raw("bias", 20, h, w, 10)
    dark_raw = make_calibration_raw("dark", 20, h, w, 20, dark_exposure_s=60.0)
    flat_raw = make_calibration_raw("flat", 20, h, w, 30)
    blurred = list(range(1, 8))
    light_raw, mats = make_light_raw(f, h, w, 0, blurred=blurred, light_exposure_s=60.0)
    masters = P.build_masters_device(bias_raw, dark_raw, flat_raw, h, w)
    # bias / dark masters of the blocks' rows, for the oracle's own master build
    cal_rows = {}
    for (x0, y0, bw, bh) in C2_BLOCKS:
        cal_rows[(x0, y0)] = {
            k: np.stack([raw[:, (y0 + r) * w + x0:(y0 + r) * w + x0 + bw].cpu().numpy() for r in range(bh)], 1)
            for k, raw in (("bias", bias_raw), ("dark", dark_raw))}
    del bias_raw, dark_raw, flat_raw
    torch.cuda.empty_cache()
    weights = np.ones(f, np.float32)
    weights[blurred] = 0.0
    moving = [i for i in range(f) if not is_identity(mats[i])]
    _say("config 2: calibrate -> warp -> integrate")
    dev = torch.device("cuda:0")
    outs = {k: torch.empty((h, w), dtype=torch.float32, device=dev) for k in ("master", "weight", "coverage",
                                                                             "low", "high")}
    outs["dq"] = torch.empty((h, w), dtype=torch.int32, device=dev)
    with ResidentCalibratedStack(f, h, w) as s:
        s.set_calibration_masters(masters["bias"], masters["dark"], masters["flat"])
        s.calibrate_frames_device(list(range(f)), light_raw, "fits_u16be_bzero", [60.0] * f, 60.0,
                                  {"master_dark_includes_bias": True})
        s.reset_warp_coverage()
        for _ in range(f - len(moving)):
            s.accumulate_full_warp_coverage_frame()
        s.warp_frames(moving, mats[moving], "lanczos3", "cpu")
        s.integrate_to_device(outs, weights, "winsorized_sigma", 3.0, 3.0, 3, 0.5, "float32",
                              dq_geometric=True, dq_active_frames=int(np.count_nonzero(weights)))
        torch.cuda.synchronize()
        geometric = s.warp_coverage_map()
    maps = {k: v.cpu().numpy() for k, v in outs.items()}
    del outs
    host_masters = {k: masters[k].cpu().numpy() for k in ("bias", "dark", "flat")}
    del masters
    torch.cuda.empty_cache()
    return {"raw": light_raw, "mats": mats, "weights": weights, "maps": maps, "geometric": geometric,
            "masters": host_masters, "cal_rows": cal_rows}


def _decode_crop(raw, f, box):
    """Raw u16 FITS bytes of frame f's crop -> float32 ADU (oracle decode)."""
    bx0, by0, bx1, by1 = box
    rows = raw[f].view(C2_H, C2_W)[by0:by1, bx0:bx1].contiguous().cpu().numpy()
    return oracle.decode_fits_u16be_bzero(rows.view(np.uint8), rows.shape)


@pytest.mark.parametrize("block", C2_BLOCKS)
def test_config2_block_end_to_end(config2, block):
    c = config2
    mb, md, mf = c["masters"]["bias"], c["masters"]["dark"], c["masters"]["flat"]
    x0, y0, bw, bh = block
    _say(f"config 2: oracle block {block}")
    # the block's bias / dark masters by the oracle, from the raw frames' rows
    rows = c["cal_rows"][(x0, y0)]
    bias_rows = oracle.decode_fits_u16be_bzero(rows["bias"].view(np.uint8), rows["bias"].shape)
    dark_rows = oracle.decode_fits_u16be_bzero(rows["dark"].view(np.uint8), rows["dark"].shape)
    eb = oracle.build_master_bias(bias_rows, {"tile_size": 512})
    assert np.array_equal(eb, mb[y0:y0 + bh, x0:x0 + bw])
    assert np.array_equal(oracle.build_master_dark(dark_rows, eb, {"tile_size": 512}), md[y0:y0 + bh, x0:x0 + bw])

    def fetch(f, box):
        bx0, by0, bx1, by1 = box
        raw = _decode_crop(c["raw"], f, box)
        return oracle.calibrate_light(raw, mb[by0:by1, bx0:bx1], md[by0:by1, bx0:bx1], mf[by0:by1, bx0:bx1],
                                      60.0, 60.0)

    samples, cov = warped_block(fetch, c["mats"], block, (C2_H, C2_W))
    exp, dq, geo = oracle_maps(samples, cov, c["weights"])
    sl = (slice(y0, y0 + bh), slice(x0, x0 + bw))
    m = c["maps"]
    got = (m["master"][sl], m["weight"][sl], m["coverage"][sl], m["low"][sl], m["high"][sl])
    assert np.array_equal(c["geometric"][sl], geo)
    report = _compare_block(got, exp, samples, cov, c["weights"], m["dq"][sl].view(np.uint32), dq)
    _say(f"config 2 block {block}: {report}")
    if block[0] == 0 or block[0] + block[2] == C2_W:
        assert np.any(cov < 0.5)                   # warp-edge blocks


# ------------------------------------------------------------- batched jobs
C3_JOBS = [
    # (height, width, frames, interpolation, zero-weight frames, masked frame: trail row)
    (1024, 1024, 14, "bilinear", [3], {5: 500}),
    (1536, 2048, 20, "lanczos3", [2, 9], {4: 760}),
    (3072, 4096, 9, "lanczos3", [], {2: 1530}),
    (1024, 1024, 33, "lanczos3", [7, 8, 30], {}),
]


def _render_job(stack, spec, seed):
    """Star field of the job through each frame's map (GPU), NaN trail rows."""
    import torch

    from glass_amd.synthetic_gpu import render_stars_

    h, w, f, mats, masked = spec
    g = torch.Generator(device="cuda").manual_seed(seed)
    sx = torch.rand(400, generator=g, device="cuda") * w
    sy = torch.rand(400, generator=g, device="cuda") * h
    peaks = 10 ** (torch.rand(400, generator=g, device="cuda") * 1.5 - 1.0)
    sig = (torch.rand(400, generator=g, device="cuda") * 2.0 + 2.0) / 2.3548
    for k in range(f):
        inv = np.linalg.inv(mats[k])
        img = 0.02 + 0.01 * torch.randn((h, w), generator=g, device="cuda")
        render_stars_(img, (inv[0, 0] * sx + inv[0, 1] * sy + inv[0, 2]).float(),
                      (inv[1, 0] * sx + inv[1, 1] * sy + inv[1, 2]).float(), peaks, sig, 8)
        if k in masked:
            img[masked[k]:masked[k] + 3, :] = math.nan
        stack.upload_calibrated_frame(k, img)


def test_config3_batched_jobs_against_oracle(gpu):
    import torch

    from glass_amd.batch import StackingBatch

    rng = np.random.default_rng(33)
    jobs, blocks = [], []
    for j, (h, w, f, interp, zero, masked) in enumerate(C3_JOBS):
        mats = [np.eye(3)]
        for _ in range(1, f):
            a = math.radians(rng.uniform(-0.5, 0.5))
            mats.append(np.array([[math.cos(a), -math.sin(a), rng.uniform(-20, 20)],
                                  [math.sin(a), math.cos(a), rng.uniform(-20, 20)], [0.0, 0.0, 1.0]]))
        s = ResidentCalibratedStack(f, h, w)
        _render_job(s, (h, w, f, mats, masked), 900 + j)
        weights = np.ones(f, np.float32)
        weights[zero] = 0.0
        # a 32-row block over a masked trail (or the middle), full width
        y0 = min(max(next(iter(masked.values()), h // 2) - 8, 0), h - 32)
        block = (0, y0, w, 32)
        crops = {}
        for k in range(f):
            from oracle.warp import source_box

            box = (0, y0, w, y0 + 32) if k == 0 else source_box(mats[k], block, (h, w), interp)
            crops[k] = (box, s.download_frame_tile(k, *box))
        blocks.append((block, crops))
        out = {k: torch.zeros((h, w), dtype=torch.float32, device="cuda") for k in ("master", "weight", "coverage",
                                                                                   "low", "high")}
        out["dq"] = torch.zeros((h, w), dtype=torch.int32, device="cuda")
        jobs.append({"stack": s, "warp_indices": list(range(1, f)), "matrices": mats[1:], "interpolation": interp,
                     "semantics": "cpu", "fill": math.nan, "weights": weights, "outputs": out,
                     "count_map_dtype": "float32", "dq_geometric": True,
                     "dq_active_frames": int(np.count_nonzero(weights)), "all_mats": mats})
    _say("config 3: batched run")
    with StackingBatch(jobs, use_graph=True, autotune=True) as b:
        b.run()
        torch.cuda.synchronize()
    for j, (job, (block, crops)) in enumerate(zip(jobs, blocks)):
        x0, y0, w, bh = block
        h = job["stack"].height
        _say(f"config 3: job {j} ({h} x {w}, {job['interpolation']}) against the oracle")

        def fetch(f, box, crops=crops):
            (cx0, cy0, cx1, cy1), data = crops[f]
            bx0, by0, bx1, by1 = box
            return data[by0 - cy0:by1 - cy0, bx0 - cx0:bx1 - cx0]

        samples, cov = warped_block(fetch, job["all_mats"], block, (h, w), job["interpolation"])
        exp, dq, geo = oracle_maps(samples, cov, job["weights"])
        sl = (slice(y0, y0 + bh), slice(x0, x0 + w))
        o = {k: v.cpu().numpy() for k, v in job["outputs"].items()}
        got = (o["master"][sl], o["weight"][sl], o["coverage"][sl], o["low"][sl], o["high"][sl])
        report = _compare_block(got, exp, samples, cov, job["weights"], o["dq"][sl].view(np.uint32), dq)
        _say(f"config 3 job {j}: {report}")
        job["stack"].close()


# ------------------------------------------------- refined registration 4096^2
C4_SIZE = 4096


def _c4_render(m, size, noise_seed=None, seed=404):
    """Frame whose pixel p shows the scene at m @ p (m: frame -> reference),
    with Gaussian noise of sigma 0.01 when noise_seed is given."""
    import torch

    from glass_amd.synthetic_gpu import render_stars_

    g = torch.Generator(device="cuda").manual_seed(seed)
    n = 1500
    sx = torch.rand(n, generator=g, device="cuda") * size
    sy = torch.rand(n, generator=g, device="cuda") * size
    peaks = 10 ** (torch.rand(n, generator=g, device="cuda") * 2.0 - 0.5)
    sig = (torch.rand(n, generator=g, device="cuda") * 2.0 + 2.0) / 2.3548
    yy, xx = torch.meshgrid(torch.arange(size, device="cuda", dtype=torch.float64),
                            torch.arange(size, device="cuda", dtype=torch.float64), indexing="ij")
    u = m[0, 0] * xx + m[0, 1] * yy + m[0, 2]
    v = m[1, 0] * xx + m[1, 1] * yy + m[1, 2]
    img = (0.05 + 0.01 * torch.sin(u / 300.0) * torch.cos(v / 410.0)).float()
    del u, v, xx, yy
    inv = np.linalg.inv(m)
    render_stars_(img, (inv[0, 0] * sx + inv[0, 1] * sy + inv[0, 2]).float(),
                  (inv[1, 0] * sx + inv[1, 1] * sy + inv[1, 2]).float(), peaks.float(), sig.float(), 8)
    if noise_seed is not None:
        gn = torch.Generator(device="cuda").manual_seed(noise_seed)
        img += 0.01 * torch.randn((size, size), generator=gn, device="cuda")
    return img


def _c4_matrix(angle_deg, scale, tx, ty):
    a = math.radians(angle_deg)
    return np.array([[scale * math.cos(a), -scale * math.sin(a), tx],
                     [scale * math.sin(a), scale * math.cos(a), ty], [0.0, 0.0, 1.0]])


def test_config4_gradient_full_frame_finite_differences(gpu):
    import torch

    from glass_amd.refine import affine_lanczos3, source_map_of
    from oracle import cpu_ref

    size = C4_SIZE
    frame = _c4_render(_c4_matrix(0.31, 1.0003, 12.4, -7.7), size)
    b = source_map_of(_c4_matrix(0.3, 1.0002, 12.1, -7.5))
    rng = np.random.default_rng(41)
    g = rng.normal(size=(size, size))
    _say("config 4: analytic gradient at 4096 x 4096")
    bt = torch.from_numpy(b.copy()).requires_grad_(True)
    out = affine_lanczos3(frame, bt)
    (out.double() * torch.from_numpy(g).cuda()).sum().backward()
    analytic = bt.grad.numpy().copy()
    img = frame.cpu().numpy()
    del out, frame
    torch.cuda.empty_cache()
    _say("config 4: FP64 central differences of the full-frame CPU warp")
    scale = np.array([size, size, 1.0, size, size, 1.0])
    numeric = np.zeros(6)
    for k in range(6):
        step = 1e-4 / scale[k]   # source coordinates move by about 1e-4 px
        bp, bm = b.copy(), b.copy()
        bp[k] += step
        bm[k] -= step
        numeric[k] = (float(np.sum(g * cpu_ref.warp_map64(img, bp))) -
                      float(np.sum(g * cpu_ref.warp_map64(img, bm)))) / (2 * step)
    rel = np.abs(analytic - numeric) / np.maximum(np.abs(numeric), 1e-6 * np.max(np.abs(numeric)))
    _say(f"config 4: gradient relative error {rel}")
    assert np.all(rel < 2e-3), (analytic, numeric, rel)


def test_config4_refine_and_stack_against_oracle(gpu):
    import torch

    size = C4_SIZE
    truths = [np.eye(3), _c4_matrix(0.42, 1.0004, 15.3, -11.2), _c4_matrix(-0.37, 0.9996, -18.1, 6.4),
              _c4_matrix(0.11, 1.0001, 3.7, 17.9)]
    frames = [_c4_render(m, size, noise_seed=4200 + k) for k, m in enumerate(truths)]
    rng = np.random.default_rng(42)
    starts = [np.eye(3)]
    for m in truths[1:]:
        da = math.radians(rng.choice([-0.02, 0.02]))
        start = np.array([[math.cos(da), -math.sin(da), 0.0], [math.sin(da), math.cos(da), 0.0], [0, 0, 1.0]]) @ m
        start[0, 2] += rng.choice([-0.3, 0.3])
        start[1, 2] += rng.choice([-0.3, 0.3])
        starts.append(start)
    corners = np.array([[0, 0, 1], [size - 1, 0, 1], [0, size - 1, 1], [size - 1, size - 1, 1]], np.float64).T
    f = len(frames)
    outs = {k: torch.empty((size, size), dtype=torch.float32, device="cuda") for k in ("master", "weight",
                                                                                      "coverage", "low", "high")}
    with ResidentCalibratedStack(f, size, size) as s:
        for k, fr in enumerate(frames):
            s.upload_calibrated_frame(k, fr)
        host = [fr.cpu().numpy() for fr in frames]
        del frames
        refined = [np.eye(3)]
        for k in range(1, f):
            m, rep = s.refine_affine(0, k, starts[k], max_iterations=20, tolerance_px=1e-4)
            err = float(np.max(np.abs((m @ corners)[:2] - (truths[k] @ corners)[:2])))
            start_err = float(np.max(np.abs((starts[k] @ corners)[:2] - (truths[k] @ corners)[:2])))
            _say(f"config 4: frame {k} corner error {start_err:.3f} -> {err:.5f} px in {rep['iterations']} steps")
            assert err < 0.01, (err, rep)
            refined.append(m)
        s.reset_warp_coverage()
        s.accumulate_full_warp_coverage_frame()
        s.warp_frames(list(range(1, f)), np.stack(refined[1:]), "lanczos3", "cpu")
        s.integrate_to_device(outs, None, "winsorized_sigma", 3.0, 3.0, 3, 0.5, "float32", dq_geometric=False)
        torch.cuda.synchronize()
    got_maps = {k: v.cpu().numpy() for k, v in outs.items()}
    weights = np.ones(f, np.float32)
    for y0 in (0, 2032, 4064):
        block = (0, y0, size, 32)

        def fetch(k, box):
            bx0, by0, bx1, by1 = box
            return host[k][by0:by1, bx0:bx1]

        sl = (slice(y0, y0 + 32), slice(0, size))
        got = tuple(got_maps[k][sl] for k in ("master", "weight", "coverage", "low", "high"))
        samples, cov = warped_block(fetch, refined, block, (size, size))
        exp, _dq, _geo = oracle_maps(samples, cov, weights)
        report = _compare_block(got, exp, samples, cov, weights)
        true_samples, true_cov = warped_block(fetch, truths, block, (size, size))
        true_exp, _, _ = oracle_maps(true_samples, true_cov, weights)
        both = (exp[1] > 0) & (true_exp[1] > 0)
        d = got[0][both].astype(np.float64) - true_exp[0][both]
        rms_true = float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(true_exp[0][both].astype(np.float64) ** 2)))
        _say(f"config 4 band {y0}: refined-matrix stack vs oracle {report}; vs oracle with true matrices "
             f"rms_rel={rms_true:.3g}")
        assert rms_true < 5e-3
