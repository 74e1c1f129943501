#!/usr/bin/env python3
"""Resident stacking benchmark: calibrate -> Lanczos3 warp -> winsorized
reject + integrate (+ weight / coverage / low / high / DQ maps) on one MI355X.

Workload (default): 200 lights at 9600 x 6422 from FITS u16 (BZERO 32768,
big-endian) raw samples already resident in HBM (24.7 GB), 20 bias / 20 dark /
20 flat frames (masters built on the GPU before timing), affine drift up to
+-60 px and +-1 deg, 3000 stars with FWHM 2-4 px, 7 blurred frames given zero
weight, Lanczos3 warp, winsorized sigma clip 3/3 over the frame axis.

One step = one pass of the path over the whole light set: calibrate all
frames from raw (f32 stack, 49.3 GB), warp the 199 non-reference frames,
integrate the 193 positive-weight frames into the output maps.

Prints one JSON line (rank 0).  --gpus > 1 is refused: the path runs on one
GPU (it would shard by output row bands, each GPU receiving its band plus the
warp halo of every raw frame).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0


def parse_args():
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--frames", type=int, default=200)
    p.add_argument("--height", type=int, default=6422)
    p.add_argument("--width", type=int, default=9600)
    p.add_argument("--calib-frames", type=int, default=20)
    p.add_argument("--zero-weight", type=int, default=7)
    p.add_argument("--interpolation", default="lanczos3", choices=["lanczos3", "bilinear"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-rows", type=int, default=128, help="rows per frame of the CPU baseline sample")
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--no-refine", action="store_true", help="skip the refined-registration (config 4) leg")
    p.add_argument("--refine-frames", type=int, default=64)
    p.add_argument("--refine-size", type=int, default=4096)
    p.add_argument("--no-batch", action="store_true", help="skip the batched multi-job (config 3) leg")
    p.add_argument("--batch-jobs", type=int, default=64)
    p.add_argument("--batch-reps", type=int, default=3)
    p.add_argument("--no-local-norm", action="store_true", help="skip the local-normalisation leg")
    p.add_argument("--no-host-raw", action="store_true", help="skip the host-raw (PCIe-inclusive) leg")
    p.add_argument("--no-registration", action="store_true", help="skip the star-registration leg")
    p.add_argument("--fused-leg", action="store_true",
                   help="also time the fused matrix-warped integrations on the calibrated stack")
    p.add_argument("--local-norm-tile", type=int, default=512)
    p.add_argument("--copy-probe-mib", type=int, default=0,
                   help="issue one untimed device-to-device copy of this many MiB before the timed steps "
                        "(a known-size reference for the FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(args, light_exposure, dark_exposure):
    """Reference CPU path restated in C/OpenMP (oracle/cpu_ref.c), timed on the
    host cores over a bounded sample of the same workload: all frames, the
    first --cpu-rows rows, sub-pixel drift so the warp footprint covers the
    sample.  Returns the baseline dict (frame-Mpix/s)."""
    import numpy as np

    from glass_amd.synthetic import affine_matrix, encode_fits_u16be_bzero, make_set, to_adu
    from oracle import cpu_ref

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cpu_ref.load()

    def timed(h, w, n_threads):
        f = args.frames
        data = make_set(seed=args.seed, n_lights=f, height=h, width=w, n_cal=3,
                        n_stars=max(4, 3000 * h * w // (args.height * args.width)), fwhm=3.0, max_shift=2.0,
                        max_rot_deg=0.005,
                        sky=0.02)
        raw = encode_fits_u16be_bzero(to_adu(data.light)).reshape(f, -1)
        bias = to_adu(data.bias.mean(0)).astype(np.float32)
        dark = to_adu(data.dark.mean(0)).astype(np.float32)
        flat = np.maximum(data.flat.mean(0) / np.median(data.flat.mean(0)), 1e-6).astype(np.float32)
        weights = np.ones(f, np.float32)
        weights[1:1 + args.zero_weight] = 0.0
        invs = [np.linalg.inv(m) for m in data.matrices]
        cpu_ref.set_threads(n_threads)
        t0 = time.perf_counter()
        stack = cpu_ref.calibrate_u16be(raw, h * w, bias, dark, flat,
                                        float(np.float32(light_exposure / dark_exposure)))
        stack = stack.reshape(f, h, w)
        for i in range(1, f):
            stack[i], _ = cpu_ref.warp(stack[i], invs[i], args.interpolation, np.nan)
        cpu_ref.integrate_winsorized(stack, weights)
        elapsed = time.perf_counter() - t0
        return f * h * w / 1e6 / elapsed, elapsed

    rate, elapsed = timed(args.cpu_rows, args.width, threads)
    # one core on an eighth of the columns (the same rows and frames, so the
    # same per-pixel work and warp footprint; bounded time)
    cols1 = max(64, args.width // 8)
    rate1, elapsed1 = timed(args.cpu_rows, cols1, 1)
    cpu_ref.set_threads(threads)
    return {
        "value": round(rate, 3),
        "unit": "frame-Mpix/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{args.frames} frames x {args.cpu_rows} x {args.width} (first {args.cpu_rows} rows; sub-pixel "
                  f"drift), calibrate+{args.interpolation} warp+winsorized integrate in oracle/cpu_ref.c, "
                  f"{elapsed:.2f} s",
        "numpy_oracle": numpy_oracle_timing(args.seed),
        "one_core": {"value": round(rate1, 3), "unit": "frame-Mpix/s", "cores": 1,
                     "sample": f"{args.frames} frames x {args.cpu_rows} x {cols1} (first rows and columns), "
                               f"{elapsed1:.2f} s"},
    }


# GPU stack band of the refined-registration leg, checked against the CPU
# oracle inside the cpu_baseline leg (the only leg that runs the oracle)
_REFINE_STACK_CHECK: dict = {}


def numpy_oracle_timing(seed):
    """The numpy restatement of the CPU path (the oracle), single process,
    timed on config 0 whole (8 x 128^2, masters from 4+4+4 frames, bilinear,
    winsorized 3/3) and on a 64-row band of config 1's frame size (32 lights
    of 64 x 2048, Lanczos3, winsorized 3/3, masters given).  frame-Mpix/s of
    the calibrate + warp + integrate part."""
    import numpy as np

    import oracle
    from glass_amd.synthetic import make_set

    def run(data, interpolation, build_masters):
        t0 = time.perf_counter()
        if build_masters:
            mb = oracle.build_master_bias(data.bias, {"tile_size": 512})
            md = oracle.build_master_dark(data.dark, mb, {"tile_size": 512})
            mf, _ = oracle.build_master_flat(data.flat, mb, {"tile_size": 512})
        else:
            mb, md = data.bias.mean(0), data.dark.mean(0)
            mf = (data.flat.mean(0) / np.median(data.flat.mean(0))).astype(np.float32)
        t1 = time.perf_counter()
        cal = [oracle.calibrate_light(li, mb, md, mf, data.light_exposure_s, data.dark_exposure_s)
               for li in data.light]
        warped, covs = [], []
        for i, c in enumerate(cal):
            if np.array_equal(data.matrices[i], np.eye(3)):
                warped.append(c)
                covs.append(np.ones_like(c))
            else:
                w, cov = oracle.warp_matrix(c, data.matrices[i], fill=0.0, interpolation=interpolation)
                warped.append(w)
                covs.append(cov)
        oracle.weighted_integrate_stack(np.stack(warped), coverage=np.stack(covs), rejection="winsorized_sigma")
        t2 = time.perf_counter()
        fpx = data.light.shape[0] * data.light.shape[1] * data.light.shape[2] / 1e6
        return round(fpx / (t2 - t1), 4), round(t2 - t1, 3), round(t1 - t0, 3)

    c0 = make_set(seed=seed)
    r0, s0, m0 = run(c0, "bilinear", True)
    c1 = make_set(seed=seed + 1, n_lights=32, height=64, width=2048, n_cal=2, n_stars=13, fwhm=2.5,
                  max_shift=2.0, max_rot_deg=0.005, max_scale=5e-4)
    r1, s1, _ = run(c1, "lanczos3", False)
    return {"cores": 1, "config0": {"value": r0, "unit": "frame-Mpix/s", "path_s": s0, "masters_s": m0,
                                    "sample": "8 x 128 x 128, bilinear, masters from 4+4+4 frames"},
            "config1_band": {"value": r1, "unit": "frame-Mpix/s", "path_s": s1,
                             "sample": "32 x 64 x 2048 (a row band of config 1's frame), Lanczos3"}}


def refine_stack_vs_oracle():
    """RMS and P99 of the refined-leg stack band against the CPU oracle
    (oracle windowed Lanczos3 warp with the same refined matrices, oracle
    winsorized integration), relative to the oracle band's RMS / range."""
    import numpy as np

    import oracle
    from oracle.warp import source_box, warp_matrix_window

    c = _REFINE_STACK_CHECK
    size = c["size"]
    x0, y0, w, h = c["band"]
    samples = np.empty((len(c["matrices"]), h, w), np.float32)
    cov = np.empty_like(samples)
    for k, m in enumerate(c["matrices"]):
        crop = c["crops"][k]
        if np.array_equal(m, np.eye(3)):
            samples[k] = crop[y0 - c["crop_y0"]:y0 - c["crop_y0"] + h]
            cov[k] = 1.0
            continue
        bx0, by0, bx1, by1 = source_box(m, c["band"], (size, size))
        assert by0 >= c["crop_y0"] and by1 <= c["crop_y0"] + crop.shape[0], "band source outside the kept rows"
        samples[k], cov[k] = warp_matrix_window(crop[by0 - c["crop_y0"]:by1 - c["crop_y0"], bx0:bx1], m, c["band"],
                                                origin=(bx0, by0), full_shape=(size, size), fill=np.nan)
    weights = np.ones(len(samples), np.float32)
    exp = oracle.weighted_integrate_stack(samples, coverage=cov, weights=weights, rejection="winsorized_sigma",
                                          low_sigma=3.0, high_sigma=3.0, min_samples=3, max_reject_fraction=0.5)
    has = exp[1] > 0
    ref = exp[0][has].astype(np.float64)
    d = np.abs(c["master"][has].astype(np.float64) - ref)
    # Weight map (kept samples after rejection) against the oracle's; a pixel
    # may differ only where one of its samples lies within 1e-5 sigma of one
    # of the oracle's rejection thresholds (the tolerance contract).  The
    # scene is noise-free, so sigma is the frames' interpolation spread
    # (~1e-6 at values ~0.05, a few hundred float32 ulps) and 1e-5 sigma is
    # far below one ulp; the GPU's Lanczos3 samples themselves differ from
    # the oracle's FP64-weight samples by a few ulps.  So a sample within
    # 4 ulps of a threshold also counts as a threshold case here; the strict
    # 1e-5-sigma figure is reported beside it.
    from oracle.rejection import center_and_scale

    ok = (cov > 0) & np.isfinite(samples)
    center, scale = center_and_scale(samples, ok, method="winsorized_sigma", low_sigma=3.0, high_sigma=3.0)
    lo32 = center - np.float32(3.0) * scale
    hi32 = center + np.float32(3.0) * scale
    lo, hi = lo32.astype(np.float64), hi32.astype(np.float64)
    xd = samples.astype(np.float64)

    def near_within(tol_lo, tol_hi):
        return np.any(ok & ((np.abs(xd - lo) <= tol_lo) | (np.abs(xd - hi) <= tol_hi)), axis=0)

    tol = 1e-5 * np.abs(scale.astype(np.float64))
    near_strict = near_within(tol, tol)
    near = near_within(np.maximum(tol, 4.0 * np.spacing(np.abs(lo32)).astype(np.float64)),
                       np.maximum(tol, 4.0 * np.spacing(np.abs(hi32)).astype(np.float64)))
    differ = c["weight"] != exp[1]
    return {"band": f"rows {y0}..{y0 + h - 1} x {w}", "pixels": int(has.sum()),
            "rms_rel": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(ref * ref))),
            "p99_of_range": float(np.percentile(d, 99) / (ref.max() - ref.min())),
            "sigma_median": float(np.median(scale)),
            "weight_map_mismatches": int(differ.sum()),
            "weight_map_mismatches_off_1e-5_sigma": int((differ & ~near_strict).sum()),
            "weight_map_mismatches_off_threshold": int((differ & ~near).sum())}


def refine_leg(args):
    """Refined-registration form (config 4): 64 frames of 4096x4096 rendered
    from one analytic scene through known affine maps (rotation U(+-0.5 deg),
    shift U(+-20 px), scale 1 +- 5e-4); each refinement starts 0.3 px / 0.02
    deg away and runs Gauss-Newton on the photometric residual through the
    Lanczos3 adjoint.  Reports throughput (frame-Mpix of refined frames per s)
    and the corner error against the known maps."""
    import math

    import numpy as np
    import torch

    import glass_amd
    from glass_amd.synthetic_gpu import render_stars_

    n, size = args.refine_frames, args.refine_size
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(args.seed + 400)
    rng = np.random.default_rng(args.seed + 401)
    n_stars = 1500
    sx = torch.rand(n_stars, generator=g, device=dev) * size
    sy = torch.rand(n_stars, generator=g, device=dev) * size
    peaks = 10 ** (torch.rand(n_stars, generator=g, device=dev) * 2.0 - 0.5)
    sig = (torch.rand(n_stars, generator=g, device=dev) * 2.0 + 2.0) / 2.3548  # FWHM 2-4 px
    yy, xx = torch.meshgrid(torch.arange(size, device=dev, dtype=torch.float64),
                            torch.arange(size, device=dev, dtype=torch.float64), indexing="ij")

    def render(m):
        # frame pixel p shows the scene at m @ p (m: frame -> scene/reference)
        u = m[0, 0] * xx + m[0, 1] * yy + m[0, 2]
        v = m[1, 0] * xx + m[1, 1] * yy + m[1, 2]
        img = (0.05 + 0.01 * torch.sin(u / 300.0) * torch.cos(v / 410.0)).float()
        inv = np.linalg.inv(m)
        px = (inv[0, 0] * sx + inv[0, 1] * sy + inv[0, 2]).float()
        py = (inv[1, 0] * sx + inv[1, 1] * sy + inv[1, 2]).float()
        render_stars_(img, px, py, peaks.float(), sig.float(), 8)
        return img

    stack = glass_amd.ResidentCalibratedStack(n + 1, size, size)
    stack.upload_calibrated_frame(0, render(np.eye(3)))
    truths, starts = [], []
    for k in range(n):
        a = math.radians(rng.uniform(-0.5, 0.5))
        sc = 1.0 + rng.uniform(-5e-4, 5e-4)
        m = np.array([[sc * math.cos(a), -sc * math.sin(a), rng.uniform(-20, 20)],
                      [sc * math.sin(a), sc * math.cos(a), rng.uniform(-20, 20)], [0.0, 0.0, 1.0]])
        stack.upload_calibrated_frame(k + 1, render(m))
        da = math.radians(rng.choice([-0.02, 0.02]))
        rot = np.array([[math.cos(da), -math.sin(da), 0.0], [math.sin(da), math.cos(da), 0.0], [0.0, 0.0, 1.0]])
        start = rot @ m
        start[0, 2] += rng.choice([-0.3, 0.3])
        start[1, 2] += rng.choice([-0.3, 0.3])
        truths.append(m)
        starts.append(start)
    torch.cuda.synchronize()
    del xx, yy
    corners = np.array([[0, 0, 1], [size - 1, 0, 1], [0, size - 1, 1], [size - 1, size - 1, 1]], dtype=np.float64).T
    # one untimed refinement (workspace allocation)
    stack.refine_affine(0, 1, starts[0], max_iterations=2, tolerance_px=1e-3)
    t0 = time.perf_counter()
    errs, iters, refined = [], [], [np.eye(3)]
    for k in range(n):
        m, rep = stack.refine_affine(0, k + 1, starts[k], max_iterations=20, tolerance_px=1e-3)
        refined.append(m)
        iters.append(rep["iterations"])
        errs.append(float(np.max(np.abs((m @ corners)[:2] - (truths[k] @ corners)[:2]))))
    elapsed = time.perf_counter() - t0
    start_err = max(float(np.max(np.abs((starts[k] @ corners)[:2] - (truths[k] @ corners)[:2]))) for k in range(n))
    # the stack of the refined frames (Lanczos3, CPU-path semantics,
    # winsorized 3/3, unit weights) on the GPU; a row band of it and the
    # source rows that band reads are kept for the CPU-oracle check of the
    # cpu_baseline leg
    y0, rows, margin = size // 2 - 16, 32, 64
    cy0, cy1 = max(0, y0 - margin), min(size, y0 + rows + margin)
    crops = np.stack([stack.download_frame_tile(k, 0, cy0, size, cy1) for k in range(n + 1)])
    stack.warp_frames(list(range(1, n + 1)), np.stack(refined[1:]), "lanczos3", "cpu")
    outs = {k: torch.empty((size, size), dtype=torch.float32, device=dev) for k in ("master", "weight", "coverage",
                                                                                     "low", "high")}
    stack.integrate_to_device(outs, None, "winsorized_sigma", 3.0, 3.0, 3, 0.5, "float32", dq_geometric=False)
    torch.cuda.synchronize()
    _REFINE_STACK_CHECK.update(size=size, band=(0, y0, size, rows), crop_y0=cy0, crops=crops, matrices=refined,
                               master=outs["master"][y0:y0 + rows].cpu().numpy(),
                               weight=outs["weight"][y0:y0 + rows].cpu().numpy())
    del outs
    stack.close()
    torch.cuda.empty_cache()
    return {
        "workload": f"{n} frames x {size}x{size}, Gauss-Newton affine refinement (Lanczos3 adjoint)",
        "value": round(n * size * size / 1e6 / elapsed, 1),
        "unit": "frame-Mpix/s",
        "ms_per_frame": round(elapsed * 1e3 / n, 3),
        "mean_iterations": round(float(np.mean(iters)), 2),
        "start_corner_err_px_max": round(start_err, 4),
        "corner_err_px_max": round(max(errs), 5),
        "corner_err_px_median": round(float(np.median(errs)), 5),
    }


def batched_leg(args):
    """Batched multi-job form (config 3): `--batch-jobs` resident jobs with
    F ~ U{8..120} frames of 1024x1024, 2048x1536 or 4096x3072, each frame
    rendered from a per-job star field through its own affine map (drift
    +-20 px, rotation +-0.5 deg), 0-10% of frames given zero weight and 0-10%
    carrying a masked (NaN) satellite trail, bilinear or Lanczos3 per job,
    winsorized sigma 3/3 with every output map.  The same jobs run as one
    StackingBatch (autotuned, captured into a hipGraph, replayed) and as serial
    per-job warp_frames + integrate_to_device calls; the maps of both forms are
    compared bit for bit and both throughputs are reported (frame-Mpix of
    input frames per second)."""
    import math

    import numpy as np
    import torch

    import glass_amd
    from glass_amd.batch import StackingBatch
    from glass_amd.synthetic_gpu import render_stars_

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(args.seed + 500)
    sizes = [(1024, 1024), (1536, 2048), (3072, 4096)]  # (height, width)

    def out_maps(h, w):
        o = {k: torch.empty((h, w), dtype=torch.float32, device=dev) for k in ("master", "weight")}
        for k in ("coverage", "low", "high"):
            o[k] = torch.empty((h, w), dtype=torch.int16, device=dev)
        o["dq"] = torch.empty((h, w), dtype=torch.int32, device=dev)
        return o

    def render_job(spec, stack):
        # frames are a function of the job's seed: re-rendered, they restore
        # the stack's inputs exactly
        h, w, f = spec["h"], spec["w"], spec["f"]
        g = torch.Generator(device=dev).manual_seed(spec["seed"])
        n_stars = 400
        sx = torch.rand(n_stars, generator=g, device=dev) * w
        sy = torch.rand(n_stars, generator=g, device=dev) * h
        peaks = 10 ** (torch.rand(n_stars, generator=g, device=dev) * 1.5 - 1.0)
        sig = (torch.rand(n_stars, generator=g, device=dev) * 2.0 + 2.0) / 2.3548
        for k in range(f):
            inv = np.linalg.inv(spec["mats"][k])
            img = 0.02 + 0.01 * torch.randn((h, w), generator=g, device=dev)
            render_stars_(img, (inv[0, 0] * sx + inv[0, 1] * sy + inv[0, 2]).float(),
                          (inv[1, 0] * sx + inv[1, 1] * sy + inv[1, 2]).float(), peaks, sig, 8)
            row = spec["masked"].get(k)
            if row is not None:
                img[row:row + 3, :] = math.nan
            stack.upload_calibrated_frame(k, img)

    jobs, specs, frame_px, n_zero_total, n_masked_total = [], [], 0, 0, 0
    for j in range(args.batch_jobs):
        h, w = sizes[int(rng.integers(3))]
        f = int(rng.integers(8, 121))
        interp = "lanczos3" if rng.random() < 0.5 else "bilinear"
        mats = [np.eye(3)]
        for _ in range(1, f):
            a = math.radians(rng.uniform(-0.5, 0.5))
            mats.append(np.array([[math.cos(a), -math.sin(a), rng.uniform(-20, 20)],
                                  [math.sin(a), math.cos(a), rng.uniform(-20, 20)], [0.0, 0.0, 1.0]]))
        masked_frames = rng.choice(np.arange(1, f), size=int(rng.integers(0, f // 10 + 1)), replace=False)
        spec = {"h": h, "w": w, "f": f, "mats": mats, "seed": args.seed + 1000 + j,
                "masked": {int(k): int(rng.integers(0, h - 3)) for k in masked_frames}}
        stack = glass_amd.ResidentCalibratedStack(f, h, w)
        render_job(spec, stack)
        weights = np.ones(f, np.float32)
        zero = rng.choice(np.arange(1, f), size=int(rng.integers(0, f // 10 + 1)), replace=False)
        weights[zero] = 0.0
        n_zero_total += zero.size
        n_masked_total += len(spec["masked"])
        frame_px += f * h * w
        specs.append(spec)
        jobs.append({"stack": stack, "warp_indices": list(range(1, f)), "matrices": mats[1:],
                     "interpolation": interp, "semantics": "cpu", "fill": math.nan, "weights": weights,
                     "outputs": out_maps(h, w),
                     "dq_geometric": True, "dq_active_frames": int(np.count_nonzero(weights))})
    torch.cuda.synchronize()
    log(f"batched leg: {len(jobs)} jobs, {frame_px / 1e9:.2f} G frame-px generated")

    stream = torch.cuda.Stream(device=dev)
    sh = stream.cuda_stream
    batch = StackingBatch(jobs, use_graph=True, autotune=True)
    t0 = time.perf_counter()
    batch.run(sh)
    torch.cuda.synchronize()
    first_s = time.perf_counter() - t0
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    reps = max(args.batch_reps, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for _ in range(reps):
        batch.run(sh)
    ev[1].record(stream)
    torch.cuda.synchronize()
    batched_ms = (time.perf_counter() - t0) * 1e3 / reps
    batched_gpu_ms = ev[0].elapsed_time(ev[1]) / reps
    info = batch.info()
    log(f"batched: {batched_ms:.2f} ms per run (first run with autotune + capture {first_s:.2f} s), {info}")

    def serial(outs):
        for job, o in zip(jobs, outs):
            s = job["stack"]
            s.reset_warp_coverage()
            s.accumulate_full_warp_coverage_frame()
            s.warp_frames(job["warp_indices"], job["matrices"], job["interpolation"], "cpu", fill=job["fill"],
                          stream=sh)
            s.integrate_to_device(o, weights=job["weights"], dq_geometric=True,
                                  dq_active_frames=job["dq_active_frames"], stream=sh)

    # The serial form warps each stack in place (slot swap).  Its first run
    # reads the frames the batch read (the batch leaves the stacks untouched)
    # and is compared with it bit for bit; before the timed run the stacks are
    # re-rendered (same scene and maps; the star scatter-add is not
    # bit-reproducible, so that run is timed, not compared).
    def restore():
        for spec, job in zip(specs, jobs):
            render_job(spec, job["stack"])
        torch.cuda.synchronize()

    serial_outs = [out_maps(j["stack"].height, j["stack"].width) for j in jobs]
    serial(serial_outs)
    torch.cuda.synchronize()
    mismatched = []
    for k, (job, o) in enumerate(zip(jobs, serial_outs)):
        for name, t in job["outputs"].items():
            u = o[name]
            same = (t == u) | (t.isnan() & u.isnan()) if t.is_floating_point() else (t == u)
            if not bool(same.all()):
                mismatched.append((k, name))
    restore()
    t0 = time.perf_counter()
    ev[0].record(stream)
    serial(serial_outs)
    ev[1].record(stream)
    torch.cuda.synchronize()
    serial_ms = (time.perf_counter() - t0) * 1e3
    serial_gpu_ms = ev[0].elapsed_time(ev[1])
    log(f"serial: {serial_ms:.2f} ms per run; mismatched maps vs batched: {len(mismatched)}")
    batch.close()
    for job in jobs:
        job["stack"].close()
    del jobs, serial_outs
    torch.cuda.empty_cache()
    return {
        "workload": f"{args.batch_jobs} jobs, F~U{{8..120}}, sizes 1024x1024/2048x1536/4096x3072, "
                    "mixed bilinear/Lanczos3, 0-10% zero-weight + 0-10% masked frames, winsorized 3/3 + maps",
        "value": round(frame_px / 1e6 / (batched_ms * 1e-3), 1),
        "unit": "frame-Mpix/s",
        "ms_per_run": round(batched_ms, 3),
        "gpu_ms_per_run": round(batched_gpu_ms, 3),
        "first_run_s": round(first_s, 3),
        "serial_value": round(frame_px / 1e6 / (serial_ms * 1e-3), 1),
        "serial_ms_per_run": round(serial_ms, 3),
        "serial_gpu_ms_per_run": round(serial_gpu_ms, 3),
        "speedup_vs_serial": round(serial_ms / batched_ms, 3),
        "frame_gpx": round(frame_px / 1e9, 3),
        "zero_weight_frames": int(n_zero_total),
        "masked_frames": int(n_masked_total),
        "maps_bit_exact_vs_serial": not mismatched,
        "schedule": info,
    }


# Measured HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, rocprofv3 --pmc, each
# scaled by a known-size copy) of the default workload's kernels.
_TRAFFIC_KERNELS = {"calibrate": "gh::calibrate_batch_kernel<2, false, 16>", "warp": "gh::warp_lanczos3_stream_kernel",
                    "integrate": "gh::integrate_winsorized_kernel<200, true>"}
# VALU issue ceiling without packed FP32: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz
VALU_PEAK_LANE_INSTS = 256 * 4 * 16 * 2.4e9


def registration_leg(args, stack, light_raw, a, frames=9):
    """Star detection (threshold 5 sigma, 500 stars) on the calibrated, not
    yet warped first `frames` lights, then the affine star transform of each
    moving frame against frame 0, compared with the generator's matrices
    (largest corner displacement over the frame).  Wall clock per frame for
    detection (on the resident frame, synchronous) and for the transform."""
    import numpy as np
    import torch

    from glass_amd.registration import estimate_star_transform

    f = min(frames, stack.frame_count)
    idx = list(range(f))
    stack.calibrate_frames_device(idx, light_raw, "fits_u16be_bzero", [a["light_exp"]] * f, a["dark_exp"],
                                  a["policy"])
    torch.cuda.synchronize()
    stack.detect_stars(0)  # first call sizes the candidate buffers
    t = time.perf_counter()
    stars = [stack.detect_stars(i, 5.0, 500) for i in idx]
    detect_s = (time.perf_counter() - t) / f
    h, w = stack.height, stack.width
    corners = np.array([[0, 0, 1], [w - 1, 0, 1], [0, h - 1, 1], [w - 1, h - 1, 1]], np.float64).T
    errs = []
    t = time.perf_counter()
    regs = [estimate_star_transform(stars[0], stars[i], "affine") for i in idx[1:]]
    fit_s = (time.perf_counter() - t) / max(1, f - 1)
    for i, reg in zip(idx[1:], regs):
        got = np.asarray(reg.matrix, np.float64).reshape(3, 3) @ corners
        exp = np.asarray(a["mats"][i], np.float64) @ corners
        errs.append(float(np.max(np.hypot(got[0] - exp[0], got[1] - exp[1]))))
    # star catalogues on the same resident frames: grid top-k + separation
    # filter (16 x 12 cells, 4 per cell, 400 kept) and the catalogue
    # translation vote (top 64 by flux), threshold median + 5 * 1.4826 MAD
    # of frame 0
    st0 = stack.frame_global_stats(0, exact_median=True)
    thr = float(st0["median"] + 5.0 * 1.4826 * st0["mad"])
    cat = {"catalogue_threshold": round(thr, 3)}
    if np.isfinite(thr):
        stack.star_grid_top_nms_candidates(0, thr, 16, 12, 4, 400, 8.0)
        t = time.perf_counter()
        grids = [stack.star_grid_top_nms_candidates(i, thr, 16, 12, 4, 400, 8.0) for i in idx]
        cat["grid_catalogue_ms_per_frame"] = round((time.perf_counter() - t) / f * 1e3, 2)
        cat["grid_catalogue_local_maxima_median"] = int(np.median([g["count"] for g in grids]))
        t = time.perf_counter()
        tr = [stack.estimate_translation_from_stars_to_reference(0, i, thr, 64, 1.0) for i in idx[1:]]
        cat["catalogue_translation_ms_per_frame"] = round((time.perf_counter() - t) / max(1, f - 1) * 1e3, 2)
        cat["catalogue_translation_mutual_min"] = int(min(x["mutual_inliers"] for x in tr)) if tr else 0
    if np.isfinite(thr):
        cat["triangle_route"] = triangle_route(stack, light_raw, a, thr, corners)
    out = {"frames": f, "detect_ms_per_frame": round(detect_s * 1e3, 2), "fit_ms_per_frame": round(fit_s * 1e3, 2),
           **cat,
           "stars_per_frame": round(float(np.mean([len(x) for x in stars])), 1),
           "matched_min": int(min(r.matched_stars for r in regs)) if regs else 0,
           "corner_err_px_max": round(max(errs), 4) if errs else None,
           "corner_err_px_median": round(float(np.median(errs)), 4) if errs else None}
    log(f"registration: {out}")
    return out


def triangle_route(stack, light_raw, a, thr, corners):
    """The resident default registration (star catalogues with centroids ->
    triangle descriptors -> similarity fits, glass_amd.triangles.
    register_resident_frames) over every calibrated, unwarped light against
    frame 0, at the route's defaults (64 stars, 5 neighbours, 1200
    descriptors, tolerance 1 px, descriptor radius 0.1).  Wall clock of the
    whole chain per moving frame, and the corner error of each fit against
    the generator's map."""
    import numpy as np
    import torch

    from glass_amd.triangles import register_resident_frames

    f = stack.frame_count
    idx = list(range(f))
    stack.calibrate_frames_device(idx, light_raw, "fits_u16be_bzero", [a["light_exp"]] * f, a["dark_exp"],
                                  a["policy"])
    torch.cuda.synchronize()
    register_resident_frames(stack, 0, [1], thr)  # sizes the workspaces
    route = register_resident_frames(stack, 0, idx[1:], thr)
    errs, inl = [], []
    for i, fit in route["fits"].items():
        got = np.asarray(fit["matrix"], np.float64) @ corners
        exp = np.asarray(a["mats"][i], np.float64) @ corners
        errs.append(float(np.max(np.hypot(got[0] - exp[0], got[1] - exp[1]))))
        inl.append(int(fit["inliers"]))
    t = route["timing_s"]
    n = max(1, f - 1)
    out = {"frames": f - 1, "fitted": len(route["fits"]),
           "catalogue_ms_per_frame": round(t["catalogues"] / f * 1e3, 3),
           "descriptors_ms_per_frame": round(t["descriptors"] / f * 1e3, 3),
           "similarity_ms_per_frame": round(t["similarity"] / n * 1e3, 3),
           "total_ms": round(t["total"] * 1e3, 2),
           "stars_min": int(min(route["catalogue_counts"].values())),
           "inliers_min": min(inl) if inl else 0,
           "corner_err_px_max": round(max(errs), 4) if errs else None,
           "corner_err_px_median": round(float(np.median(errs)), 4) if errs else None}
    log(f"triangle route: {out}")
    return out


def fused_leg(args, stack, light_raw, a):
    """integrate_matrix_warped_mean / _sigma_clip over the calibrated,
    unwarped stack (every frame resampled inside the integration, no warped
    stack written), bilinear and Lanczos3; wall clock per call (each call is
    synchronous), second of two runs."""
    import torch

    f = stack.frame_count
    stack.calibrate_frames_device(list(range(f)), light_raw, "fits_u16be_bzero", [a["light_exp"]] * f, a["dark_exp"],
                                  a["policy"])
    torch.cuda.synchronize()
    out = {}
    for interp in ("bilinear", "lanczos3"):
        for rej in ("mean", "sigma_clip"):
            for _ in range(2):
                t = time.perf_counter()
                if rej == "mean":
                    stack.integrate_matrix_warped_mean(a["mats"], a["weights"], interp, -1.0, "master_only")
                else:
                    stack.integrate_matrix_warped_sigma_clip(a["mats"], a["weights"], interp, -1.0, 3.0, 3.0, True,
                                                             "master_only")
                dt = time.perf_counter() - t
            out[f"{interp}_{rej}_ms"] = round(dt * 1e3, 1)
            log(f"fused {interp} {rej}: {dt * 1e3:.1f} ms")
    return out


def host_raw_leg(args, stack, light_raw, a):
    """The plain step with the raw lights handed over in pinned host memory
    instead of HBM: every frame crosses PCIe through the library's upload
    ring while earlier frames calibrate and warp (8-frame chunks on one
    stream), then the integration runs as in the timed steps.  This is the PCIe-inclusive rate (never `value`).  Two runs,
    the second reported; wall clock around each, HIP events for the stages."""
    import torch

    f = stack.frame_count
    raw_bytes = light_raw.numel() * light_raw.element_size()
    t = time.perf_counter()
    host = torch.empty(raw_bytes, dtype=torch.uint8, pin_memory=True)
    host.copy_(light_raw.reshape(-1).view(torch.uint8))
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t
    sh = a["stream"].cuda_stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    maps_host = {k: torch.empty(v.shape, dtype=v.dtype, pin_memory=True) for k, v in a["outs"].items()}
    maps_bytes = sum(v.numel() * v.element_size() for v in a["outs"].values())
    out = {}
    for run in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev[0].record(a["stream"])
        stack.reset_warp_coverage()
        for _ in range(f - len(a["moving"])):
            stack.accumulate_full_warp_coverage_frame()
        # chunks of 8 frames: the warp of a chunk runs while the next
        # chunk's frames are still crossing PCIe
        fb = raw_bytes // f
        moving = set(a["moving"])
        for c0 in range(0, f, 8):
            chunk = list(range(c0, min(f, c0 + 8)))
            stack.calibrate_frames_host(chunk, host[c0 * fb:(chunk[-1] + 1) * fb], "fits_u16be_bzero",
                                        [a["light_exp"]] * len(chunk), a["dark_exp"], a["policy"], stream=sh)
            warp = [i for i in chunk if i in moving]
            if warp:
                stack.warp_frames(warp, a["mats"][warp], args.interpolation, "cpu", stream=sh)
        ev[1].record(a["stream"])
        stack.integrate_to_device(a["outs"], a["weights"], "winsorized_sigma", 3.0, 3.0, 3, 0.5, "uint16",
                                  dq_geometric=True, dq_active_frames=a["n_active"], stream=sh)
        ev[2].record(a["stream"])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # the output maps back to pinned host memory (the end of a run)
        with torch.cuda.stream(a["stream"]):
            for k, v in a["outs"].items():
                maps_host[k].copy_(v, non_blocking=True)
            ev[3].record(a["stream"])
        torch.cuda.synchronize()
        wall_maps = time.perf_counter() - t0
        upload_ms = ev[0].elapsed_time(ev[1])
        out = {"ms": round(wall * 1e3, 2), "frame_mpix_s": round(f * stack.pixels_per_frame / 1e6 / wall, 1),
               "upload_calibrate_warp_ms": round(upload_ms, 2),
               "upload_gbps": round(raw_bytes / (upload_ms * 1e-3) / 1e9, 1),
               "integrate_ms": round(ev[1].elapsed_time(ev[2]), 2), "raw_gb": round(raw_bytes / 1e9, 2),
               "maps_download_ms": round(ev[2].elapsed_time(ev[3]), 2), "maps_gb": round(maps_bytes / 1e9, 3),
               "end_to_end_with_maps_ms": round(wall_maps * 1e3, 2),
               "pinned_setup_s": round(setup_s, 2)}
        log(f"host raw run {run}: {out}")
    del host
    return out


def local_norm_leg(args, stack, weights):
    """Resident local normalisation of every active frame to frame 0 on the
    calibrated + warped stack left by the timed steps: the grid model
    (pair statistics of all sources in one batched call, one batched
    application) and the global model (per-frame statistics and application,
    the reference's frame loop).  Each call is synchronous; times are wall
    clock around the calls.  Algorithmic bytes: statistics read every source
    and the reference once (4 B/px each), the application reads and writes
    each source (8 B/px)."""
    import numpy as np

    from glass_amd.local_norm import resident_local_normalization

    npix = stack.pixels_per_frame
    tile = args.local_norm_tile
    sources = np.asarray([i for i in range(stack.frame_count) if weights[i] > 0 and i != 0], np.int32)
    n = int(sources.size)
    # statistics alone, twice (the first call also allocates its buffers)
    for _ in range(2):
        t = time.perf_counter()
        batch = stack.frame_pair_grid_stats_batch(0, sources, tile, tile)
        stats_s = time.perf_counter() - t
    # the batched application alone (unit coefficients leave every value as it is)
    rows, cols = batch["grid_rows"], batch["grid_cols"]
    ones = np.ones((n, rows, cols), np.float32)
    zeros = np.zeros((n, rows, cols), np.float32)
    for _ in range(2):
        t = time.perf_counter()
        stack.apply_grid_normalization_frames(sources, ones, zeros, tile, tile)
        apply_s = time.perf_counter() - t
    out = {"sources": n, "tile": tile, "grid": [rows, cols]}
    for mode in ("grid_mean_std", "global_mean_std"):
        t = time.perf_counter()
        res = resident_local_normalization(stack, 0, weights, mode, tile)
        total_s = time.perf_counter() - t
        statuses = {}
        for r in res["frames"]:
            statuses[r["status"]] = statuses.get(r["status"], 0) + 1
        out[mode] = {"total_ms": round(total_s * 1e3, 2), "stats_ms": round(res["timing"]["stats_s"] * 1e3, 2),
                     "apply_ms": round(res["timing"]["apply_s"] * 1e3, 2), "statuses": statuses}
        log(f"local norm {mode}: {total_s * 1e3:.1f} ms ({statuses})")
    stats_bytes = 4.0 * npix * (n + 1)
    apply_bytes = 8.0 * npix * n
    out["grid_stats_ms"] = round(stats_s * 1e3, 2)
    out["grid_stats_gbps"] = round(stats_bytes / stats_s / 1e9, 1)
    out["grid_apply_ms"] = round(apply_s * 1e3, 2)
    out["grid_apply_gbps"] = round(apply_bytes / apply_s / 1e9, 1)
    return out


def measured_traffic(args, stage):
    path = ROOT / "profiles" / "hbm_traffic_200x9600x6422.json"
    default = (args.frames, args.height, args.width, args.interpolation, args.zero_weight) == (200, 6422, 9600,
                                                                                                "lanczos3", 7)
    if not default or not path.exists():
        return None, None
    k = json.loads(path.read_text())["kernels"].get(_TRAFFIC_KERNELS[stage])
    if k is None:
        return None, None
    return int(k["read_bytes"] + k["write_bytes"]), str(path.relative_to(ROOT))


def measured_valu(args, stage):
    """VALU wave-instructions per launch of a stage's kernel from the committed
    SQ_INSTS_VALU profile of the default workload (None otherwise)."""
    path = ROOT / "profiles" / "hbm_traffic_200x9600x6422.json"
    default = (args.frames, args.height, args.width, args.interpolation, args.zero_weight) == (200, 6422, 9600,
                                                                                                "lanczos3", 7)
    if not default or not path.exists():
        return None
    k = json.loads(path.read_text())["kernels"].get(_TRAFFIC_KERNELS[stage], {})
    return k.get("valu_wave_insts")


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the resident stacking path runs on one GPU; --gpus > 1 is not supported "
                                   "(more GPUs would shard by output row bands with a warp halo)"}))
        sys.exit(2)
    import numpy as np
    import torch

    import glass_amd
    from glass_amd import pipeline as P
    from glass_amd.synthetic_gpu import make_calibration_raw, make_light_raw

    if not glass_amd.cuda_available():
        raise SystemExit("no HIP device / libglass_hip.so not loadable")
    dev = torch.device("cuda:0")
    f, h, w = args.frames, args.height, args.width
    npix = h * w
    light_exp, dark_exp = 60.0, 60.0
    log(f"generating {args.calib_frames}x3 calibration frames and {f} lights at {w}x{h} on the GPU")
    t = time.perf_counter()
    bias_raw = make_calibration_raw("bias", args.calib_frames, h, w, args.seed + 10)
    dark_raw = make_calibration_raw("dark", args.calib_frames, h, w, args.seed + 20, dark_exposure_s=dark_exp)
    flat_raw = make_calibration_raw("flat", args.calib_frames, h, w, args.seed + 30)
    blurred = list(range(1, 1 + args.zero_weight))
    light_raw, mats = make_light_raw(f, h, w, args.seed, blurred=blurred, light_exposure_s=light_exp)
    torch.cuda.synchronize()
    log(f"generated in {time.perf_counter() - t:.1f} s; building masters")
    t = time.perf_counter()
    masters = P.build_masters_device(bias_raw, dark_raw, flat_raw, h, w)
    torch.cuda.synchronize()
    master_s = time.perf_counter() - t
    del bias_raw, dark_raw, flat_raw
    torch.cuda.empty_cache()
    log(f"masters built in {master_s:.2f} s")

    weights = np.ones(f, np.float32)
    weights[blurred] = 0.0
    n_active = int(np.count_nonzero(weights))
    moving = [i for i in range(f) if not np.array_equal(mats[i], np.eye(3))]
    stack = glass_amd.ResidentCalibratedStack(f, h, w)
    stack.set_calibration_masters(masters["bias"], masters["dark"], masters["flat"])
    outs = {
        "master": torch.empty((h, w), dtype=torch.float32, device=dev),
        "weight": torch.empty((h, w), dtype=torch.float32, device=dev),
        "coverage": torch.empty((h, w), dtype=torch.int16, device=dev),
        "low": torch.empty((h, w), dtype=torch.int16, device=dev),
        "high": torch.empty((h, w), dtype=torch.int16, device=dev),
        "dq": torch.empty((h, w), dtype=torch.int32, device=dev),
    }
    # a non-null stream: every kernel of the step is launched on it and the
    # HIP events below are recorded on it
    stream = torch.cuda.Stream(device=dev)
    sh = stream.cuda_stream
    policy = {"master_dark_includes_bias": True}
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    stage_ms = {"calibrate": [], "warp": [], "integrate": []}

    step_args = {"moving": moving, "mats": mats, "outs": outs, "weights": weights, "n_active": n_active,
                 "light_exp": light_exp, "dark_exp": dark_exp, "policy": policy, "stream": stream}

    def step(timed):
        if timed:
            ev[0].record(stream)
        stack.calibrate_frames_device(list(range(f)), light_raw, "fits_u16be_bzero", [light_exp] * f, dark_exp,
                                      policy, stream=sh)
        if timed:
            ev[1].record(stream)
        stack.reset_warp_coverage()
        for _ in range(f - len(moving)):
            stack.accumulate_full_warp_coverage_frame()
        stack.warp_frames(moving, mats[moving], args.interpolation, "cpu", stream=sh)
        if timed:
            ev[2].record(stream)
        stack.integrate_to_device(outs, weights, "winsorized_sigma", 3.0, 3.0, 3, 0.5, "uint16",
                                  dq_geometric=True, dq_active_frames=n_active, stream=sh)
        if timed:
            ev[3].record(stream)
            torch.cuda.synchronize()
            stage_ms["calibrate"].append(ev[0].elapsed_time(ev[1]))
            stage_ms["warp"].append(ev[1].elapsed_time(ev[2]))
            stage_ms["integrate"].append(ev[2].elapsed_time(ev[3]))

    if args.copy_probe_mib > 0:
        probe_src = torch.ones(args.copy_probe_mib * 2**18, dtype=torch.float32, device=dev)
        probe_dst = torch.empty_like(probe_src)
        probe_dst.copy_(probe_src)
        torch.cuda.synchronize()
        del probe_src, probe_dst
        log(f"copy probe: {args.copy_probe_mib} MiB device-to-device")
    for i in range(args.warmup):
        step(False)
        torch.cuda.synchronize()
        log(f"warmup step {i} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(True)
        log(f"step {i}: calibrate {stage_ms['calibrate'][-1]:.1f} ms, warp {stage_ms['warp'][-1]:.1f} ms, "
            f"integrate {stage_ms['integrate'][-1]:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    frame_mpix = f * npix / 1e6
    value = frame_mpix / (ms_per_step / 1e3)

    # Algorithmic bytes (staged resident model): calibrate 2 B raw + 4 B f32 per
    # frame-pixel (+12 B/px masters once), warp 8 B per moved frame-pixel,
    # integrate 4 B per active frame-pixel + 18 B/px of output maps.
    med = {k: float(np.median(v)) for k, v in stage_ms.items()}
    bytes_cal = 6.0 * f * npix + 12.0 * npix
    bytes_warp = 8.0 * len(moving) * npix
    bytes_int = 4.0 * n_active * npix + 18.0 * npix
    gbps = {
        "calibrate": bytes_cal / (med["calibrate"] * 1e-3) / 1e9,
        "warp": bytes_warp / (med["warp"] * 1e-3) / 1e9,
        "integrate": bytes_int / (med["integrate"] * 1e-3) / 1e9,
    }
    dominant = max(med, key=med.get)
    launches = {"calibrate": -(-f // 16), "warp": -(-len(moving) // 8), "integrate": 1}[dominant]
    per_launch_bytes = {"calibrate": bytes_cal, "warp": bytes_warp, "integrate": bytes_int}[dominant] / launches
    avg_launch_ms = med[dominant] / launches
    total_bytes = bytes_cal + bytes_warp + bytes_int
    traffic, traffic_src = measured_traffic(args, dominant)
    line = {
        "metric": "calib+warp+reject+integrate frame-Mpix/s & HBM %peak, 200x9600x6422",
        "value": round(value, 2),
        "unit": "frame-Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (u16 raw, f64 statistics)",
        "data": "synthetic (GPU-generated, seeded)",
        "config": {
            "workload": f"{f} lights x {w}x{h} u16 FITS raw resident -> calibrate -> {args.interpolation} warp -> "
                        f"winsorized sigma 3/3 integrate + maps",
            "frames": f, "width": w, "height": h, "active_frames": n_active, "warped_frames": len(moving),
            "interpolation": args.interpolation, "calibration_frames": args.calib_frames,
        },
        "stages_ms": {k: round(v, 3) for k, v in med.items()},
        "stages_gbps": {k: round(v, 1) for k, v in gbps.items()},
        "path_hbm_frac": round(total_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
        "master_build_s": round(master_s, 3),
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": round(per_launch_bytes / (avg_launch_ms * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(per_launch_bytes / (avg_launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "launches_per_step": launches,
            "avg_launch_ms": round(avg_launch_ms, 4),
        },
    }
    valu = measured_valu(args, dominant)
    if valu:
        rate = valu * 64 / (avg_launch_ms * 1e-3)
        line["compute_roofline"] = {
            "kernel": dominant, "bound": "valu", "achieved": round(rate / 1e12, 2), "peak": round(
                VALU_PEAK_LANE_INSTS / 1e12, 2), "unit": "T lane-instr/s", "frac": round(rate / VALU_PEAK_LANE_INSTS, 4),
            "valu_wave_insts_per_launch": valu, "source": "profiles/hbm_traffic_200x9600x6422.json "
                                                           "(SQ_INSTS_VALU, profiles/sq_counters_warp_integrate_"
                                                           "200x9600x6422.md)"}
    if not args.no_registration:
        log("star-registration leg: detect stars and fit affine matrices on calibrated, unwarped lights")
        line["registration"] = registration_leg(args, stack, light_raw, step_args)
    if args.fused_leg:
        log("fused matrix-warped integrations on the calibrated (unwarped) stack")
        line["fused_matrix_warped"] = fused_leg(args, stack, light_raw, step_args)
    if not args.no_host_raw:
        log("host-raw leg: raw lights in pinned host memory, upload overlapped with calibration")
        line["host_raw"] = host_raw_leg(args, stack, light_raw, step_args)
    if not args.no_local_norm:
        log("local-normalisation leg (grid and global models on the warped stack)")
        line["local_normalization"] = local_norm_leg(args, stack, weights)
    stack.close()
    del light_raw
    torch.cuda.empty_cache()
    if not args.no_refine:
        log("refined-registration leg (config 4)")
        line["refined_registration"] = refine_leg(args)
    if not args.no_batch:
        log("batched multi-job leg (config 3)")
        line["batched"] = batched_leg(args)
    if not args.no_cpu_baseline:
        log("timing the CPU baseline sample")
        line["cpu_baseline"] = cpu_baseline(args, light_exp, dark_exp)
        if _REFINE_STACK_CHECK and "refined_registration" in line:
            log("refined-leg stack band against the CPU oracle")
            line["refined_registration"]["stack_vs_oracle"] = refine_stack_vs_oracle()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
