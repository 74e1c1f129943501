"""Batched multi-job stacking (gh_batch_*) against the per-job path: the same
jobs through warp_frames + integrate_to_device must give identical maps, and
every replay of the captured hipGraph (with eager work and synchronisations
between replays) must equal the eager result."""

from __future__ import annotations

import math

import numpy as np
import pytest

from glass_amd import ResidentCalibratedStack
from glass_amd.batch import StackingBatch

pytestmark = pytest.mark.gpu

KEYS = ("master", "weight", "coverage", "low", "high", "dq")


def _rot(deg, tx, ty):
    a = math.radians(deg)
    return np.array([[math.cos(a), -math.sin(a), tx], [math.sin(a), math.cos(a), ty], [0, 0, 1.0]])


def _outputs(h, w):
    import torch

    o = {k: torch.zeros((h, w), dtype=torch.float32, device="cuda") for k in ("master", "weight")}
    for k in ("coverage", "low", "high"):
        o[k] = torch.zeros((h, w), dtype=torch.int16, device="cuda")
    o["dq"] = torch.zeros((h, w), dtype=torch.int32, device="cuda")
    return o


SPECS = [
    # (h, w, frames, interpolation, semantics, fill, weights kind, nan pixel)
    (70, 90, 5, "lanczos3", "cpu", 0.0, None, True),
    (96, 200, 12, "lanczos3", "cpu", math.nan, "zero", True),
    (128, 160, 17, "bilinear", "cpu", 0.0, "zero", True),
    (33, 47, 9, "lanczos3", "resident", math.nan, None, True),
    (200, 130, 40, "lanczos3", "cpu", 0.0, "varied", False),
    (80, 100, 70, "lanczos3", "cpu", 0.0, "zero", False),
]


def _make_jobs(seed):
    rng = np.random.default_rng(seed)
    jobs = []
    for h, w, f, interp, sem, fill, wkind, nan in SPECS:
        frames = rng.normal(1000, 10, size=(f, h, w)).astype(np.float32)
        frames[rng.random(frames.shape) < 0.003] += 400.0
        if nan:
            frames[1, h // 3, w // 4] = np.nan
        s = ResidentCalibratedStack(f, h, w)
        for i in range(f):
            s.upload_calibrated_frame(i, frames[i])
        moving = list(range(1, f))
        mats = [_rot(rng.uniform(-1, 1), rng.uniform(-3, 3), rng.uniform(-3, 3)) for _ in moving]
        weights = None
        if wkind == "zero":
            weights = np.ones(f, np.float32)
            weights[2] = 0.0
        elif wkind == "varied":
            weights = rng.uniform(0.5, 2.0, f).astype(np.float32)
        jobs.append({"stack": s, "warp_indices": moving, "matrices": mats, "interpolation": interp,
                     "semantics": sem, "fill": fill, "weights": weights, "outputs": _outputs(h, w),
                     "dq_geometric": True, "dq_active_frames": f})
    return jobs


def _serial(job):
    s = job["stack"]
    s.reset_warp_coverage()
    for _ in range(s.frame_count - len(job["warp_indices"])):
        s.accumulate_full_warp_coverage_frame()
    s.warp_frames(job["warp_indices"], job["matrices"], job["interpolation"], job["semantics"], fill=job["fill"])
    out = _outputs(s.height, s.width)
    s.integrate_to_device(out, weights=job["weights"], dq_geometric=True, dq_active_frames=s.frame_count)
    return out


def _host(o):
    import torch

    torch.cuda.synchronize()
    return {k: o[k].cpu().numpy() for k in KEYS}


@pytest.mark.parametrize("use_graph,autotune", [(True, True), (False, False)])
def test_batch_matches_per_job_path(gpu, use_graph, autotune):
    jobs = _make_jobs(1)
    with StackingBatch(jobs, use_graph=use_graph, autotune=autotune) as b:
        b.run()
        got = [_host(j["outputs"]) for j in jobs]
        info = b.info()
    assert info["jobs"] == len(jobs) and info["graph"] == (1 if use_graph else 0)
    for job, g in zip(jobs, got):
        exp = _host(_serial(job))
        for k in KEYS:
            assert np.array_equal(g[k], exp[k], equal_nan=True), (k, job["interpolation"], job["semantics"])
        job["stack"].close()


def test_graph_replays_match_eager(gpu):
    import torch

    jobs = _make_jobs(2)
    with StackingBatch(jobs, use_graph=True, autotune=False) as b:
        b.run()
        first = [_host(j["outputs"]) for j in jobs]
        for rep in range(3):
            for j in jobs:  # eager work between replays
                for k in KEYS:
                    j["outputs"][k].fill_(7)
            torch.cuda.synchronize()
            b.run()
            again = [_host(j["outputs"]) for j in jobs]
            for g, e in zip(again, first):
                for k in KEYS:
                    assert np.array_equal(g[k], e[k], equal_nan=True), (rep, k)
    for j in jobs:
        j["stack"].close()


def _set_interleaved(b, frames_per_item, classes):
    # the first stride the library accepts (coprime to the Lanczos3 work items)
    for p in (97, 89, 83, 79, 73, 71):
        try:
            b.set_schedule(512, p, frames_per_item, classes)
            return p
        except ValueError:
            continue
    raise AssertionError("no coprime stride")


@pytest.mark.parametrize("frames_per_item,classes,interleave", [(2, 0, False), (4, 1, True), (8, 2, True),
                                                                (2, 2, False)])
def test_pinned_schedules_match_per_job_path(gpu, frames_per_item, classes, interleave):
    # Every schedule the autotuner may pick gives the per-job maps bit for bit:
    # frames per warp work item, item order, and the integration's capacity
    # classes (each job at its smallest capacity, powers of two, one launch).
    jobs = _make_jobs(3)
    with StackingBatch(jobs, use_graph=True, autotune=True) as b:
        b.set_schedule(512, 1, frames_per_item, classes)
        if interleave:
            assert _set_interleaved(b, frames_per_item, classes) > 1
            assert b.info()["warp_item_stride"] > 1
        b.run()
        got = [_host(j["outputs"]) for j in jobs]
        info = b.info()
    assert info["warp_frames_per_item"] == frames_per_item and info["integrate_classes"] == classes
    assert info["autotune_trials"] == 0 and info["graph"] == 1
    # one register launch per class and weight kind: 5, 11, 16, 9 and 69
    # positive-weight frames with unit weights and 40 with varied ones make
    # classes 16 / 48 / 80 at the smallest capacities, 16 / 64 / 128 at
    # powers of two, or 80 for both weight kinds in the single-capacity set
    assert info["integrate_launches"] == {0: 3, 1: 3, 2: 2}[classes]
    for job, g in zip(jobs, got):
        exp = _host(_serial(job))
        for k in KEYS:
            assert np.array_equal(g[k], exp[k], equal_nan=True), (k, job["interpolation"], frames_per_item, classes)
        job["stack"].close()


def test_schedule_arguments_are_checked(gpu):
    jobs = _make_jobs(4)[:1]
    with StackingBatch(jobs, use_graph=False, autotune=False) as b:
        for bad in [(512, 1, 3, 0), (512, 1, 8, 3), (0, 1, 8, 0), (512, 0, 8, 0)]:
            with pytest.raises(ValueError):
                b.set_schedule(*bad)
    jobs[0]["stack"].close()
