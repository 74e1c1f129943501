"""The end-to-end tolerance contract of the resident stacking path, as a
comparator of the GPU maps against the CPU oracle's on the same inputs:

* integrated image: RMS of the difference <= 1e-5 of the RMS of the oracle
  image, and the 99th percentile of |difference| <= 1e-4 of the oracle
  image's range (over pixels the oracle gives weight);
* rejection-count maps (low, high, post-rejection coverage) and the
  zero-weight map identical, except at pixels holding a sample within
  1e-5 sigma of one of the oracle's two rejection thresholds (center -
  low_sigma * scale, center + high_sigma * scale); such pixels are reported.

The samples are the oracle's warped stack (the same array the oracle
integrated) and its validity mask; the thresholds come from the oracle's own
center_and_scale on them, with the frame-level arithmetic choice.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

import oracle
from oracle.rejection import center_and_scale

RMS_REL = 1.0e-5
P99_OF_RANGE = 1.0e-4
THRESHOLD_SIGMA = 1.0e-5


@dataclass
class ContractReport:
    rms_rel: float
    p99_of_range: float
    pixels: int
    count_mismatches: int = 0
    exempt: int = 0
    exempt_pixels: list = field(default_factory=list)

    def __str__(self):
        return (f"rms_rel={self.rms_rel:.3g} p99/range={self.p99_of_range:.3g} pixels={self.pixels} "
                f"count mismatches={self.count_mismatches} (exempt {self.exempt})")


def near_threshold(samples, valid, weights, low_sigma=3.0, high_sigma=3.0, all_valid=None,
                   eps=THRESHOLD_SIGMA):
    """Per pixel: whether a valid sample of a positive-weight frame lies within
    eps * scale of the oracle's low or high rejection threshold."""
    x = np.asarray(samples, np.float32)
    w = np.asarray(weights, np.float32)
    ok = np.asarray(valid, bool) & np.isfinite(x) & (np.isfinite(w) & (w > 0))[:, None, None]
    center, scale = center_and_scale(x, ok, method="winsorized_sigma", low_sigma=low_sigma,
                                     high_sigma=high_sigma, all_valid=all_valid)
    lo = (center - np.float32(low_sigma) * scale).astype(np.float64)
    hi = (center + np.float32(high_sigma) * scale).astype(np.float64)
    tol = eps * np.abs(scale.astype(np.float64))
    xd = x.astype(np.float64)
    close = ok & ((np.abs(xd - lo[None]) <= tol[None]) | (np.abs(xd - hi[None]) <= tol[None]))
    return np.any(close, axis=0)


def check_contract(got, exp, samples=None, valid=None, weights=None, low_sigma=3.0, high_sigma=3.0,
                   all_valid=None) -> ContractReport:
    """got / exp: (master, weight, coverage, low, high) maps of one region.
    Raises AssertionError when the contract does not hold."""
    gm, gw, gc, gl, gh = (np.asarray(a, np.float64) for a in got[:5])
    em, ew, ec, el, eh = (np.asarray(a, np.float64) for a in exp[:5])
    assert gm.shape == em.shape
    has = ew > 0
    diff = np.abs(gm - em)[has]
    ref = em[has]
    rms_ref = float(np.sqrt(np.mean(ref * ref))) if ref.size else 0.0
    rms_rel = float(np.sqrt(np.mean(diff * diff)) / rms_ref) if ref.size and rms_ref > 0 else 0.0
    rng = float(ref.max() - ref.min()) if ref.size else 0.0
    p99 = float(np.percentile(diff, 99)) if diff.size else 0.0
    report = ContractReport(rms_rel, p99 / rng if rng > 0 else 0.0, int(ref.size))
    mismatch = (gc != ec) | (gl != el) | (gh != eh) | ((gw > 0) != (ew > 0))
    report.count_mismatches = int(mismatch.sum())
    if report.count_mismatches:
        assert samples is not None, "count maps differ and no samples were given for the exemption check"
        near = near_threshold(samples, valid, weights, low_sigma, high_sigma, all_valid)
        report.exempt = int((mismatch & near).sum())
        report.exempt_pixels = [tuple(int(v) for v in p) for p in np.argwhere(mismatch & near)[:20]]
        assert report.exempt == report.count_mismatches, \
            f"{report.count_mismatches - report.exempt} count-map mismatches away from a threshold: {report}"
    assert report.rms_rel <= RMS_REL, str(report)
    assert p99 <= P99_OF_RANGE * rng, str(report)
    return report


def oracle_block(stack, coverage, weights, low_sigma=3.0, high_sigma=3.0, min_samples=3,
                 max_reject_fraction=0.5, all_valid=None):
    """The oracle's five maps of one block (weighted_integrate_stack)."""
    return oracle.weighted_integrate_stack(stack, coverage=coverage, weights=weights, rejection="winsorized_sigma",
                                           low_sigma=low_sigma, high_sigma=high_sigma, min_samples=min_samples,
                                           max_reject_fraction=max_reject_fraction, all_valid=all_valid)
