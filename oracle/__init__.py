"""CPU oracle for the resident stacking path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy (and in ``cpu_ref.c`` for timing), the
reference GLASS CPU path for: master calibration, light calibration, affine
warp (bilinear / Lanczos3), winsorized sigma-clip rejection with weighted-mean
integration, per-frame quality statistics and the resident DQ map.  Every
function cites the reference file:line it follows.

It is the *checker*, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import, call, link or execute anything under ``oracle/``.  The product path
(``glass_amd``) never imports it and fails loudly when its HIP library is
missing.

Pinning status (what checks each function; see tests/test_oracle_*.py):
  calibrate_light            -- reference literals (540.0, 15.0) and the
                                s2_gate_314 golden calibrated FITS (bit-exact)
  build_master_*             -- s2_gate_314 golden master FITS (bit-exact)
  center_and_scale /
  weighted_integrate_stack   -- reference literals (test_cpu_integration,
                                test_stack_engine, test_cuda_resident_stack),
                                the per-pixel brute-force restatement with
                                seeds 594 / 626, closed forms
  warp_matrix                -- per-pixel brute-force restatements of the
                                reference test oracles, closed forms
                                (integer shifts, constant images, linear ramps)
  scan_quality_stats         -- closed forms (np.median / exact sort)
  resident_dq_map            -- explicit rule table cases
  registration.detect_stars  -- reference literal cases (elongated star,
                                star fields), closed forms (centroid of a
                                centred star, float32 median pair average)
  registration.estimate_star_transform
                             -- golden vectors of the reference's own
                                detect_stars / estimate_star_transform
                                (tests/golden/registration_golden.npz,
                                bit-exact incl. lstsq / svd), reference
                                literal cases, exact affine recovery
  warp_matrix_resident clamp -- closed forms (step-edge overshoot lands on
                                the bounds, ramps unchanged)
"""

from .calibration import calibrate_light, decode_fits_u16be_bzero, decode_u16le
from .rejection import center_and_scale
from .integration import weighted_integrate_stack, stack_engine_combine, mean_integrate
from .masters import build_master_bias, build_master_dark, build_master_flat, exact_median
from .warp import warp_matrix, warp_matrix_resident, invert_matrix
from .quality import scan_quality_stats, quality_weights, simple_snr_weight, frame_global_stats
from .dq import resident_dq_map, DQ_NO_DATA, DQ_WARP_EDGE, DQ_LOW_REJECTED, DQ_HIGH_REJECTED

__all__ = [
    "calibrate_light",
    "decode_fits_u16be_bzero",
    "decode_u16le",
    "center_and_scale",
    "weighted_integrate_stack",
    "stack_engine_combine",
    "mean_integrate",
    "build_master_bias",
    "build_master_dark",
    "build_master_flat",
    "exact_median",
    "warp_matrix",
    "warp_matrix_resident",
    "invert_matrix",
    "scan_quality_stats",
    "quality_weights",
    "simple_snr_weight",
    "frame_global_stats",
    "resident_dq_map",
    "DQ_NO_DATA",
    "DQ_WARP_EDGE",
    "DQ_LOW_REJECTED",
    "DQ_HIGH_REJECTED",
]
