/*
 * glass_hip.h -- C ABI of libglass_hip.so, the MI355X (gfx950) resident
 * stacking path: master calibration -> affine warp (bilinear / Lanczos3) ->
 * winsorized sigma-clip rejection with weighted-mean integration, plus the
 * output maps (weight, coverage, low/high rejection counts, DQ).
 *
 * The entry points replace the hot-path methods of the reference's pybind11
 * class ResidentCalibratedStack (cpp/src/native_bindings.cpp:2915, bound at
 * :20624-21213), which the reference reaches through glass_cuda.py:2653.
 * Each declaration below cites the reference method it replaces.
 *
 * Conventions
 *  - Every function returns an int status: GH_OK (0) or one of the error codes
 *    below; gh_last_error() returns the message of the last failure on the
 *    calling thread.  GH_ERR_VALUE maps to Python ValueError
 *    (std::invalid_argument in the reference), GH_ERR_INDEX to IndexError
 *    (std::out_of_range), GH_ERR_RUNTIME to RuntimeError (check_cuda,
 *    native_bindings.cpp:1011).
 *  - Frames live in HBM as float32 row-major [height][width] slots owned by
 *    the stack object.  Pointers marked "host or device" are classified with
 *    hipPointerGetAttributes; everything else is host memory.
 *  - `stream` is a hipStream_t (NULL = the stack's own stream).  Calls that
 *    return host data synchronise before returning.
 *  - Plain C types only: no torch or numpy types cross this boundary.
 */
#ifndef GLASS_HIP_H
#define GLASS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GH_OK 0
#define GH_ERR_VALUE 1
#define GH_ERR_INDEX 2
#define GH_ERR_RUNTIME 3

typedef struct gh_stack gh_stack;

/* Raw sample formats accepted by the calibration entry points. */
#define GH_FMT_F32 0            /* float32 */
#define GH_FMT_U16 1            /* little-endian uint16 ADU */
#define GH_FMT_FITS_U16BE_BZERO 2 /* FITS BITPIX=16, BZERO=32768, big-endian */

/* Interpolators of the matrix warp. */
#define GH_INTERP_BILINEAR 1
#define GH_INTERP_LANCZOS3 3

/* Warp flags. */
#define GH_TRANSLATE_INTEGER 0
#define GH_WARP_TRACK_COVERAGE 1u     /* add per-pixel footprint to the geometric coverage map */
#define GH_WARP_SKIP_NONFINITE_TAPS 2u /* resident semantics: non-finite Lanczos taps skipped */
#define GH_WARP_AFFINE 4u             /* last inverse row is exactly (0,0,1): no projective divide */

/* Rejection modes of gh_integrate. */
#define GH_REJECT_NONE 0
#define GH_REJECT_WINSORIZED_SIGMA 1
#define GH_REJECT_SIGMA_CLIP 2

/* Count-map dtypes of gh_integrate. */
#define GH_COUNT_F32 0
#define GH_COUNT_U16 1

typedef struct gh_calib_policy {
  int32_t master_dark_includes_bias; /* default 1 (src/glass/models.py:68) */
  float flat_floor;                  /* default 1e-6 */
  float pedestal;                    /* default 0 */
  int32_t accumulate_stats;          /* 1: fused FP64 sum / sum^2 / finite count per frame */
} gh_calib_policy;

typedef struct gh_integrate_params {
  int32_t rejection;          /* GH_REJECT_* */
  float low_sigma;            /* default 3 */
  float high_sigma;           /* default 3 */
  int32_t min_samples;        /* default 3, >= 1 */
  float max_reject_fraction;  /* default 0.5, in [0,1] */
  int32_t count_dtype;        /* GH_COUNT_F32 or GH_COUNT_U16 */
  int32_t statistics_variant; /* -1 auto, 0 per-pixel valid-count arithmetic
                                 (rejection.py:242-296), 1 all-valid arithmetic
                                 (rejection.py:219-239) */
  int32_t require_positive_scale; /* 1: StackEngine rule, no rejection when the
                                     scale is not finite and > 0 (stack_engine.py:586) */
  int32_t tile_size;          /* 0: the statistics variant is chosen once for the
                                 whole frame (weighted_integrate_stack); >0: per
                                 tile of this size (CPUStackEngine tiles) */
} gh_integrate_params;

/* Output pointers of gh_integrate.  Each may be NULL (not produced).  Device
 * pointers are written on the stream; host pointers are filled before return. */
typedef struct gh_integrate_outputs {
  float* master;           /* [H*W] f32 */
  float* weight;           /* [H*W] f32 */
  void* coverage;          /* [H*W] f32 or u16 (post-rejection sample count) */
  void* low;               /* [H*W] f32 or u16 */
  void* high;              /* [H*W] f32 or u16 */
  uint32_t* dq;            /* [H*W] u32 DQ flags (resident_cuda.py:3110-3297) */
  int32_t dq_use_geometric; /* 1: include the geometric warp coverage map in DQ */
  int32_t dq_active_frames; /* active frame count for partial-coverage WARP_EDGE */
} gh_integrate_outputs;

typedef struct gh_frame_statistics {
  double mean;          /* FP64, finite pixels */
  double std;           /* sqrt(max(E[x^2] - mean^2, 0)) */
  double sum;
  double sum_sq;
  int64_t valid_pixels;
  int64_t total_pixels;
  int64_t nonfinite_pixels;
  double median;        /* exact (radix select), even counts average the middle pair */
  double mad;           /* exact median of |x - f32(median)| */
  float min;
  float max;
} gh_frame_statistics;

/* ---- library ------------------------------------------------------------ */
int gh_last_error(char* buffer, size_t size);
int gh_version(void);
/* Number of visible HIP devices (cudaGetDeviceCount, native_bindings.cpp:15572). */
int gh_device_count(int* count);

typedef struct gh_device_properties {
  char name[256];             /* marketing name */
  char arch[64];              /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
  int64_t memory_total_bytes;
  int32_t multiprocessor_count; /* compute units */
  int32_t major;
  int32_t minor;
  int32_t wavefront_size;
} gh_device_properties;

/* Properties of one device: replaces get_device_info / list_devices
 * (device_info_dict, native_bindings.cpp:1932-1944, :15576-15594).
 * GH_ERR_INDEX for an id outside [0, device count). */
int gh_device_info(int device, gh_device_properties* out);

/* Pinned (page-locked, portable) host memory: replaces host_pinned_empty_f32 /
 * host_pinned_empty_u8 (native_bindings.cpp:1946-1990).  GH_ERR_VALUE for 0 bytes. */
int gh_host_alloc(size_t bytes, void** out);
int gh_host_free(void* ptr);

/* ---- lifecycle: ResidentCalibratedStack(frame_count, height, width)
 *      (native_bindings.cpp:2917-2934, destructor :2939-2996) ----------------- */
int gh_stack_create(int64_t frame_count, int64_t height, int64_t width, gh_stack** out);
int gh_stack_destroy(gh_stack* stack);
/* frame_count / height / width / loaded_count / bytes_allocated /
 * warp_coverage_frame_count properties (native_bindings.cpp:20626-20639). */
int gh_stack_info(const gh_stack* stack, int64_t out[8]);
/* Device address of frame `index` (for zero-copy tensor handoff); valid until
 * the next warp of that frame (warps write into a fresh slot). */
int gh_stack_frame_ptr(gh_stack* stack, int64_t index, float** device_ptr);
/* 1 if frame `index` holds any non-finite sample (tracked by every writer). */
int gh_stack_frame_nonfinite(gh_stack* stack, int64_t index, int64_t* count);

/* set_calibration_masters(bias=None, dark=None, flat=None)
 * (native_bindings.cpp:3152, :15398-15414).  NULL = absent.  host or device. */
int gh_set_masters(gh_stack* stack, const float* bias, const float* dark, const float* flat);

/* upload_calibrated_frame(i, f32[H,W]) (native_bindings.cpp:3158). host or device. */
int gh_upload_frame(gh_stack* stack, int64_t index, const float* frame);

/* download_frame_tile / download_frames_tile (native_bindings.cpp:3082-3150):
 * frames indices[0..count) tile [y0,y1)x[x0,x1) into host `out`
 * ([count][y1-y0][x1-x0] f32). */
int gh_download_tile(gh_stack* stack, const int64_t* indices, int64_t count, int64_t x0, int64_t y0,
                     int64_t x1, int64_t y1, float* out);

/* calibrate_frame / calibrate_frames_* / ..._fits_u16be_bzero_...
 * (native_bindings.cpp:5979-8778; kernels cpp/cuda/calibration_kernels.cu:5-186):
 * frame k of the batch is read from `raw` + k*raw_frame_stride_bytes (host or
 * device, format GH_FMT_*) and written, calibrated, into slot indices[k].
 * dark_scales[k] multiplies the dark (1.0 when unscaled).  Host raw frames are
 * uploaded through a ring of device slots on a copy stream, so the upload of
 * frame k + 1 overlaps the calibration of frame k (pinned host memory,
 * e.g. hipHostMalloc, gives the full link rate; the reference's pinned ring
 * and lane streams, resident_cuda.py:1159-1897, native_bindings.cpp:15185-15242).
 * A NULL stream makes the call synchronous; on a caller stream the call is
 * asynchronous and a host `raw` buffer must stay valid until that stream's
 * work has completed. */
int gh_calibrate_frames(gh_stack* stack, const int64_t* indices, int64_t count, const void* raw,
                        int64_t raw_frame_stride_bytes, int32_t format, const float* dark_scales,
                        const gh_calib_policy* policy, void* stream);

/* Per-frame affine normalisation x = (x - offset[k]) / divisor[k] in f32 for
 * frames indices[k]: the per-flat median normalisation of the CPU master-flat
 * build (src/glass/engine/pipeline.py:880-884, `data / np.float32(scalar)`;
 * resident form rc.py:6736-6755). */
int gh_normalize_frames(gh_stack* stack, const int64_t* indices, int64_t count, const float* offsets,
                        const float* divisors, void* stream);

/* apply_matrix_{bilinear,lanczos3}_frame[s] (native_bindings.cpp:8835-9922;
 * kernels cpp/cuda/warp_kernels.cu:206-914): frame indices[k] is resampled onto
 * the reference grid through inverse[k] (row-major 3x3 FP64 inverse of the
 * moving->reference matrix).  Out-of-footprint pixels get `fill`.
 * clamping_threshold < 0 disables the Lanczos3 clamp. */
int gh_warp_frames(gh_stack* stack, const int64_t* indices, int64_t count, const double* inverse,
                   int32_t interpolation, float fill, float clamping_threshold, uint32_t flags, void* stream);

/* reset_warp_coverage / accumulate_full_warp_coverage_frame / warp_coverage_map
 * (native_bindings.cpp:3051-3080).  `out` host or device, [H*W] f32. */
/* apply_translation_frame(i, dx, dy, fill) (mode GH_TRANSLATE_INTEGER, whole
 * pixels) and apply_translation_bilinear_frame(i, dx, dy, fill) (mode
 * GH_INTERP_BILINEAR, dx / dy rounded to float32) (native_bindings.cpp:
 * 8778-8833; kernels cpp/cuda/warp_kernels.cu:103-190): the frame becomes, in
 * place, in[y - dy][x - dx] (integer) or its bilinear sample at
 * (x - dx, y - dy) computed in float32 (valid 0 <= s <= dim - 1, far
 * neighbour clamped), `fill` outside; the geometric coverage map gains 1 where
 * the pixel is inside and counts one more warped frame. */
int gh_translate_frame(gh_stack* stack, int64_t index, int32_t mode, double dx, double dy, float fill,
                       void* stream);

/* estimate_translation_to_reference(reference, moving, max_shift_x, max_shift_y,
 * sample_stride) (subpixel = 0: integer offsets within +-radius) and
 * estimate_translation_subpixel_to_reference(reference, moving, center_dx,
 * center_dy, radius_steps, step, sample_stride) (subpixel = 1: offsets
 * center + k * step, k within +-radius, sampled bilinearly in float32)
 * (native_bindings.cpp:10418-10568; kernels
 * cpp/cuda/registration_kernels.cu:10-264): the offset (dx, dy), moving ->
 * reference, that maximises the normalised cross-correlation of the reference
 * with the moving frame shifted by it, over the reference pixels sampled every
 * `sample_stride` (pairs finite in both).  The first best candidate in
 * row-major order wins.  FP64 sums; `score` rounded to float32. */
int gh_estimate_translation(gh_stack* stack, int64_t reference_index, int64_t moving_index, int32_t subpixel,
                            int32_t radius_x, int32_t radius_y, float center_dx, float center_dy, float step,
                            int32_t sample_stride, float* dx, float* dy, float* score, int64_t* candidates);

int gh_reset_warp_coverage(gh_stack* stack);
int gh_accumulate_full_warp_coverage(gh_stack* stack);
int gh_warp_coverage_map(gh_stack* stack, float* out);

/* integrate_mean(weights) and integrate_hardened_winsorized_sigma(weights, low,
 * high, min_samples, max_reject_fraction, count_map_dtype, ...)
 * (native_bindings.cpp:12935-12990, :13395-14335; kernel
 * cpp/cuda/integration_kernels.cu:2292-2779).  weights: host f32[frame_count]
 * or NULL (unit).  Zero / non-finite weights exclude the frame. */
int gh_integrate(gh_stack* stack, const float* weights, const gh_integrate_params* params,
                 const gh_integrate_outputs* outputs, void* stream);

/* gh_integrate without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (serving loops, code that
 * must not call hipMalloc).  The frame-pointer table, frame weights and
 * per-tile statistics flags live in a fixed library-owned device array of
 * 208 KiB per GPU (up to 16384 positive-weight frames and 16384 statistics
 * tiles; larger requests are GH_ERR_VALUE).  Every output must be device
 * memory or NULL (host outputs are GH_ERR_VALUE).  Results equal gh_integrate's.
 * All calls on one GPU share that array: calls that overlap on one GPU (other
 * streams or threads) race; consecutive calls on one stream are ordered. */
int gh_integrate_noalloc(gh_stack* stack, const float* weights, const gh_integrate_params* params,
                         const gh_integrate_outputs* outputs, void* stream);

/* Parameters of gh_integrate_clip. */
typedef struct gh_clip_params {
  int32_t sampler;            /* GH_CLIP_STACK, GH_INTERP_BILINEAR or GH_INTERP_LANCZOS3 */
  int32_t rejection;          /* GH_REJECT_NONE (weighted mean) or GH_REJECT_SIGMA_CLIP */
  int32_t winsorize;          /* sigma clip: one winsorized re-estimate of centre and scale */
  float low_sigma, high_sigma;
  float clamping_threshold;   /* Lanczos3 sampler: clamp to [min - t*r, max + t*r] of the taps; < 0 off */
} gh_clip_params;

#define GH_CLIP_STACK 0

/* integrate_sigma_clip (native_bindings.cpp:13227-13393; kernel
 * cpp/cuda/integration_kernels.cu:1743-1900) and the fused matrix-warped forms
 * integrate_matrix_warped_mean / integrate_matrix_warped_sigma_clip
 * (native_bindings.cpp:14632-15115; kernels integration_kernels.cu:3840-4123):
 * per pixel over the frames of positive finite weight, in frame order, with
 * float32 sums -- weighted mean of the finite samples, or the mean / population
 * std sigma clip (optionally one winsorized re-estimate), rejected sample
 * counts and the weighted mean of the rest.  The samples are the resident
 * frames as stored (GH_CLIP_STACK) or the frames resampled on the fly through
 * inverses[k] (FP64 3x3, output -> source, k over all frames) with the
 * resident tap rules (non-finite Lanczos3 taps skipped); the stack is not
 * modified.  Outputs (host or device, NULL = not produced) are f32 [H*W]:
 * master, weight sum, kept count, low / high rejection counts and geometric
 * coverage (samples inside the footprint).  Every frame must be loaded. */
int gh_integrate_clip(gh_stack* stack, const float* weights, const double* inverses, const gh_clip_params* params,
                      float* master, float* weight, float* coverage, float* low, float* high, float* geometric,
                      void* stream);

/* integrate_tile_local_mean / integrate_tile_local_sigma_clip
 * (native_bindings.cpp:12998-13225, :14337-14630; kernels
 * cpp/cuda/integration_kernels.cu:1658-1741, :3622-3838): gh_integrate_clip
 * over the stack as stored (params->sampler GH_CLIP_STACK) with a per-pixel
 * weight: a frame f with target_mask[f] != 0 has weight
 * weights[f] * tile_multipliers[t] at the pixels of tile t (1 outside every
 * tile), and a frame whose weight at a pixel is not positive and finite does
 * not enter that pixel's statistics.  tile_extents[t] = (x0, y0, x1, y1), a
 * half-open rectangle inside the frame; tiles must not overlap; multipliers
 * finite and >= 0; target_mask must select a frame.  Outputs as in
 * gh_integrate_clip (the reference's mean returns master and weight only). */
int gh_integrate_tile_local(gh_stack* stack, const float* weights, const uint8_t* target_mask,
                            const int32_t* tile_extents, const float* tile_multipliers, int64_t tile_count,
                            const gh_clip_params* params, float* master, float* weight, float* coverage, float* low,
                            float* high, void* stream);

/* Star candidate catalogues of resident frame `index`
 * (native_bindings.cpp:11356-12122; kernels
 * cpp/cuda/star_detect_kernels.cu:41-742).  Candidates are interior pixels,
 * finite, above `threshold` and not below any 8-neighbour; they rank by flux
 * (descending), then y, then x.
 *   GH_CANDIDATES_ALL       star_candidates: the first max_candidates in
 *                           raster order (the reference's order is that of
 *                           its atomic appends);
 *   GH_CANDIDATES_TOP       star_top_candidates: the best max_candidates in
 *                           rank order;
 *   GH_CANDIDATES_TOP_NMS   star_top_nms_candidates[_centroid]: the best
 *                           max_candidates (= scan_candidates), then greedily
 *                           those at squared distance >= min_separation_px^2
 *                           (float32) from every one kept, up to
 *                           max_output_candidates;
 *   GH_CANDIDATES_GRID_NMS  star_grid_top_nms_candidates[_deterministic]
 *                           [_centroid]: the best candidates_per_cell of each
 *                           cell (min(x * cols / W, cols - 1), same for y) in
 *                           rank order, then the same greedy filter.
 * centroid_radius > 0 moves each output position to its window centroid
 * (radius clamped to 1..8; background the window's float32 median of finite
 * samples, or the frame's finite mean with centroid_global_mean_background;
 * FP64 weight sums) and writes centroid_status (1 = moved; may be NULL).
 * x, y, flux hold max_candidates (ALL, TOP) or max_output_candidates entries.
 * Synchronous. */
#define GH_CANDIDATES_ALL 0
#define GH_CANDIDATES_TOP 1
#define GH_CANDIDATES_TOP_NMS 2
#define GH_CANDIDATES_GRID_NMS 3

typedef struct gh_candidate_params {
  int32_t mode;
  float threshold;
  int32_t max_candidates;          /* ALL / TOP: catalogue size; TOP_NMS: scan_candidates */
  int32_t max_output_candidates;   /* the NMS modes */
  int32_t grid_cols, grid_rows, candidates_per_cell;
  float min_separation_px;
  int32_t centroid_radius;         /* 0: off */
  int32_t centroid_global_mean_background;
} gh_candidate_params;

typedef struct gh_candidate_result {
  int64_t count;                   /* local maxima above the threshold */
  int64_t stored_count;            /* entries written */
  int64_t selected_count;          /* ranked entries scanned by the separation filter */
  int64_t refined_count;           /* centroids moved */
  float centroid_background;       /* global-mean background used, else NaN */
  float max_shift_px;              /* largest centroid move */
} gh_candidate_result;

int gh_star_candidates(gh_stack* stack, int64_t index, const gh_candidate_params* params, float* x, float* y,
                       float* flux, uint8_t* centroid_status, gh_candidate_result* result);

/* estimate_translation_from_stars_to_reference (native_bindings.cpp:12645-
 * 12880; kernels cpp/cuda/registration_kernels.cu:265-520, :2007-2110):
 * reference - moving translation from star catalogues of both frames (the
 * best max_candidates, or with grid_cols / grid_rows > 0 the brightest
 * candidate of each cell in rank order, padded with (0, 0) entries to
 * min(count, cells) like the reference's cell slots).  Every pair offset
 * inside |dx| <= max_abs_dx, |dy| <= max_abs_dy (< 0: no limit; max_abs_dy
 * < 0 takes max_abs_dx) and within prior_radius_px of the prior (< 0: none)
 * votes; the offset with the most offsets within tolerance_px (float32
 * squared distances; first in pair order) is (dx, dy).  Mutual nearest
 * matches within tolerance under it give refined_dx / refined_dy (their mean
 * offset; the best offset when there are none) and rms_px.  Throws when a
 * catalogue is empty.  Synchronous. */
typedef struct gh_star_translation_params {
  float threshold;
  int32_t max_candidates;
  float tolerance_px, max_abs_dx, max_abs_dy, prior_dx, prior_dy, prior_radius_px;
  int32_t grid_cols, grid_rows;
} gh_star_translation_params;

typedef struct gh_star_translation_result {
  float dx, dy, refined_dx, refined_dy, rms_px;
  float max_abs_dx, max_abs_dy, prior_radius_px;  /* after the defaults above */
  int64_t inliers, mutual_inliers;
  int64_t reference_count, moving_count, reference_total_count, moving_total_count;
} gh_star_translation_result;

int gh_estimate_translation_from_stars(gh_stack* stack, int64_t reference_index, int64_t moving_index,
                                       const gh_star_translation_params* params, gh_star_translation_result* result);

/* refine_matrix_translation_candidates_to_reference (native_bindings.cpp:
 * 10079-10184; helpers :1459-1700; kernel cpp/cuda/warp_kernels.cu:1051-1172):
 * for each seed matrix (float32 [seed_count][9], output -> source), the
 * corrections (dx, dy) of a float32 grid of radius search_radius_px and step
 * coarse_step_px are added to the matrix translation and scored on the
 * reference's pixels every coarse_sample_stride: the moving frame sampled
 * bilinearly through the float32 inverse, FP64 sums giving rms, mean |diff|
 * and ncc over the pixels inside the source frame with finite values.  The
 * best (lowest rms, then highest ncc, then most pixels) is rescored on a fine
 * grid around it (fine_radius_px > 0, stride final_sample_stride), or alone at
 * the final stride.  seeds[k] receives each seed's best; *selected_index the
 * best seed.  Synchronous. */
typedef struct gh_align_search_params {
  float search_radius_px, coarse_step_px, fine_radius_px, fine_step_px;
  int32_t coarse_sample_stride, final_sample_stride;
} gh_align_search_params;

typedef struct gh_align_candidate {
  float matrix[9];
  float dx, dy;                    /* correction added to the seed's translation */
  int64_t valid_pixels, sampled_pixels;
  int32_t sample_stride;
  double rms, mean_abs_diff, ncc;  /* NaN when undefined */
  int64_t coarse_candidates, fine_candidates;
} gh_align_candidate;

int gh_refine_matrix_translation_candidates(gh_stack* stack, int64_t reference_index, int64_t moving_index,
                                            const float* matrices, int64_t seed_count,
                                            const gh_align_search_params* params, gh_align_candidate* seeds,
                                            int64_t* selected_index);

/* refine_matrix_translation_candidates_batch_to_reference
 * (native_bindings.cpp:10186-10416; kernel cpp/cuda/warp_kernels.cu:1198):
 * moving frame moving_indices[k] searched from seed matrices[k] as above;
 * out[k] its best.  The coarse candidates of every frame share one launch, as
 * do the fine ones.  Synchronous. */
int gh_refine_matrix_translation_candidates_batch(gh_stack* stack, int64_t reference_index,
                                                  const int64_t* moving_indices, int64_t count, const float* matrices,
                                                  const gh_align_search_params* params, gh_align_candidate* out);

/* frame_global_stats(i) (native_bindings.cpp:10569-10643) plus exact median /
 * MAD (src/glass/engine/quality.py:126-200).  FP64. */
int gh_frame_stats(gh_stack* stack, int64_t index, int32_t exact_median, gh_frame_statistics* out);

/* frame_global_stats (native_bindings.cpp:10569-10643) of many frames in one
 * pass: out[k] for frame indices[k], bit-identical to gh_frame_stats with
 * exact_median = 0 (median and mad are 0).  The reference asks for these one
 * frame at a time in its local-normalisation and weighting loops
 * (resident_cuda.py:12928-12996); this is their batched form.  Synchronous. */
int gh_frames_stats(gh_stack* stack, const int64_t* indices, int64_t count, gh_frame_statistics* out);

/* frame_pair_grid_stats(reference, source, tile_height, tile_width) and
 * frame_pair_grid_stats_batch(reference, sources, tile_height, tile_width)
 * (native_bindings.cpp:10645-10784, :10786-11021; kernel
 * cpp/cuda/local_norm_kernels.cu:220-304): per tile of the
 * ceil(H / tile_height) x ceil(W / tile_width) grid (the last row / column of
 * tiles ends at the frame edge), over the pixels finite in both the source and
 * the reference frame, FP64 sums turned into float32 mean and standard
 * deviation of each frame (variance clamped at 0; all 0 for an empty tile) and
 * the paired pixel count.  Outputs are host arrays [count][grid_rows][grid_cols];
 * any may be NULL.  Synchronous. */
int gh_pair_grid_stats(gh_stack* stack, int64_t reference_index, const int64_t* source_indices, int64_t count,
                       int32_t tile_height, int32_t tile_width, float* source_mean, float* source_std,
                       float* reference_mean, float* reference_std, uint64_t* valid_pixels);

/* apply_global_normalization_frame(index, scale, offset) and
 * apply_grid_normalization_frame[s](indices, scales, offsets, tile_height,
 * tile_width) (native_bindings.cpp:11023-11052, :11054-11151, :11153-11290;
 * kernels cpp/cuda/local_norm_kernels.cu:6-71): frame indices[k] becomes, in
 * place, x * scale + offset (product and sum rounded separately) with the
 * coefficients of tile (min(y / tile_height, rows - 1), min(x / tile_width,
 * cols - 1)) of grid k, for the grid of gh_pair_grid_stats.  The global model is
 * tile_height = H, tile_width = W (one tile).  scales / offsets: host or device
 * float32 [count][grid_rows][grid_cols].  A frame named twice is normalised
 * twice, in order.  Host coefficients, or a NULL stream, make the call
 * synchronous; device coefficients on a caller stream are applied
 * asynchronously on that stream. */
int gh_apply_normalization(gh_stack* stack, const int64_t* indices, int64_t count, const float* scales,
                           const float* offsets, int32_t tile_height, int32_t tile_width, void* stream);

/* Fused calibration statistics of the last gh_calibrate_frames call with
 * accumulate_stats=1: sum, sum_sq (FP64) and finite count per frame index. */
int gh_calibration_stats(gh_stack* stack, int64_t index, double* sum, double* sum_sq, int64_t* finite);

/* Wait for all work queued on the stack's stream. */
int gh_synchronize(gh_stack* stack);

/* ------------------------------------------------ star detection / matching */
#define GH_MODEL_TRANSLATION 0
#define GH_MODEL_SIMILARITY 1
#define GH_MODEL_AFFINE 2

/* One detected star (src/glass/cpu/star_detect.py:8-18 Star). */
typedef struct gh_star {
  double x, y, flux, peak, fwhm_px, eccentricity, snr, sharpness;
  int64_t npix;
  int32_t has_fwhm, has_eccentricity; /* 0: the reference field is None */
} gh_star;

/* detect_stars(data, threshold_sigma, max_stars, background, noise,
 * window_radius, min_separation_px, origin_x, origin_y)
 * (src/glass/cpu/star_detect.py:121-168, robust_background :28-42) on resident
 * frame `index`: float32 median / MAD background (NaN `background` / `noise`
 * selects the robust estimate), GPU 3x3 local-maximum candidates, FP64 window
 * moments.  Up to `capacity` stars go to `out` (brightest first); *count gets
 * the list length.  background_out (nullable) receives {median, mad, noise}.
 * window_radius must be in [1, 8]. */
int gh_detect_stars(gh_stack* stack, int64_t index, double threshold_sigma, int64_t max_stars, double background,
                    double noise, int32_t window_radius, double min_separation_px, int64_t origin_x,
                    int64_t origin_y, gh_star* out, int64_t capacity, int64_t* count, double* background_out);

/* estimate_star_transform(reference, moving, transform_model, min_inliers,
 * tolerance_px, max_stars) (src/glass/cpu/registration.py:476-566) on star
 * positions ref_xy[2*n_ref], mov_xy[2*n_mov] (brightest first), FP64, host
 * only.  model: GH_MODEL_*.  matrix[9] (row-major, moving -> reference);
 * pairs[2 * max_stars] receives (moving, reference) index pairs, *n_pairs the
 * inlier count; *rms the pair RMS (NaN when none); *ok = inliers >= min_inliers. */
int gh_estimate_star_transform(const double* ref_xy, int64_t n_ref, const double* mov_xy, int64_t n_mov,
                               int32_t model, int64_t min_inliers, double tolerance_px, int64_t max_stars,
                               double* matrix, int64_t* pairs, int64_t* n_pairs, double* rms, int32_t* ok);

/* ------------------------------------- triangle-asterism similarity (host arrays)
 * The resident default registration route similarity_cuda_triangle: local
 * triangle descriptors of compact star catalogues, similarity transforms from
 * matched descriptors, and the two-star-pair catalogue seed matcher
 * (src/glass_cuda.py:1726-2140; cpp/src/native_bindings.cpp:17044-18678;
 * cpp/cuda/registration_kernels.cu:746-1821).  Inputs and outputs are host
 * memory; each call synchronises before it returns.  Calls on one process
 * serialise on a per-device workspace. */

typedef struct gh_triangle_info {
  int64_t count;     /* descriptors written (<= max_descriptors) */
  int64_t raw_count; /* anchors x C(neighbors, 3) */
  int32_t neighbors; /* neighbour count used: clamp(min(stars, neighbors), 3, 16) */
  int32_t reserved;
} gh_triangle_info;

/* triangle_asterism_descriptors_f32(x, y, max_stars, neighbors,
 * max_descriptors) (glass_cuda.py:1835-1901, native_bindings.cpp:17044-17253):
 * the first min(n, max_stars) stars; every star and its nearest neighbours
 * give C(neighbors, 3) triangles; triangles with area < 2 px^2 or a side
 * <= 1e-6 drop out; one descriptor per star triple (the first in anchor /
 * combination order), ordered by area descending then by the sorted triple.
 * descriptors[2 * max_descriptors] (longest/middle, middle/shortest),
 * indices[3 * max_descriptors] (longest side's far-from-third vertex, the
 * other end, the third vertex), areas[max_descriptors]. */
int gh_triangle_descriptors(const float* x, const float* y, int64_t n, int32_t max_stars, int32_t neighbors,
                            int32_t max_descriptors, float* descriptors, int32_t* indices, float* areas,
                            gh_triangle_info* info);

/* triangle_asterism_descriptors_batch_f32 (glass_cuda.py:1935-1986,
 * native_bindings.cpp:17255-17595): `batch` catalogues in one launch; outputs
 * of catalogue b at descriptors + 2 * b * max_descriptors, indices + 3 * b *
 * max_descriptors, areas + b * max_descriptors, info[b]. */
int gh_triangle_descriptors_batch(const float* const* x, const float* const* y, const int64_t* n, int64_t batch,
                                  int32_t max_stars, int32_t neighbors, int32_t max_descriptors, float* descriptors,
                                  int32_t* indices, float* areas, gh_triangle_info* info);

/* One similarity estimate: matrix row-major moving -> reference
 * [[a, -b, tx], [b, a, ty], [0, 0, 1]] (identity when nothing matched). */
typedef struct gh_similarity_fit {
  float matrix[9];
  float scale;
  float rotation_rad;
  float rms_px;                 /* NaN when nothing matched */
  int32_t inliers;              /* mutual nearest-neighbour matches of the best seed */
  int64_t best_candidate_index; /* -1 when nothing matched */
  int64_t candidate_count;      /* seeds enumerated */
  int64_t scored_candidates;    /* seeds that passed the descriptor / pair gates */
  int32_t refined_inliers;      /* catalogue estimate: one-sided matches of the refit */
  int32_t refit_status;         /* catalogue estimate: 0 ok, 1 < 2 matches, 2 degenerate,
                                   3 rejected (seed kept); -1 not run */
  float refit_rms_px;
  int32_t reserved;
} gh_similarity_fit;

/* estimate_similarity_from_triangle_descriptors_f32(reference_x, reference_y,
 * moving_x, moving_y, reference_descriptors, reference_indices,
 * moving_descriptors, moving_indices, tolerance_px, descriptor_radius)
 * (glass_cuda.py:2050-2140, native_bindings.cpp:17597-17848): every
 * (reference, moving) descriptor pair within descriptor_radius, with both
 * orders of the reference triangle's first two vertices, seeds an FP64
 * three-point similarity fit scored by mutual nearest neighbours within
 * tolerance_px; the best seed has most inliers, then lowest RMS, then lowest
 * candidate index (rd * n_mov_desc + md) * 2 + order. */
int gh_similarity_from_triangles(const float* rx, const float* ry, int64_t nr, const float* mx, const float* my,
                                 int64_t nm, const float* rdesc, const int32_t* ridx, int64_t nrd, const float* mdesc,
                                 const int32_t* midx, int64_t nmd, float tolerance_px, float descriptor_radius,
                                 gh_similarity_fit* out);

/* estimate_similarity_from_triangle_descriptors_batch_f32 (glass_cuda.py:
 * 2143-2240, native_bindings.cpp:17850-18252): one shared reference catalogue
 * and descriptor set, `batch` moving catalogues; out[b] equals the single
 * call on moving catalogue b. */
int gh_similarity_from_triangles_batch(const float* rx, const float* ry, int64_t nr, const float* rdesc,
                                       const int32_t* ridx, int64_t nrd, int64_t batch, const float* const* mx,
                                       const float* const* my, const int64_t* nm, const float* const* mdesc,
                                       const int32_t* const* midx, const int64_t* nmd, float tolerance_px,
                                       float descriptor_radius, gh_similarity_fit* out);

/* Parameters of gh_similarity_from_catalogs with the binding's conventions
 * (native_bindings.cpp:18283-18302): prior_radius_px < 0 disables the prior
 * (prior_dx / prior_dy then read 0), max_abs_rotation_rad < 0 disables the
 * rotation gate, max_scale FLT_MAX is unbounded. */
typedef struct gh_catalog_similarity_params {
  float tolerance_px;
  float min_pair_distance;
  float prior_dx, prior_dy, prior_radius_px;
  float min_scale, max_scale;
  float max_abs_rotation_rad;
  int32_t top_k;
} gh_catalog_similarity_params;

typedef struct gh_similarity_candidate {
  int64_t candidate_index;
  int32_t inliers;
  float rms_px;
  float a, b, tx, ty; /* matrix [[a, -b, tx], [b, a, ty], [0, 0, 1]] */
} gh_similarity_candidate;

/* estimate_similarity_from_catalogs_f32 (glass_cuda.py:1726-1824,
 * native_bindings.cpp:18254-18678): every ordered reference star pair against
 * every ordered moving star pair seeds a similarity (scale / rotation / prior
 * gates), scored like the triangle seeds; the best seed is refitted in FP64
 * over its one-sided matches and the refit replaces it unless it has fewer
 * matches or a higher RMS.  top[top_k] receives the diversified top-k seed
 * list (per inlier count the ceil(top_k / 4) lowest-RMS seeds from the highest
 * count down, then the overall best), *top_count its length.  Requires
 * nr (nr - 1) nm (nm - 1) < 2^31. */
int gh_similarity_from_catalogs(const float* rx, const float* ry, int64_t nr, const float* mx, const float* my,
                                int64_t nm, const gh_catalog_similarity_params* params, gh_similarity_fit* out,
                                gh_similarity_candidate* top, int32_t* top_count);

/* --------------------------------------- affine Lanczos3 adjoint / refinement
 * No counterpart in the reference: the warp of src/glass/engine/warp.py:121-228
 * (CPU-path Lanczos3, normalised weights) for an affine source map
 * s = B (x, y, 1), B = map6 row-major 2x3 (reference pixel -> source pixel),
 * and its adjoint with respect to the pixels and the 6 map parameters, for
 * photometric registration refinement (the analogue of the reference's
 * refine_matrix_translation_candidates_*, native_bindings.cpp:10079-10416).
 * Images are device f32 [h*w]; nothing is allocated inside these calls. */

/* Device doubles of workspace gh_affine_lanczos3_backward needs for h x w. */
int gh_affine_workspace_size(int64_t height, int64_t width, int64_t* doubles);

/* out = warp(src; map6), `fill` outside the footprint 2 <= s < dim-3. */
int gh_affine_lanczos3_forward(const float* src, int64_t height, int64_t width, const double* map6, float fill,
                               float* out, void* stream);

/* grad_src (device f32, nullable) += adjoint of grad_out; grad_map6 (device f64
 * [6], nullable) = d <grad_out, warp(src; map6)> / d map6.  workspace: device
 * f64 of gh_affine_workspace_size doubles. */
int gh_affine_lanczos3_backward(const float* src, int64_t height, int64_t width, const double* map6,
                                const float* grad_out, float* grad_src, double* grad_map6, double* workspace,
                                int64_t workspace_doubles, void* stream);

/* Gauss-Newton refinement of the moving->reference affine matrix of resident
 * frame `moving_index` against frame `reference_index`: minimises
 * sum (warp(moving) - reference)^2 over the pixels in the footprint.
 * matrix_in / matrix_out: 3x3 row-major (host); stops when the update moves
 * every image corner by less than tolerance_px or after max_iterations.
 * report (host, nullable) [4]: iterations, final RMS residual, pixels used,
 * last update (px). */
int gh_refine_affine(gh_stack* stack, int64_t reference_index, int64_t moving_index, const double* matrix_in,
                     int32_t max_iterations, double tolerance_px, double* matrix_out, double* report);

/* ------------------------------------------------------ batched multi-job
 * No counterpart in the reference (which runs one resident job per call):
 * many stacking jobs -- each a loaded gh_stack, the frames to warp with their
 * FP64 inverses, frame weights and integration parameters/outputs -- run as
 * one batched warp launch over all jobs' (frame group, tile) items, one
 * launch that decides each job's statistics arithmetic on the device from the
 * warped frames' non-finite flags, and one integration launch per
 * sort-network capacity class over the pixels of all jobs of the class (a
 * device job table; rejection-free jobs and jobs of more than 256
 * positive-weight frames launch on their own), optionally captured once into
 * a hipGraph and replayed.  Results equal the per-job calls (gh_warp_frames +
 * gh_integrate with statistics_variant -1, tile_size 0) whatever the
 * schedule.  The batch owns the warped copies of the listed frames; the
 * stacks' frames are read, not modified. */
typedef struct gh_batch gh_batch;

typedef struct gh_batch_job {
  gh_stack* stack;
  const int64_t* warp_indices;   /* host [n_warp]; other frames are used as loaded */
  const double* inverse;         /* host [n_warp * 9], FP64 inverses (last row 0 0 1) */
  int64_t n_warp;
  int32_t interpolation;         /* GH_INTERP_BILINEAR or GH_INTERP_LANCZOS3 */
  uint32_t warp_flags;           /* GH_WARP_SKIP_NONFINITE_TAPS (resident semantics) */
  float fill;
  float clamping_threshold;      /* < 0: off (Lanczos3 resident semantics only) */
  const float* weights;          /* host [frame_count] or NULL (unit) */
  gh_integrate_params params;    /* tile_size must be 0 */
  gh_integrate_outputs outputs;  /* device pointers; dq_use_geometric uses the
                                    batch's coverage (unwarped frames count 1) */
} gh_batch_job;

#define GH_BATCH_USE_GRAPH 1u  /* capture the launch sequence once, replay it */
#define GH_BATCH_AUTOTUNE 2u   /* time candidate schedules on the first run (warp grid, item
                                  order and frames per item; integration capacity classes) */

int gh_batch_create(const gh_batch_job* jobs, int64_t n_jobs, uint32_t flags, gh_batch** out);
/* Enqueue one execution on `stream`, asynchronously; NULL runs it on the
   batch's own stream and returns when it has finished. */
int gh_batch_run(gh_batch* batch, void* stream);
/* out[10]: jobs, warp work items (both warp launches), integration launches,
 * graph captured (0/1), Lanczos3 warp grid, Lanczos3 warp item stride, jobs
 * whose statistics arithmetic is decided
 * on the device, autotune trials run, frames per warp work item, integration
 * capacity class set (0 each job at its smallest register capacity, 1 powers
 * of two, 2 one launch at the largest). */
int gh_batch_info(gh_batch* batch, int64_t* out);
/* Fixes the schedule (no autotuning; the next run captures the graph anew):
 * warp_grid workgroups in [1, 2^31 - 1] and warp_item_stride 1 or coprime to
 * the Lanczos3 jobs' work-item count (their launch; the bilinear jobs' items
 * run one workgroup each, in order), frames_per_item 2, 4 or 8,
 * integrate_classes 0, 1 or 2 (as in gh_batch_info).  Results do not depend
 * on it. */
int gh_batch_set_schedule(gh_batch* batch, int64_t warp_grid, int32_t warp_item_stride, int32_t frames_per_item,
                          int32_t integrate_classes);
int gh_batch_destroy(gh_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* GLASS_HIP_H */
