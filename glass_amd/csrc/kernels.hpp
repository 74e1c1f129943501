// Host-side launchers of the gfx950 kernels (one .hip translation unit each).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gh {

// ---------------------------------------------------------------- calibrate
constexpr int kCalibBatch = 16;

struct CalibBatch {
  const void* src[kCalibBatch];  // raw frame k
  float* dst[kCalibBatch];       // destination slot of frame k
  float dark_scale[kCalibBatch];
  unsigned* nonfinite[kCalibBatch];  // frame k's "holds a non-finite sample" flag (pre-cleared)
  int frame[kCalibBatch];            // frame index of batch entry k (statistics row)
};

struct CalibParams {
  const float* bias;
  const float* dark;
  const float* flat;
  int has_bias, has_dark, has_flat;
  int dark_includes_bias;
  float flat_floor;
  float pedestal;
  int format;  // GH_FMT_*
  int accumulate_stats;
};

// Per-block partial statistics, reduced in fixed order by frame_stats_reduce.
struct StatPartial {
  double sum;
  double sum_sq;
  unsigned long long finite;
  unsigned long long nonfinite;
};

// Blocks per frame used by the calibrate kernels for `npix` pixels.
unsigned calibrate_blocks_per_frame(long long npix);

void launch_calibrate(const CalibBatch& batch, int count, const CalibParams& p, long long npix,
                      StatPartial* partials, hipStream_t stream);

// Reduces rows [0, rows) of `partials` (blocks_per_row each) into
// out[batch.frame[row]], in a fixed order (deterministic FP64 sums).
void launch_stats_reduce(const StatPartial* partials, const CalibBatch& batch, int rows, unsigned blocks_per_row,
                         StatPartial* out, hipStream_t stream);

// Clears flags[i] for the listed frames (kernel, so it is graph-capturable).
void launch_clear_flags(unsigned* flags, const CalibBatch& batch, int count, hipStream_t stream);

// Per-frame sum / sum^2 / finite / nonfinite / min / max of an f32 frame.
struct FrameMoments {
  double sum;
  double sum_sq;
  unsigned long long finite;
  unsigned long long nonfinite;
  float min;
  float max;
};
void launch_frame_moments(const float* frame, long long npix, FrameMoments* scratch_partials,
                          unsigned blocks, FrameMoments* out, hipStream_t stream);
unsigned frame_moments_blocks(long long npix);

// Exact k-th smallest finite value of |x - center| (abs_mode) or x by 4-pass
// 8-bit radix select over order-preserving keys; hist must hold 256 u32.
void launch_radix_select(const float* frame, long long npix, int abs_mode, float center,
                         unsigned long long rank, unsigned* hist, unsigned* state, hipStream_t stream);

// -------------------------------------------------------------------- warp
constexpr int kWarpBatch = 8;

struct WarpBatch {
  const float* src[kWarpBatch];
  float* dst[kWarpBatch];
  double inv[kWarpBatch][9];
  unsigned* nonfinite[kWarpBatch];  // per-frame "holds a non-finite sample" flag (device)
  // [count] device snapshot of the source frames' flags taken before the
  // destination flags were cleared (launch_snapshot_warp_flags).
  const unsigned* src_flags;
};

struct WarpParams {
  int height, width;
  int interpolation;  // 1 bilinear, 3 lanczos3
  float fill;
  float clamp;        // < 0: off
  int skip_nonfinite;
  int affine;         // 1: no projective divide
  float* coverage;    // geometric coverage accumulator (nullable)
  // Normalised Lanczos3 weight table (see make_lanczos_lut); non-null selects
  // the LDS-tiled kernel, valid for affine maps of all-finite frames without clamp.
  const float4* lanczos_lut;
};

// Per-interval quadratic fits of the normalised Lanczos3 tap weights.
// Interval i holds kLanczosLutStride floats; (a, b, c) of tap k sit at
// [3k, 3k + 3), so that for f = (i + u) / intervals,
// w_k(f) ~= a + u (b + u c) with w_k(f) = L(f - k + 2) / sum_j L(f - j + 2),
// L(x) = sinc(x) sinc(x/3) (engine/warp.py:73-84), k = 0..5 for taps
// floor(s) - 2 .. floor(s) + 3.  Each quadratic interpolates w_k at u = 0,
// 1/2, 1 (exact delta weights at integer offsets); fitted in FP64 on the
// host, max error after f32 evaluation ~1e-7.
constexpr int kLanczosLutIntervals = 128;
constexpr int kLanczosLutStride = 20;  // 18 used, padded to whole float4s
// Row kLanczosLutIntervals (one past the last interval) holds the f == 1 weights.
void make_lanczos_lut(float* table /* (kLanczosLutIntervals + 1) * kLanczosLutStride floats */);

// snap[j] = *b.nonfinite[j]; *b.nonfinite[j] = 0 for j < count.
void launch_snapshot_warp_flags(const WarpBatch& b, int count, unsigned* snap, hipStream_t stream);

void launch_warp(const WarpBatch& batch, int count, const WarpParams& p, hipStream_t stream);

// Whether an affine Lanczos3 group of `count` inverse maps (up to kWarpBatch,
// launched together) takes the streamed kernel rather than the tiled one.
bool lanczos_stream_eligible(const double (*inv)[9], int count, int width, int height);

// ---- batched multi-job warp (one persistent launch over many jobs)
struct BatchWarpFrame {
  const float* src;
  float* dst;
  double inv[9];
  const unsigned* src_flag;  // the source frame's non-finite flag (read only)
  unsigned* dst_flag;        // raised when an output sample is non-finite
  int stream;                // Lanczos3: the frame's 8-frame group takes the streamed kernel
  int pad;
};
struct BatchWarpJob {
  int height, width;
  int tiles_x, tiles;        // output tiles of warp_tile_shape
  int interpolation;         // 1 bilinear, 3 lanczos3
  int skip_nonfinite;
  float fill, clamp;
  float* coverage;           // nullable, accumulated atomically
  long long first_item;      // prefix of the items of earlier jobs
  int first_frame, n_frames; // range in the frame array
  int frames_per_item;       // frames of one work item (a divisor of kWarpBatch)
  int pad;
};
// Output tile width / height of the tiled warp kernels.
void warp_tile_shape(int* tile_w, int* tile_h);
long long warp_batched_items(const BatchWarpJob& job);  // ceil(frames / frames_per_item) * tiles
unsigned warp_batched_max_blocks(bool lanczos);  // resident blocks that fill the device
// The items of the Lanczos3 jobs (lanczos) or of the bilinear jobs; item_job[i]
// is the job of work item i (jobs in first_item order).
void launch_warp_batched(bool lanczos, const BatchWarpJob* jobs, const unsigned short* item_job,
                         const BatchWarpFrame* frames, long long total_items, const float4* lut, unsigned grid,
                         int order_stride, hipStream_t stream);

void launch_fill(float* dst, long long n, float value, hipStream_t stream);
void launch_add_scalar(float* dst, long long n, float value, hipStream_t stream);
// x = (x - offset) / divisor in place (f32), flagging non-finite results.
void launch_normalize(float* frame, long long npix, float offset, float divisor, unsigned* nonfinite_flag,
                      hipStream_t stream);

// Sets *flag to 1 when the frame holds a non-finite sample (never clears it).
void launch_flag_nonfinite(const float* frame, long long npix, unsigned* flag, hipStream_t stream);

// --------------------------------------------------------------- integrate
struct IntegrateArgs {
  const float* const* frame_ptr; // [n_active] device slot pointers, frame order
  const float* frame_w;          // [n_active] positive finite weights
  int n_active;
  long long npix;
  int width;
  int rejection;                 // 0 none, 1 winsorized
  float low_sigma, high_sigma;
  int min_samples;
  float max_frac;
  int unit_weights;
  int variant_all_valid;         // statistics arithmetic when tile_flags == nullptr
  const unsigned char* tile_flags;  // per tile: 1 all-valid arithmetic (nullable)
  int tile_size, tiles_x;
  int require_positive_scale;
  // outputs (device, nullable)
  float* master;
  float* weight;
  void* coverage;
  void* low;
  void* high;
  int count_u16;
  unsigned* dq;
  const float* geometric;        // nullable
  int dq_active_frames;
  // When non-null, the statistics arithmetic is read from this device word at
  // run time (decided on the device after a batched warp, see
  // launch_decide_variants), overriding variant_all_valid.
  const int* variant_ptr;
};

// Statistics variant of one batched job: all-valid arithmetic when every
// frame weight is positive and none of its n_flags active frames holds a
// non-finite sample.
struct VariantJob {
  const unsigned* const* flags;  // device table of per-frame flag words
  int n_flags;
  int weights_positive;
};
void launch_decide_variants(const VariantJob* jobs, int n_jobs, int* out, hipStream_t stream);

// Normalised cross-correlation over a grid of candidate translations:
// integer offsets (ox, oy) in [-radius, radius], or float32 offsets
// center + (ox, oy) * step sampled bilinearly.
struct NccArgs {
  const float* reference;
  const float* moving;
  int width, height, stride;
  int radius_x, radius_y;
  float center_dx, center_dy, step;  // subpixel candidates only
};
int ncc_chunks(const NccArgs& a);  // row chunks per candidate (partials: candidates * chunks * 6 doubles)
void launch_ncc_search(const NccArgs& a, bool subpixel, double* partials, double* scores, hipStream_t stream);

// Integer or float32-bilinear translation of one frame into dst (fill
// outside); coverage (nullable) += 1 inside; *flag |= any non-finite output.
void launch_translate(const float* src, float* dst, int width, int height, bool bilinear, int idx, int idy, float fdx,
                      float fdy, float fill, float* coverage, unsigned* flag, hipStream_t stream);

// Mean / sigma-clip integration that samples the frames itself (resident
// stack as stored, or warped on the fly through FP64 inverses).
struct ClipArgs {
  const float* const* frames;  // [n] device table, frames of positive finite weight in order
  const double* inverses;      // [n][9] (samplers 1, 2)
  const float* weights;        // [n]
  int n, width, height;
  long long npix;
  float clamp;                 // Lanczos3 clamp (< 0: off)
  int affine;                  // every inverse has last row (0, 0, 1): no projective divide
  int mode;                    // 0 weighted mean, 1 sigma clip
  int winsorize;
  float low_sigma, high_sigma;
  float *master, *weight, *coverage, *low, *high, *geometric;  // nullable outputs
  const float* tile_map;          // tile-local form: per-pixel multiplier of the target entries
  const unsigned char* target;    // tile-local form: [n], nonzero = weight scaled by tile_map
};
// sampler: 0 stack as stored, 1 bilinear, 2 Lanczos3 (resident tap rules)
void launch_clip_integrate(const ClipArgs& a, int sampler, hipStream_t stream);
// Tile-local form over the stack as stored (a.tile_map and a.target set).
void launch_clip_integrate_tiles(const ClipArgs& a, hipStream_t stream);
// map[p] = 1, then multipliers[t] over tile t's half-open rectangle
// extents[4t..4t+3] = (x0, y0, x1, y1); tiles do not overlap; max_area is the
// largest tile's pixel count.
void launch_tile_map(float* map, const int* extents, const float* multipliers, int tiles, long long max_area,
                     int width, long long npix, hipStream_t stream);

// Resident local normalisation.  A batch names up to kLnBatch resident
// frames (sources for the statistics, targets for the application) and their
// non-finite flags (application only).
constexpr int kLnBatch = 64;
constexpr int kLnGroup = 4;  // sources per statistics block (one reference read)
struct LnBatch {
  const float* frame[kLnBatch];
  unsigned* flag[kLnBatch];
};
// FP64 moments of each frame of the batch (bit-identical to
// launch_frame_moments of the same frame): partials [n][blocks], out [n].
void launch_frame_moments_batch(const LnBatch& b, int n, long long npix, FrameMoments* scratch_partials,
                                unsigned blocks, FrameMoments* out, hipStream_t stream);
// Per-tile FP64 sums (s, s^2, r, r^2) and counts of (source, reference) pairs
// over pixels finite in both; batch entry j writes record (first + j).
void launch_pair_grid_sums(const LnBatch& b, int n, const float* reference, int width, int height, int tile_w,
                           int tile_h, int grid_cols, int grid_rows, double* sums, unsigned long long* counts,
                           int first, hipStream_t stream);
// In place x * scale + offset with per-tile coefficients (device grids
// [first + j][grid_rows][grid_cols] for batch entry j).
void launch_apply_grid(const LnBatch& b, int n, const float* scales, const float* offsets, int width, int height,
                       int tile_w, int tile_h, int grid_cols, int grid_rows, int first, hipStream_t stream);

// Frame-count limits of the winsorized integration.
int integrate_max_frames();         // positive-weight frames the winsorized path accepts
int integrate_register_capacity();  // frames the register kernels hold; frame tables pad to it
void launch_integrate(const IntegrateArgs& a, hipStream_t stream);

// Job table of one register-kernel launch over the pixels of many jobs: block
// b belongs to job block_job[b] (arguments jobs[job]) and covers its pixels
// [256 (b - first_block[job]), +256).  All null: one job, the kernel argument.
struct IntegrateJobTable {
  const IntegrateArgs* jobs;
  const unsigned short* block_job;
  const unsigned* first_block;
};
// Capacities of the register-sort winsorized kernels (samples held in VGPRs);
// a job takes the smallest one holding its positive-weight frames.
constexpr int kIntegrateCaps[] = {16, 24, 32, 40, 48, 56, 64, 72, 80, 88, 96, 104, 112, 120, 128, 160, 200, 256};
constexpr int kIntegrateCapCount = sizeof(kIntegrateCaps) / sizeof(kIntegrateCaps[0]);
int integrate_cap_for(int n_active);  // smallest capacity >= n_active (n_active <= 256)
// The register kernel of capacity `cap` (one of kIntegrateCaps) over `grid`
// blocks of 256 pixels, for one job (empty table) or the jobs of the table.
void launch_integrate_cap(int cap, const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid,
                          hipStream_t stream);
// Per capacity set (one translation unit each, compiled in parallel): false
// when `cap` is not in the set.
bool launch_integrate_caps_small(int cap, const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid,
                                 hipStream_t stream);
bool launch_integrate_caps_mid(int cap, const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid,
                               hipStream_t stream);
bool launch_integrate_caps_upper(int cap, const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid,
                                 hipStream_t stream);
bool launch_integrate_caps_large(int cap, const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid,
                                 hipStream_t stream);

// Per-tile "any non-finite sample" flags over n_active frames (for the
// CPUStackEngine per-tile statistics choice); flags[t] = 1 if all finite.
void launch_tile_all_finite(const float* const* frame_ptr, int n_active, int height, int width, int tile_size,
                            unsigned char* flags, hipStream_t stream);

}  // namespace gh
