// Star detection (GPU candidate pass + host FP64 measurement) and star-list
// transform estimation (host FP64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <vector>

namespace gh {

constexpr int kModelTranslation = 0;
constexpr int kModelSimilarity = 1;
constexpr int kModelAffine = 2;
constexpr int kMaxStarRadius = 8;
constexpr int kMaxStarWindow = (2 * kMaxStarRadius + 1) * (2 * kMaxStarRadius + 1);

struct Star {
  double x = 0, y = 0, flux = 0, peak = 0, fwhm_px = 0, eccentricity = 0, snr = 0, sharpness = 0;
  int64_t npix = 0;
  bool has_fwhm = false, has_eccentricity = false;
};

// One candidate pixel (x, y) and its (2r+1)^2 window, row-major; window cells
// outside the frame hold anything (the measurement clips the window).
struct CandidatePatch {
  int x = 0, y = 0;
  const float* values = nullptr;
};

struct TransformResult {
  double matrix[9];
  std::vector<std::pair<int64_t, int64_t>> pairs;
  double rms = INFINITY;
  bool ok = false;
};

double pairwise_sum(const double* a, size_t n);
float pairwise_sum_f32(const float* a, size_t n);
float numpy_std_f32(const float* finite, size_t n);

bool measure_candidate(const CandidatePatch& c, int height, int width, int radius, double background, double noise,
                       int64_t origin_x, int64_t origin_y, Star& out);
std::vector<Star> finish_star_list(std::vector<Star> stars, double min_separation_px, int64_t max_stars);

TransformResult estimate_star_transform(const double* ref_xy, int64_t n_ref, const double* mov_xy, int64_t n_mov,
                                        int model, int64_t min_inliers, double tolerance_px, int64_t max_stars);

// GPU: append every core pixel that is finite, above `threshold` (float32)
// and >= its 8 neighbours to cand[] (flat pixel index; order arbitrary),
// counting all of them in *count even past `capacity`.
void launch_star_candidates(const float* frame, int height, int width, float threshold, unsigned* cand,
                            unsigned capacity, unsigned* count, hipStream_t stream);
// GPU: windows[k] = the (2r+1)^2 window around cand[k], k < n.
void launch_gather_windows(const float* frame, int height, int width, const unsigned* cand, unsigned n, int radius,
                           float* windows, hipStream_t stream);

}  // namespace gh
