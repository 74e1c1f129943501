// C ABI of refine_matrix_translation_candidates_to_reference
// (native_bindings.cpp:10079-10184, helpers :1459-1700): for each seed
// matrix, a coarse grid of translation corrections and a fine grid around
// the coarse best, scored by the bilinear alignment metrics of the moving
// frame against the reference; the best seed wins.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

#include "candidates.hpp"
#include "stack.hpp"

using ghi::guarded;

namespace {

using Mat = std::array<float, 9>;

// translation_offsets (native_bindings.cpp:1489-1506): float32 loops, dx outer.
std::vector<std::pair<float, float>> translation_offsets(float cx, float cy, float radius, float step) {
  std::vector<std::pair<float, float>> out;
  for (float dx = cx - radius; dx <= cx + radius + step * 0.5f; dx += step)
    for (float dy = cy - radius; dy <= cy + radius + step * 0.5f; dy += step) out.emplace_back(dx, dy);
  if (out.empty()) out.emplace_back(cx, cy);
  return out;
}

// invert_matrix3x3 (native_bindings.cpp:1422-1457): FP64 adjugate, float32 result.
Mat invert(const Mat& m) {
  const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
  const double c00 = e * i - f * h, c01 = -(d * i - f * g), c02 = d * h - e * g;
  const double c10 = -(b * i - c * h), c11 = a * i - c * g, c12 = -(a * h - b * g);
  const double c20 = b * f - c * e, c21 = -(a * f - c * d), c22 = a * e - b * d;
  const double det = a * c00 + b * c01 + c * c02;
  if (std::abs(det) <= 1.0e-20) throw gh::ValueError("matrix must be invertible");
  const double k = 1.0 / det;
  return {static_cast<float>(c00 * k), static_cast<float>(c10 * k), static_cast<float>(c20 * k),
          static_cast<float>(c01 * k), static_cast<float>(c11 * k), static_cast<float>(c21 * k),
          static_cast<float>(c02 * k), static_cast<float>(c12 * k), static_cast<float>(c22 * k)};
}

// metrics_from_sums (:1507-1544) and better_matrix_metric (:1545-1566).
void metrics(const double* s, gh_align_candidate& out) {
  const uint64_t n = static_cast<uint64_t>(s[7]);
  out.valid_pixels = static_cast<int64_t>(n);
  out.rms = out.mean_abs_diff = out.ncc = NAN;
  if (n == 0) return;
  const double count = static_cast<double>(n);
  out.rms = std::sqrt(s[5] / count);
  out.mean_abs_diff = s[6] / count;
  if (n > 1) {
    const double num = s[4] - (s[0] * s[1] / count);
    const double rv = std::max(s[2] - (s[0] * s[0] / count), 0.0);
    const double mv = std::max(s[3] - (s[1] * s[1] / count), 0.0);
    const double den = std::sqrt(rv * mv);
    if (den > 0.0) out.ncc = num / den;
  }
}

bool better(const gh_align_candidate& c, const gh_align_candidate& b) {
  const bool cr = std::isfinite(c.rms), br = std::isfinite(b.rms);
  if (cr != br) return cr;
  if (cr && std::abs(c.rms - b.rms) > 1.0e-12) return c.rms < b.rms;
  const bool cn = std::isfinite(c.ncc), bn = std::isfinite(b.ncc);
  if (cn != bn) return cn;
  if (cn && std::abs(c.ncc - b.ncc) > 1.0e-12) return c.ncc > b.ncc;
  return c.valid_pixels > b.valid_pixels;
}

// score_matrix_translation_candidates_f32 (:1577-1700): the best of the
// seed's translated candidates at one sample stride.
gh_align_candidate score(gh_stack* s, const float* ref, const float* mov, const Mat& base,
                         const std::vector<std::pair<float, float>>& offsets, int stride) {
  const int w = static_cast<int>(s->width), h = static_cast<int>(s->height);
  stride = std::max(stride, 1);
  const int sw = (w + stride - 1) / stride, sh = (h + stride - 1) / stride;
  const int n = static_cast<int>(offsets.size());
  std::vector<Mat> mats(n);
  std::vector<float> inv(static_cast<size_t>(n) * 9);
  for (int i = 0; i < n; ++i) {
    mats[i] = base;
    mats[i][2] += offsets[i].first;
    mats[i][5] += offsets[i].second;
    const Mat v = invert(mats[i]);
    std::copy(v.begin(), v.end(), inv.begin() + static_cast<size_t>(i) * 9);
  }
  // enough (candidate, row chunk) blocks to fill the chip
  const int want = std::max(1, std::min(sh, (2048 + n - 1) / n));
  const int rows = (sh + want - 1) / want;
  const int chunks = (sh + rows - 1) / rows;
  hipStream_t st = s->stream;
  s->align_inv.ensure(inv.size(), "hipMalloc(alignment inverses)");
  s->align_partials.ensure(static_cast<size_t>(n) * chunks * 8, "hipMalloc(alignment partials)");
  check(hipMemcpyAsync(s->align_inv.ptr, inv.data(), inv.size() * sizeof(float), hipMemcpyHostToDevice, st),
        "hipMemcpy(alignment inverses)");
  gh::AlignArgs a{ref, mov, s->align_inv.ptr, w, h, stride, sw, sh, rows, chunks, n};
  gh::launch_alignment_partials(a, s->align_partials.ptr, st);
  std::vector<double> part(static_cast<size_t>(n) * chunks * 8);
  check(hipMemcpyAsync(part.data(), s->align_partials.ptr, part.size() * sizeof(double), hipMemcpyDeviceToHost, st),
        "hipMemcpy(alignment partials)");
  check(hipStreamSynchronize(st), "alignment metrics");
  gh_align_candidate best{};
  bool have = false;
  for (int i = 0; i < n; ++i) {
    double sums[8] = {};
    for (int k = 0; k < chunks; ++k)
      for (int j = 0; j < 8; ++j) sums[j] += part[(static_cast<size_t>(i) * chunks + k) * 8 + j];
    gh_align_candidate c{};
    std::copy(mats[i].begin(), mats[i].end(), c.matrix);
    c.dx = offsets[i].first;
    c.dy = offsets[i].second;
    c.sampled_pixels = static_cast<int64_t>(sw) * sh;
    c.sample_stride = stride;
    metrics(sums, c);
    if (!have || better(c, best)) {
      best = c;
      have = true;
    }
  }
  return best;
}

}  // namespace

extern "C" {

int gh_refine_matrix_translation_candidates(gh_stack* s, int64_t reference_index, int64_t moving_index,
                                            const float* matrices, int64_t seed_count,
                                            const gh_align_search_params* params, gh_align_candidate* seeds,
                                            int64_t* selected_index) {
  return guarded([&] {
    if (!params || !seeds || !selected_index) throw gh::ValueError("parameters or result pointer is NULL");
    const gh_align_search_params p = *params;
    s->require_loaded(reference_index, "resident matrix multi-refine");
    s->require_loaded(moving_index, "resident matrix multi-refine");
    if (p.search_radius_px < 0.0f || p.fine_radius_px < 0.0f)
      throw gh::ValueError("search radii must be non-negative");
    if (!(p.coarse_step_px > 0.0f) || !(p.fine_step_px > 0.0f))
      throw gh::ValueError("search steps must be positive");
    if (p.coarse_sample_stride <= 0 || p.final_sample_stride <= 0)
      throw gh::ValueError("sample strides must be positive");
    if (!matrices || seed_count <= 0) throw gh::ValueError("matrices must contain at least one matrix");
    if (s->npix >= (int64_t(1) << 31)) throw gh::ValueError("frames of 2^31 pixels or more are not supported");
    const auto coarse = translation_offsets(0.0f, 0.0f, p.search_radius_px, p.coarse_step_px);
    const float* ref = s->slot[reference_index];
    const float* mov = s->slot[moving_index];
    int64_t selected = -1;
    for (int64_t k = 0; k < seed_count; ++k) {
      Mat base;
      std::copy(matrices + 9 * k, matrices + 9 * k + 9, base.begin());
      const gh_align_candidate cbest = score(s, ref, mov, base, coarse, p.coarse_sample_stride);
      gh_align_candidate best = cbest;
      int64_t fine = 0;
      if (p.fine_radius_px > 0.0f) {
        const auto offs = translation_offsets(cbest.dx, cbest.dy, p.fine_radius_px, p.fine_step_px);
        fine = static_cast<int64_t>(offs.size());
        best = score(s, ref, mov, base, offs, p.final_sample_stride);
      } else if (p.final_sample_stride != p.coarse_sample_stride) {
        fine = 1;
        best = score(s, ref, mov, base, {{cbest.dx, cbest.dy}}, p.final_sample_stride);
      }
      best.coarse_candidates = static_cast<int64_t>(coarse.size());
      best.fine_candidates = fine;
      seeds[k] = best;
      if (selected < 0 || better(best, seeds[selected])) selected = k;
    }
    *selected_index = selected;
  });
}

}  // extern "C"
