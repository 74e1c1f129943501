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

// One translated candidate: the moving frame it samples, its matrix and the
// correction added to the seed's translation.
struct Cand {
  const float* moving;
  Mat matrix;
  float dx, dy;
};

void add_candidates(std::vector<Cand>& out, const float* moving, const Mat& base,
                    const std::vector<std::pair<float, float>>& offsets) {
  for (const auto& o : offsets) {
    Cand c{moving, base, o.first, o.second};
    c.matrix[2] += o.first;
    c.matrix[5] += o.second;
    out.push_back(c);
  }
}

// Metrics of every candidate at one sample stride, one launch
// (score_matrix_translation_candidates_f32, :1577-1700, scores one seed's
// candidates per launch).
std::vector<gh_align_candidate> evaluate(gh_stack* s, const float* ref, const std::vector<Cand>& cands, int stride) {
  const int w = static_cast<int>(s->width), h = static_cast<int>(s->height);
  stride = std::max(stride, 1);
  const int sw = (w + stride - 1) / stride, sh = (h + stride - 1) / stride;
  const int n = static_cast<int>(cands.size());
  std::vector<float> inv(static_cast<size_t>(n) * 9);
  std::vector<const float*> frames(n);
  for (int i = 0; i < n; ++i) {
    const Mat v = invert(cands[i].matrix);
    std::copy(v.begin(), v.end(), inv.begin() + static_cast<size_t>(i) * 9);
    frames[i] = cands[i].moving;
  }
  // enough (candidate, row chunk) blocks to fill the chip
  const int want = std::max(1, std::min(sh, (2048 + n - 1) / n));
  const int rows = (sh + want - 1) / want;
  const int chunks = (sh + rows - 1) / rows;
  hipStream_t st = s->stream;
  s->align_inv.ensure(inv.size(), "hipMalloc(alignment inverses)");
  s->align_frames.ensure(frames.size(), "hipMalloc(alignment frame table)");
  s->align_partials.ensure(static_cast<size_t>(n) * chunks * 8, "hipMalloc(alignment partials)");
  check(hipMemcpyAsync(s->align_inv.ptr, inv.data(), inv.size() * sizeof(float), hipMemcpyHostToDevice, st),
        "hipMemcpy(alignment inverses)");
  check(hipMemcpyAsync(s->align_frames.ptr, frames.data(), frames.size() * sizeof(const float*),
                       hipMemcpyHostToDevice, st),
        "hipMemcpy(alignment frame table)");
  gh::AlignArgs a{ref, s->align_frames.ptr, s->align_inv.ptr, w, h, stride, sw, sh, rows, chunks, n};
  gh::launch_alignment_partials(a, s->align_partials.ptr, st);
  std::vector<double> part(static_cast<size_t>(n) * chunks * 8);
  check(hipMemcpyAsync(part.data(), s->align_partials.ptr, part.size() * sizeof(double), hipMemcpyDeviceToHost, st),
        "hipMemcpy(alignment partials)");
  check(hipStreamSynchronize(st), "alignment metrics");
  std::vector<gh_align_candidate> out(n);
  for (int i = 0; i < n; ++i) {
    double sums[8] = {};
    for (int k = 0; k < chunks; ++k)
      for (int j = 0; j < 8; ++j) sums[j] += part[(static_cast<size_t>(i) * chunks + k) * 8 + j];
    gh_align_candidate& c = out[i];
    c = gh_align_candidate{};
    std::copy(cands[i].matrix.begin(), cands[i].matrix.end(), c.matrix);
    c.dx = cands[i].dx;
    c.dy = cands[i].dy;
    c.sampled_pixels = static_cast<int64_t>(sw) * sh;
    c.sample_stride = stride;
    metrics(sums, c);
  }
  return out;
}

// The best of out[first, first + count) (first of equals).
gh_align_candidate best_of(const std::vector<gh_align_candidate>& out, size_t first, size_t count) {
  gh_align_candidate best = out[first];
  for (size_t i = first + 1; i < first + count; ++i)
    if (better(out[i], best)) best = out[i];
  return best;
}

void validate(const gh_align_search_params& p) {
  if (p.search_radius_px < 0.0f || p.fine_radius_px < 0.0f) throw gh::ValueError("search radii must be non-negative");
  if (!(p.coarse_step_px > 0.0f) || !(p.fine_step_px > 0.0f)) throw gh::ValueError("search steps must be positive");
  if (p.coarse_sample_stride <= 0 || p.final_sample_stride <= 0)
    throw gh::ValueError("sample strides must be positive");
}

// Coarse grid for every job (one launch), then the fine grid (or the final
// stride) around each job's coarse best (one launch); job k = (moving frame,
// seed matrix).  Results in seeds[k].
void search(gh_stack* s, const float* ref, const std::vector<const float*>& moving, const float* matrices,
            const gh_align_search_params& p, gh_align_candidate* seeds) {
  const auto coarse = translation_offsets(0.0f, 0.0f, p.search_radius_px, p.coarse_step_px);
  const size_t jobs = moving.size();
  std::vector<Mat> base(jobs);
  std::vector<Cand> cands;
  for (size_t k = 0; k < jobs; ++k) {
    std::copy(matrices + 9 * k, matrices + 9 * k + 9, base[k].begin());
    add_candidates(cands, moving[k], base[k], coarse);
  }
  const auto first = evaluate(s, ref, cands, p.coarse_sample_stride);
  std::vector<gh_align_candidate> best(jobs);
  for (size_t k = 0; k < jobs; ++k) best[k] = best_of(first, k * coarse.size(), coarse.size());
  std::vector<size_t> start(jobs + 1, 0);
  bool second = true;
  cands.clear();
  for (size_t k = 0; k < jobs; ++k) {
    if (p.fine_radius_px > 0.0f) {
      add_candidates(cands, moving[k], base[k], translation_offsets(best[k].dx, best[k].dy, p.fine_radius_px,
                                                                     p.fine_step_px));
    } else if (p.final_sample_stride != p.coarse_sample_stride) {
      add_candidates(cands, moving[k], base[k], {{best[k].dx, best[k].dy}});
    } else {
      second = false;
    }
    start[k + 1] = cands.size();
  }
  if (second) {
    const auto fine = evaluate(s, ref, cands, p.final_sample_stride);
    for (size_t k = 0; k < jobs; ++k) best[k] = best_of(fine, start[k], start[k + 1] - start[k]);
  }
  for (size_t k = 0; k < jobs; ++k) {
    best[k].coarse_candidates = static_cast<int64_t>(coarse.size());
    best[k].fine_candidates = second ? static_cast<int64_t>(start[k + 1] - start[k]) : 0;
    seeds[k] = best[k];
  }
}

}  // namespace

extern "C" {

int gh_refine_matrix_translation_candidates(gh_stack* s, int64_t reference_index, int64_t moving_index,
                                            const float* matrices, int64_t seed_count,
                                            const gh_align_search_params* params, gh_align_candidate* seeds,
                                            int64_t* selected_index) {
  return guarded([&] {
    if (!params || !seeds || !selected_index) throw gh::ValueError("parameters or result pointer is NULL");
    s->require_loaded(reference_index, "resident matrix multi-refine");
    s->require_loaded(moving_index, "resident matrix multi-refine");
    validate(*params);
    if (!matrices || seed_count <= 0) throw gh::ValueError("matrices must contain at least one matrix");
    if (s->npix >= (int64_t(1) << 31)) throw gh::ValueError("frames of 2^31 pixels or more are not supported");
    const std::vector<const float*> moving(static_cast<size_t>(seed_count), s->slot[moving_index]);
    search(s, s->slot[reference_index], moving, matrices, *params, seeds);
    int64_t selected = 0;
    for (int64_t k = 1; k < seed_count; ++k)
      if (better(seeds[k], seeds[selected])) selected = k;
    *selected_index = selected;
  });
}

int gh_refine_matrix_translation_candidates_batch(gh_stack* s, int64_t reference_index, const int64_t* moving_indices,
                                                  int64_t count, const float* matrices,
                                                  const gh_align_search_params* params, gh_align_candidate* out) {
  return guarded([&] {
    if (!params) throw gh::ValueError("parameters pointer is NULL");
    s->require_loaded(reference_index, "resident matrix batch refine");
    validate(*params);
    if (count <= 0) return;
    if (!moving_indices || !matrices || !out) throw gh::ValueError("batch refine requires one seed matrix per moving frame");
    if (s->npix >= (int64_t(1) << 31)) throw gh::ValueError("frames of 2^31 pixels or more are not supported");
    std::vector<const float*> moving(static_cast<size_t>(count));
    for (int64_t k = 0; k < count; ++k) {
      s->require_loaded(moving_indices[k], "resident matrix batch refine");
      moving[k] = s->slot[moving_indices[k]];
    }
    search(s, s->slot[reference_index], moving, matrices, *params, out);
  });
}

}  // extern "C"
