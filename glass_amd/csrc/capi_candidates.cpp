// C ABI of the resident star candidate catalogues
// (native_bindings.cpp:11356-12122): star_candidates, star_top_candidates,
// star_top_nms_candidates[_centroid] and star_grid_top_nms_candidates
// [_deterministic][_centroid].

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "candidates.hpp"
#include "kernels.hpp"
#include "stack.hpp"

using ghi::guarded;

namespace {

void validate(const gh_candidate_params& p) {
  switch (p.mode) {
    case GH_CANDIDATES_ALL:
    case GH_CANDIDATES_TOP:
      if (p.max_candidates <= 0) throw gh::ValueError("max_candidates must be positive");
      break;
    case GH_CANDIDATES_TOP_NMS:
      if (p.max_candidates <= 0 || p.max_output_candidates <= 0)
        throw gh::ValueError("candidate counts must be positive");
      if (p.max_candidates < p.max_output_candidates)
        throw gh::ValueError("scan_candidates must be greater than or equal to max_output_candidates");
      if (p.min_separation_px < 0.0f) throw gh::ValueError("min_separation_px must be non-negative");
      break;
    case GH_CANDIDATES_GRID_NMS:
      if (p.grid_cols <= 0 || p.grid_rows <= 0 || p.candidates_per_cell <= 0 || p.max_output_candidates <= 0)
        throw gh::ValueError("grid dimensions and candidate counts must be positive");
      if (p.min_separation_px < 0.0f) throw gh::ValueError("min_separation_px must be non-negative");
      if (static_cast<int64_t>(p.grid_cols) * p.grid_rows > INT32_MAX)
        throw gh::ValueError("grid_cols * grid_rows must fit in int32");
      break;
    default:
      throw gh::ValueError("unknown candidate mode");
  }
}

// star_catalog_nms_kernel (star_detect_kernels.cu:607-648): greedy over the
// ranked list, float32 squared distances.
int64_t separation_filter(const std::vector<float>& x, const std::vector<float>& y, const std::vector<float>& f,
                          int64_t max_output, float min_separation, float* ox, float* oy, float* of) {
  const float sep2 = min_separation * min_separation;
  int64_t selected = 0;
  for (size_t i = 0; i < x.size() && selected < max_output; ++i) {
    if (!std::isfinite(f[i]) || f[i] <= -1.0e30f || !std::isfinite(x[i]) || !std::isfinite(y[i])) continue;
    bool keep = true;
    for (int64_t j = 0; j < selected; ++j) {
      const float dx = x[i] - ox[j], dy = y[i] - oy[j];
      const float dx2 = dx * dx, dy2 = dy * dy;
      if (dx2 + dy2 < sep2) {
        keep = false;
        break;
      }
    }
    if (!keep) continue;
    ox[selected] = x[i];
    oy[selected] = y[i];
    of[selected] = f[i];
    ++selected;
  }
  return selected;
}

// Steps shared by the catalogues: rank keys of every local maximum, the
// ranking (raster order for the plain list), the grid cut, and the (x, y,
// flux) of the listed prefix on the host.  Returns the local-maximum count.
unsigned catalogue(gh_stack* s, int64_t index, const gh_candidate_params& p, std::vector<float>& hx,
                   std::vector<float>& hy, std::vector<float>& hf) {
  const float* frame = s->slot[index];
  const int h = static_cast<int>(s->height), w = static_cast<int>(s->width);
  hipStream_t st = s->stream;
  s->cand_count.ensure(2, "hipMalloc(candidate count)");
  unsigned* d_count = s->cand_count.ptr;
  // 1. rank keys of every local maximum (a second pass when the list outgrew
  //    the buffer)
  unsigned total = 0;
  size_t capacity = std::max<size_t>(s->cand_keys.count, std::min<size_t>(static_cast<size_t>(s->npix), 1u << 20));
  for (int pass = 0; pass < 2; ++pass) {
    s->cand_keys.ensure(capacity, "hipMalloc(candidate keys)");
    capacity = s->cand_keys.count;
    check(hipMemsetAsync(d_count, 0, 2 * sizeof(unsigned), st), "hipMemset(candidate count)");
    if (h >= 3 && w >= 3)
      gh::launch_local_max_keys(frame, h, w, p.threshold, s->cand_keys.ptr, static_cast<unsigned>(capacity), d_count,
                                st);
    check(hipMemcpyAsync(&total, d_count, sizeof(unsigned), hipMemcpyDeviceToHost, st), "hipMemcpy(count)");
    check(hipStreamSynchronize(st), "local maxima");
    if (total <= capacity) break;
    capacity = total;
  }
  const unsigned n = total;
  // 2. rank order (raster order for the plain list)
  const bool grid = p.mode == GH_CANDIDATES_GRID_NMS;
  const unsigned cells = grid ? static_cast<unsigned>(p.grid_cols) * static_cast<unsigned>(p.grid_rows) : 0u;
  if (n > 0) {
    size_t temp = gh::sort_keys_temp_bytes(n);
    if (grid) temp = std::max(temp, gh::grid_select_temp_bytes(n, cells));
    s->cand_temp.ensure(temp, "hipMalloc(candidate sort)");
    s->cand_sorted.ensure(n, "hipMalloc(ranked candidates)");
    gh::sort_keys(s->cand_temp.ptr, temp, s->cand_keys.ptr, s->cand_sorted.ptr, n, p.mode == GH_CANDIDATES_ALL, st);
  }
  const unsigned long long* ranked = s->cand_sorted.ptr;
  unsigned listed = 0;  // ranked entries handed to the output / separation filter
  if (p.mode == GH_CANDIDATES_ALL || p.mode == GH_CANDIDATES_TOP || p.mode == GH_CANDIDATES_TOP_NMS) {
    listed = std::min<unsigned>(n, static_cast<unsigned>(p.max_candidates));
  } else if (n > 0) {
    s->cand_u32.ensure(4 * static_cast<size_t>(n) + cells, "hipMalloc(cell scratch)");
    s->cand_keep.ensure(n, "hipMalloc(cell flags)");
    s->cand_selected.ensure(n, "hipMalloc(cell top-k)");
    gh::GridSelectScratch ws{};
    ws.cell = s->cand_u32.ptr;
    ws.pos = ws.cell + n;
    ws.cell_sorted = ws.pos + n;
    ws.pos_sorted = ws.cell_sorted + n;
    ws.start = ws.pos_sorted + n;
    ws.keep = s->cand_keep.ptr;
    ws.temp = s->cand_temp.ptr;
    ws.temp_bytes = s->cand_temp.count;
    gh::grid_select(ranked, n, w, h, p.grid_cols, p.grid_rows, static_cast<unsigned>(p.candidates_per_cell), ws,
                    s->cand_selected.ptr, d_count + 1, st);
    check(hipMemcpyAsync(&listed, d_count + 1, sizeof(unsigned), hipMemcpyDeviceToHost, st), "hipMemcpy(cells)");
    check(hipStreamSynchronize(st), "grid top-k");
    ranked = s->cand_selected.ptr;
  }
  // 3. (x, y, flux) of the listed prefix
  s->cand_xyf.ensure(3 * static_cast<size_t>(std::max<unsigned>(listed, 1)), "hipMalloc(candidate list)");
  float* dx = s->cand_xyf.ptr;
  float* dy = dx + listed;
  float* df = dy + listed;
  hx.assign(listed, 0.0f);
  hy.assign(listed, 0.0f);
  hf.assign(listed, 0.0f);
  if (listed) {
    gh::launch_decode_keys(ranked, listed, frame, w, dx, dy, df, st);
    check(hipMemcpyAsync(hx.data(), dx, listed * sizeof(float), hipMemcpyDeviceToHost, st), "hipMemcpy(x)");
    check(hipMemcpyAsync(hy.data(), dy, listed * sizeof(float), hipMemcpyDeviceToHost, st), "hipMemcpy(y)");
    check(hipMemcpyAsync(hf.data(), df, listed * sizeof(float), hipMemcpyDeviceToHost, st), "hipMemcpy(flux)");
  }
  check(hipStreamSynchronize(st), "candidate list");
  return total;
}

}  // namespace

extern "C" {

int gh_star_candidates(gh_stack* s, int64_t index, const gh_candidate_params* params, float* x, float* y, float* flux,
                       uint8_t* centroid_status, gh_candidate_result* result) {
  return guarded([&] {
    if (!params || !result) throw gh::ValueError("parameters or result pointer is NULL");
    const gh_candidate_params p = *params;
    s->require_loaded(index, "star detection");
    validate(p);
    if (!x || !y || !flux) throw gh::ValueError("output pointers must not be NULL");
    if (s->npix >= (int64_t(1) << 31)) throw gh::ValueError("frames of 2^31 pixels or more are not supported");
    *result = gh_candidate_result{};
    result->centroid_background = NAN;
    std::vector<float> hx, hy, hf;
    result->count = catalogue(s, index, p, hx, hy, hf);
    result->selected_count = static_cast<int64_t>(hx.size());
    const unsigned listed = static_cast<unsigned>(hx.size());
    const float* frame = s->slot[index];
    const int h = static_cast<int>(s->height), w = static_cast<int>(s->width);
    hipStream_t st = s->stream;
    int64_t stored = 0;
    if (p.mode == GH_CANDIDATES_ALL || p.mode == GH_CANDIDATES_TOP) {
      stored = listed;
      std::copy(hx.begin(), hx.end(), x);
      std::copy(hy.begin(), hy.end(), y);
      std::copy(hf.begin(), hf.end(), flux);
    } else {
      stored = separation_filter(hx, hy, hf, p.max_output_candidates, p.min_separation_px, x, y, flux);
    }
    result->stored_count = stored;
    // 4. window centroids of the output
    if (p.centroid_radius > 0 && stored > 0) {
      float background = NAN;
      if (p.centroid_global_mean_background) {
        const unsigned blocks = gh::frame_moments_blocks(s->npix);
        s->moments.ensure(blocks + 1, "hipMalloc(frame moments)");
        gh::launch_frame_moments(frame, s->npix, s->moments.ptr + 1, blocks, s->moments.ptr, st);
        gh::FrameMoments m{};
        check(hipMemcpyAsync(&m, s->moments.ptr, sizeof(m), hipMemcpyDeviceToHost, st), "hipMemcpy(moments)");
        check(hipStreamSynchronize(st), "centroid background");
        if (m.finite > 0) background = static_cast<float>(m.sum / static_cast<double>(m.finite));
        result->centroid_background = background;
      }
      const size_t k = static_cast<size_t>(stored);
      s->cand_xyf.ensure(3 * k, "hipMalloc(candidate list)");
      s->cand_status.ensure(k, "hipMalloc(centroid status)");
      float* cx = s->cand_xyf.ptr;
      float* cy = cx + k;
      float* cf = cy + k;
      check(hipMemcpyAsync(cx, x, k * sizeof(float), hipMemcpyHostToDevice, st), "hipMemcpy(centroid x)");
      check(hipMemcpyAsync(cy, y, k * sizeof(float), hipMemcpyHostToDevice, st), "hipMemcpy(centroid y)");
      check(hipMemcpyAsync(cf, flux, k * sizeof(float), hipMemcpyHostToDevice, st), "hipMemcpy(centroid flux)");
      gh::launch_refine_centroids(frame, w, h, cx, cy, cf, s->cand_status.ptr, static_cast<int>(k), p.centroid_radius,
                                  background, st);
      std::vector<float> before_x(x, x + k), before_y(y, y + k);
      std::vector<uint8_t> status(k);
      check(hipMemcpyAsync(x, cx, k * sizeof(float), hipMemcpyDeviceToHost, st), "hipMemcpy(centroid x)");
      check(hipMemcpyAsync(y, cy, k * sizeof(float), hipMemcpyDeviceToHost, st), "hipMemcpy(centroid y)");
      check(hipMemcpyAsync(status.data(), s->cand_status.ptr, k, hipMemcpyDeviceToHost, st), "hipMemcpy(status)");
      check(hipStreamSynchronize(st), "centroid refine");
      double max_shift = 0.0;
      for (size_t i = 0; i < k; ++i) {
        if (!status[i]) continue;
        ++result->refined_count;
        max_shift = std::max(max_shift, std::hypot(static_cast<double>(x[i]) - before_x[i],
                                                   static_cast<double>(y[i]) - before_y[i]));
      }
      result->max_shift_px = static_cast<float>(max_shift);
      if (centroid_status) std::copy(status.begin(), status.end(), centroid_status);
    }
  });
}

// estimate_translation_from_stars_to_reference (native_bindings.cpp:12645-
// 12880; kernels registration_kernels.cu:265-520, :2007-2110): catalogues of
// both frames (top flux, or the brightest per grid cell in rank order padded
// with (0, 0) entries up to min(count, cells) as the reference's cell slots
// are), every (reference, moving) offset inside the limits, the offset with
// the most offsets within tolerance (first in pair order), mutual nearest
// matches under it, and their mean offset and rms.  The vote runs on the GPU;
// the matching sums run on the host in moving order (the reference adds them
// with float atomics, in no fixed order).
int gh_estimate_translation_from_stars(gh_stack* s, int64_t reference_index, int64_t moving_index,
                                       const gh_star_translation_params* params, gh_star_translation_result* r) {
  return guarded([&] {
    if (!params || !r) throw gh::ValueError("parameters or result pointer is NULL");
    gh_star_translation_params p = *params;
    s->require_loaded(reference_index, "resident star-catalog reference registration");
    s->require_loaded(moving_index, "resident star-catalog moving registration");
    if (p.max_candidates <= 0) throw gh::ValueError("max_candidates must be positive");
    if (p.tolerance_px < 0.0f) throw gh::ValueError("tolerance_px must be non-negative");
    if (p.max_abs_dx < 0.0f) p.max_abs_dx = -1.0f;
    if (p.max_abs_dy < 0.0f) p.max_abs_dy = p.max_abs_dx;
    if (p.prior_radius_px < 0.0f) p.prior_radius_px = -1.0f;
    const bool grid = p.grid_cols > 0 || p.grid_rows > 0;
    if (grid && (p.grid_cols <= 0 || p.grid_rows <= 0))
      throw gh::ValueError("resident star grid dimensions must both be positive");
    if (s->npix >= (int64_t(1) << 31)) throw gh::ValueError("frames of 2^31 pixels or more are not supported");
    const int64_t capacity = grid ? static_cast<int64_t>(p.grid_cols) * p.grid_rows : p.max_candidates;
    gh_candidate_params cp{};
    cp.threshold = p.threshold;
    if (grid) {
      cp.mode = GH_CANDIDATES_GRID_NMS;
      cp.grid_cols = p.grid_cols;
      cp.grid_rows = p.grid_rows;
      cp.candidates_per_cell = 1;
    } else {
      cp.mode = GH_CANDIDATES_TOP;
      cp.max_candidates = p.max_candidates;
    }
    std::vector<float> rx, ry, rf, mx, my, mf;
    const unsigned rtotal = catalogue(s, reference_index, cp, rx, ry, rf);
    const unsigned mtotal = catalogue(s, moving_index, cp, mx, my, mf);
    const size_t rn = static_cast<size_t>(std::min<int64_t>(rtotal, capacity));
    const size_t mn = static_cast<size_t>(std::min<int64_t>(mtotal, capacity));
    rx.resize(rn, 0.0f);
    ry.resize(rn, 0.0f);
    mx.resize(mn, 0.0f);
    my.resize(mn, 0.0f);
    *r = gh_star_translation_result{};
    r->reference_total_count = rtotal;
    r->moving_total_count = mtotal;
    r->reference_count = static_cast<int64_t>(rn);
    r->moving_count = static_cast<int64_t>(mn);
    r->max_abs_dx = p.max_abs_dx;
    r->max_abs_dy = p.max_abs_dy;
    r->prior_radius_px = p.prior_radius_px;
    if (rn == 0 || mn == 0) throw std::runtime_error("resident star-catalog registration found no stars");
    // pair offsets (catalog_pair_offsets_kernel, :265-305)
    const size_t pairs = rn * mn;
    if (pairs > static_cast<size_t>(INT32_MAX)) throw gh::ValueError("too many catalogue pairs");
    std::vector<float> off(2 * pairs);
    for (size_t i = 0; i < pairs; ++i) {
      const size_t a = i / mn, b = i % mn;
      float dx = rx[a] - mx[b], dy = ry[a] - my[b];
      if ((p.max_abs_dx >= 0.0f && std::fabs(dx) > p.max_abs_dx) ||
          (p.max_abs_dy >= 0.0f && std::fabs(dy) > p.max_abs_dy)) {
        dx = dy = NAN;
      } else if (p.prior_radius_px >= 0.0f) {
        const float ex = dx - p.prior_dx, ey = dy - p.prior_dy;
        const float e2 = ex * ex, f2 = ey * ey, rr = p.prior_radius_px * p.prior_radius_px;
        if (e2 + f2 > rr) dx = dy = NAN;
      }
      off[i] = dx;
      off[pairs + i] = dy;
    }
    hipStream_t st = s->stream;
    s->cand_xyf.ensure(2 * pairs, "hipMalloc(catalogue offsets)");
    s->cand_u32.ensure(pairs, "hipMalloc(offset scores)");
    check(hipMemcpyAsync(s->cand_xyf.ptr, off.data(), 2 * pairs * sizeof(float), hipMemcpyHostToDevice, st),
          "hipMemcpy(offsets)");
    const float tol2 = p.tolerance_px * p.tolerance_px;
    gh::launch_offset_scores(s->cand_xyf.ptr, s->cand_xyf.ptr + pairs, static_cast<int>(pairs), tol2,
                             reinterpret_cast<int*>(s->cand_u32.ptr), st);
    std::vector<int> scores(pairs);
    check(hipMemcpyAsync(scores.data(), s->cand_u32.ptr, pairs * sizeof(int), hipMemcpyDeviceToHost, st),
          "hipMemcpy(scores)");
    check(hipStreamSynchronize(st), "catalogue offset vote");
    // best offset (catalog_offset_best_kernel, :337-362): first of the most
    int best = -1, best_score = -1;
    for (size_t i = 0; i < pairs; ++i)
      if (scores[i] > best_score) {
        best_score = scores[i];
        best = static_cast<int>(i);
      }
    const float bdx = best < 0 ? NAN : off[best], bdy = best < 0 ? NAN : off[pairs + best];
    r->dx = bdx;
    r->dy = bdy;
    r->inliers = best < 0 ? 0 : best_score;
    // mutual nearest matches within tolerance (:363-431; ties go to the later index)
    std::vector<int> mbest(mn, -1), rbest(rn, -1);
    for (size_t b = 0; b < mn; ++b) {
      const float tx = mx[b] + bdx, ty = my[b] + bdy;
      float d2best = tol2;
      for (size_t a = 0; a < rn; ++a) {
        const float dx = rx[a] - tx, dy = ry[a] - ty;
        const float d2x = dx * dx, d2y = dy * dy;
        const float d2 = d2x + d2y;
        if (d2 <= d2best) {
          d2best = d2;
          mbest[b] = static_cast<int>(a);
        }
      }
    }
    for (size_t a = 0; a < rn; ++a) {
      float d2best = tol2;
      for (size_t b = 0; b < mn; ++b) {
        const float tx = mx[b] + bdx, ty = my[b] + bdy;
        const float dx = rx[a] - tx, dy = ry[a] - ty;
        const float d2x = dx * dx, d2y = dy * dy;
        const float d2 = d2x + d2y;
        if (d2 <= d2best) {
          d2best = d2;
          rbest[a] = static_cast<int>(b);
        }
      }
    }
    float sx = 0.0f, sy = 0.0f;
    int mutual = 0;
    for (size_t b = 0; b < mn; ++b) {
      const int a = mbest[b];
      if (a < 0 || rbest[a] != static_cast<int>(b)) continue;
      sx += rx[a] - mx[b];
      sy += ry[a] - my[b];
      ++mutual;
    }
    const float scale = mutual > 0 ? 1.0f / static_cast<float>(mutual) : 0.0f;
    const float rdx = mutual > 0 ? sx * scale : bdx, rdy = mutual > 0 ? sy * scale : bdy;
    float s2 = 0.0f;
    for (size_t b = 0; b < mn; ++b) {
      const int a = mbest[b];
      if (a < 0 || rbest[a] != static_cast<int>(b)) continue;
      const float dx = rx[a] - (mx[b] + rdx), dy = ry[a] - (my[b] + rdy);
      const float d2x = dx * dx, d2y = dy * dy;
      s2 += d2x + d2y;
    }
    r->refined_dx = rdx;
    r->refined_dy = rdy;
    r->mutual_inliers = mutual;
    r->rms_px = mutual > 0 ? std::sqrt(s2 / static_cast<float>(mutual)) : NAN;
  });
}

}  // extern "C"
