// C ABI of the resident local normalisation (pair-grid statistics and the
// in-place global / grid application, native_bindings.cpp:10645-11290), the
// batched frame statistics and the mean / sigma-clip integrations that sample
// the frames themselves (native_bindings.cpp:13227-13393, :14632-15115).

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "stack.hpp"

using ghi::guarded;

namespace {

// Tiles of the statistics / coefficient grid; a tile dimension beyond the
// frame maps every pixel to tile 0 on that axis, so it is clamped to the frame
// (same grid, no 32-bit overflow of tile corners).
struct LnGrid {
  int tile_h, tile_w, rows, cols;
};

LnGrid ln_grid(const gh_stack* s, int32_t tile_height, int32_t tile_width) {
  if (tile_height <= 0 || tile_width <= 0) throw gh::ValueError("tile dimensions must be positive");
  LnGrid g{};
  g.tile_h = static_cast<int>(std::min<int64_t>(tile_height, s->height));
  g.tile_w = static_cast<int>(std::min<int64_t>(tile_width, s->width));
  g.rows = static_cast<int>((s->height + g.tile_h - 1) / g.tile_h);
  g.cols = static_cast<int>((s->width + g.tile_w - 1) / g.tile_w);
  if (g.rows <= 0 || g.cols <= 0)
    throw std::runtime_error("resident frame shape produced an empty local-normalization grid");
  return g;
}

}  // namespace

extern "C" {

int gh_pair_grid_stats(gh_stack* s, int64_t reference_index, const int64_t* source_indices, int64_t count,
                       int32_t tile_height, int32_t tile_width, float* source_mean, float* source_std,
                       float* reference_mean, float* reference_std, uint64_t* valid_pixels) {
  return guarded([&] {
    s->require_loaded(reference_index, "grid local normalization reference statistics");
    if (count < 0) throw gh::ValueError("source count must not be negative");
    if (count > 0 && !source_indices) throw gh::ValueError("source indices pointer is NULL");
    const LnGrid g = ln_grid(s, tile_height, tile_width);
    for (int64_t k = 0; k < count; ++k) {
      if (source_indices[k] < 0 || source_indices[k] >= s->frames)
        throw gh::IndexError("source index is out of resident stack range");
      s->require_loaded(source_indices[k], "grid local normalization source statistics");
    }
    if (count == 0) return;
    const int64_t tiles = static_cast<int64_t>(g.rows) * g.cols;
    const int64_t chunk = std::min<int64_t>(count, gh::kLnBatch);
    s->ln_sums.ensure(static_cast<size_t>(chunk * tiles * 4), "hipMalloc(local normalization sums)");
    s->ln_counts.ensure(static_cast<size_t>(chunk * tiles), "hipMalloc(local normalization counts)");
    std::vector<double> sums(static_cast<size_t>(chunk * tiles * 4));
    std::vector<unsigned long long> counts(static_cast<size_t>(chunk * tiles));
    s->wait_caller();
    for (int64_t start = 0; start < count; start += chunk) {
      const int n = static_cast<int>(std::min<int64_t>(chunk, count - start));
      gh::LnBatch b{};
      for (int j = 0; j < n; ++j) b.frame[j] = s->slot[source_indices[start + j]];
      gh::launch_pair_grid_sums(b, n, s->slot[reference_index], static_cast<int>(s->width),
                                static_cast<int>(s->height), g.tile_w, g.tile_h, g.cols, g.rows, s->ln_sums.ptr,
                                s->ln_counts.ptr, 0, s->stream);
      check(hipMemcpyAsync(sums.data(), s->ln_sums.ptr, static_cast<size_t>(n * tiles * 4) * sizeof(double),
                           hipMemcpyDeviceToHost, s->stream),
            "hipMemcpy(local normalization sums)");
      check(hipMemcpyAsync(counts.data(), s->ln_counts.ptr,
                           static_cast<size_t>(n * tiles) * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           s->stream),
            "hipMemcpy(local normalization counts)");
      check(hipStreamSynchronize(s->stream), "pair grid statistics");
      // Means and standard deviations as native_bindings.cpp:10740-10768:
      // FP64 moments about zero, variance clamped at 0, rounded to float32.
      for (int64_t r = 0; r < n * tiles; ++r) {
        const int64_t o = start * tiles + r;
        const unsigned long long c = counts[r];
        float sm = 0.0f, sd = 0.0f, rm = 0.0f, rd = 0.0f;
        if (c != 0) {
          const double inv = 1.0 / static_cast<double>(c);
          const double s_mean = sums[r * 4 + 0] * inv;
          const double r_mean = sums[r * 4 + 2] * inv;
          double s_var = sums[r * 4 + 1] * inv - s_mean * s_mean;
          double r_var = sums[r * 4 + 3] * inv - r_mean * r_mean;
          if (s_var < 0.0) s_var = 0.0;
          if (r_var < 0.0) r_var = 0.0;
          sm = static_cast<float>(s_mean);
          sd = static_cast<float>(std::sqrt(s_var));
          rm = static_cast<float>(r_mean);
          rd = static_cast<float>(std::sqrt(r_var));
        }
        if (source_mean) source_mean[o] = sm;
        if (source_std) source_std[o] = sd;
        if (reference_mean) reference_mean[o] = rm;
        if (reference_std) reference_std[o] = rd;
        if (valid_pixels) valid_pixels[o] = c;
      }
    }
  });
}

int gh_frames_stats(gh_stack* s, const int64_t* indices, int64_t count, gh_frame_statistics* out) {
  return guarded([&] {
    if (count < 0) throw gh::ValueError("frame count must not be negative");
    if (count == 0) return;
    if (!indices || !out) throw gh::ValueError("indices and output pointers must not be NULL");
    for (int64_t k = 0; k < count; ++k) s->require_loaded(indices[k], "global frame statistics");
    const unsigned blocks = gh::frame_moments_blocks(s->npix);
    const int64_t chunk = std::min<int64_t>(count, gh::kLnBatch);
    s->moments.ensure(static_cast<size_t>(chunk) * (blocks + 1), "hipMalloc(frame moments)");
    gh::FrameMoments* partials = s->moments.ptr + chunk;
    std::vector<gh::FrameMoments> m(static_cast<size_t>(chunk));
    s->wait_caller();
    for (int64_t start = 0; start < count; start += chunk) {
      const int n = static_cast<int>(std::min<int64_t>(chunk, count - start));
      gh::LnBatch b{};
      for (int j = 0; j < n; ++j) b.frame[j] = s->slot[indices[start + j]];
      gh::launch_frame_moments_batch(b, n, s->npix, partials, blocks, s->moments.ptr, s->stream);
      check(hipMemcpyAsync(m.data(), s->moments.ptr, static_cast<size_t>(n) * sizeof(gh::FrameMoments),
                           hipMemcpyDeviceToHost, s->stream),
            "hipMemcpy(frame moments)");
      check(hipStreamSynchronize(s->stream), "frames statistics");
      for (int j = 0; j < n; ++j) ghi::fill_frame_statistics(m[j], s->npix, out + start + j);
    }
  });
}

int gh_apply_normalization(gh_stack* s, const int64_t* indices, int64_t count, const float* scales,
                           const float* offsets, int32_t tile_height, int32_t tile_width, void* stream_ptr) {
  return guarded([&] {
    if (count < 0) throw gh::ValueError("frame count must not be negative");
    const LnGrid g = ln_grid(s, tile_height, tile_width);
    if (count == 0) return;
    if (!indices) throw gh::ValueError("frame indices pointer is NULL");
    if (!scales || !offsets) throw gh::ValueError("scales and offsets must not be NULL");
    for (int64_t k = 0; k < count; ++k) {
      if (indices[k] < 0 || indices[k] >= s->frames) throw gh::IndexError("frame index is out of resident stack range");
      s->require_loaded(indices[k], "grid local normalization");
    }
    const int64_t grid = static_cast<int64_t>(g.rows) * g.cols;
    const bool dev_s = is_device_pointer(scales), dev_o = is_device_pointer(offsets);
    if (dev_s != dev_o) throw gh::ValueError("scales and offsets must both be host or both be device pointers");
    hipStream_t st = s->pick(stream_ptr);
    const float* sc = scales;
    const float* of = offsets;
    if (!dev_s) {
      // Host coefficients are staged in the stack's buffer; the call then
      // completes before returning, so the buffer is free for the next call.
      const size_t n = static_cast<size_t>(count * grid);
      s->ln_coef.ensure(2 * n, "hipMalloc(local normalization coefficients)");
      check(hipMemcpyAsync(s->ln_coef.ptr, scales, n * sizeof(float), hipMemcpyHostToDevice, st),
            "hipMemcpy(local normalization scales)");
      check(hipMemcpyAsync(s->ln_coef.ptr + n, offsets, n * sizeof(float), hipMemcpyHostToDevice, st),
            "hipMemcpy(local normalization offsets)");
      sc = s->ln_coef.ptr;
      of = s->ln_coef.ptr + n;
    }
    if (st == s->stream) s->wait_caller();
    // A frame named twice is normalised twice, in order, as consecutive
    // launches do in the reference: a repeat starts a new launch.
    int64_t start = 0;
    while (start < count) {
      gh::LnBatch b{};
      int n = 0;
      while (start + n < count && n < gh::kLnBatch) {
        const int64_t f = indices[start + n];
        bool repeat = false;
        for (int j = 0; j < n; ++j) repeat = repeat || b.frame[j] == s->slot[f];
        if (repeat) break;
        b.frame[n] = s->slot[f];
        b.flag[n] = s->nonfinite + f;
        ++n;
      }
      gh::launch_apply_grid(b, n, sc, of, static_cast<int>(s->width), static_cast<int>(s->height), g.tile_w,
                            g.tile_h, g.cols, g.rows, static_cast<int>(start), st);
      for (int j = 0; j < n; ++j) s->calib_stats_valid[indices[start + j]] = 0;
      start += n;
    }
    s->mark_caller(st);
    if (!dev_s || !stream_ptr) check(hipStreamSynchronize(st), "apply normalization");
  });
}

}  // extern "C"

namespace {

// Tile-local weight policy of gh_integrate_tile_local.
struct TilePolicy {
  const uint8_t* target_mask;  // [frames]
  const int32_t* extents;      // [tiles][4]
  const float* multipliers;    // [tiles]
  int64_t tiles;
  int64_t max_area;            // filled in by validate_tiles
};

// Checks of native_bindings.cpp:13042-13096 (same order and messages).
void validate_tiles(const gh_stack* s, TilePolicy& t) {
  if (!t.target_mask) throw gh::ValueError("target_mask must have shape (frame_count,)");
  int64_t targets = 0;
  for (int64_t f = 0; f < s->frames; ++f) targets += t.target_mask[f] != 0;
  if (targets == 0) throw gh::ValueError("target_mask must select at least one resident frame");
  if (t.tiles <= 0 || t.tiles > INT32_MAX || !t.extents)
    throw gh::ValueError("tile_extents must contain between 1 and INT_MAX tiles");
  if (!t.multipliers) throw gh::ValueError("tile_multipliers must have shape (tile_count,)");
  t.max_area = 0;
  for (int64_t i = 0; i < t.tiles; ++i) {
    const int32_t* e = t.extents + 4 * i;
    if (e[0] < 0 || e[1] < 0 || e[2] <= e[0] || e[3] <= e[1] || e[2] > s->width || e[3] > s->height)
      throw gh::ValueError("tile_extents must be positive half-open rectangles inside the image");
    if (!std::isfinite(t.multipliers[i]) || t.multipliers[i] < 0.0f)
      throw gh::ValueError("tile_multipliers must be finite non-negative values");
    t.max_area = std::max<int64_t>(t.max_area, int64_t(e[2] - e[0]) * (e[3] - e[1]));
  }
  for (int64_t i = 0; i < t.tiles; ++i) {
    const int32_t* l = t.extents + 4 * i;
    for (int64_t j = i + 1; j < t.tiles; ++j) {
      const int32_t* r = t.extents + 4 * j;
      if (std::max(l[0], r[0]) < std::min(l[2], r[2]) && std::max(l[1], r[1]) < std::min(l[3], r[3]))
        throw gh::ValueError("tile_extents must not overlap for deterministic tile-local integration");
    }
  }
}

size_t round256(size_t b) { return (b + 255) / 256 * 256; }

void integrate_clip(gh_stack* s, const float* weights, const double* inverses, const gh_clip_params& prm,
                    TilePolicy* tiles, float* const user[6], void* stream_ptr) {
  if (prm.sampler != GH_CLIP_STACK && prm.sampler != GH_INTERP_BILINEAR && prm.sampler != GH_INTERP_LANCZOS3)
    throw gh::ValueError("interpolation must be bilinear or lanczos3");
  if (prm.rejection != GH_REJECT_NONE && prm.rejection != GH_REJECT_SIGMA_CLIP)
    throw gh::ValueError("unsupported rejection for the clip integration");
  if (s->loaded_count != s->frames)
    throw std::runtime_error("all resident frames must be loaded before integration");
  if (prm.rejection == GH_REJECT_SIGMA_CLIP && !(prm.low_sigma > 0.0f && prm.high_sigma > 0.0f))
    throw gh::ValueError("sigma thresholds must be positive");
  if (prm.sampler != GH_CLIP_STACK && !inverses) throw gh::ValueError("inverse matrices pointer is NULL");
  if (s->npix >= (int64_t(1) << 31)) throw gh::ValueError("frames of 2^31 pixels or more are not supported");
  if (tiles) {
    if (prm.sampler != GH_CLIP_STACK) throw gh::ValueError("tile-local integration samples the stack as stored");
    validate_tiles(s, *tiles);
  }
  // Frames of positive finite weight, in order (the kernels skip the rest;
  // a tile multiplier >= 0 cannot bring a skipped frame back).
  std::vector<const float*> table;
  std::vector<float> w;
  std::vector<double> inv;
  std::vector<unsigned char> target;
  for (int64_t f = 0; f < s->frames; ++f) {
    const float wf = weights ? weights[f] : 1.0f;
    if (!(wf > 0.0f) || !std::isfinite(wf)) continue;
    table.push_back(s->slot[f]);
    w.push_back(wf);
    if (tiles) target.push_back(tiles->target_mask[f] != 0 ? 1 : 0);
    if (prm.sampler != GH_CLIP_STACK) {
      for (int e = 0; e < 9; ++e) {
        if (!std::isfinite(inverses[f * 9 + e])) throw gh::ValueError("warp inverse matrix must be finite");
        inv.push_back(inverses[f * 9 + e]);
      }
    }
  }
  hipStream_t st = s->pick(stream_ptr);
  if (st == s->stream) s->wait_caller();
  const size_t n = table.size();
  // one allocation: pointer table, weights, inverses | target flags, tile
  // extents, tile multipliers
  const size_t table_bytes = round256(n * sizeof(const float*));
  const size_t w_bytes = round256(n * sizeof(float));
  const size_t inv_bytes = round256(inv.size() * sizeof(double));
  const size_t target_bytes = round256(target.size());
  const size_t ntiles = tiles ? static_cast<size_t>(tiles->tiles) : 0;
  const size_t ext_bytes = round256(ntiles * 4 * sizeof(int32_t));
  const size_t mult_bytes = round256(ntiles * sizeof(float));
  s->clip_tables.ensure(table_bytes + w_bytes + inv_bytes + target_bytes + ext_bytes + mult_bytes + 8,
                        "hipMalloc(clip tables)");
  unsigned char* base = s->clip_tables.ptr;
  unsigned char* const w_at = base + table_bytes;
  unsigned char* const inv_at = w_at + w_bytes;
  unsigned char* const target_at = inv_at + inv_bytes;
  unsigned char* const ext_at = target_at + target_bytes;
  unsigned char* const mult_at = ext_at + ext_bytes;
  s->wait_tables();
  if (n) {
    check(hipMemcpy(base, table.data(), n * sizeof(const float*), hipMemcpyHostToDevice), "hipMemcpy(clip table)");
    check(hipMemcpy(w_at, w.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy(clip weights)");
    if (!inv.empty())
      check(hipMemcpy(inv_at, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice),
            "hipMemcpy(clip inverses)");
    if (!target.empty())
      check(hipMemcpy(target_at, target.data(), n, hipMemcpyHostToDevice), "hipMemcpy(tile-local target mask)");
  }
  if (tiles) {
    check(hipMemcpy(ext_at, tiles->extents, ntiles * 4 * sizeof(int32_t), hipMemcpyHostToDevice),
          "hipMemcpy(tile-local extents)");
    check(hipMemcpy(mult_at, tiles->multipliers, ntiles * sizeof(float), hipMemcpyHostToDevice),
          "hipMemcpy(tile-local multipliers)");
  }
  gh::ClipArgs a{};
  a.frames = reinterpret_cast<const float* const*>(base);
  a.weights = reinterpret_cast<const float*>(w_at);
  a.inverses = inv.empty() ? nullptr : reinterpret_cast<const double*>(inv_at);
  a.n = static_cast<int>(n);
  a.width = static_cast<int>(s->width);
  a.height = static_cast<int>(s->height);
  a.npix = s->npix;
  a.clamp = prm.clamping_threshold;
  // affine inverses skip the projective divide (x / 1.0 is x exactly)
  a.affine = 1;
  for (size_t k = 0; k < n && !inv.empty(); ++k)
    if (!(inv[k * 9 + 6] == 0.0 && inv[k * 9 + 7] == 0.0 && inv[k * 9 + 8] == 1.0)) a.affine = 0;
  a.mode = prm.rejection == GH_REJECT_SIGMA_CLIP ? 1 : 0;
  a.winsorize = prm.winsorize ? 1 : 0;
  a.low_sigma = prm.low_sigma;
  a.high_sigma = prm.high_sigma;
  float* dev[6] = {};
  // Host outputs go through device temporaries (the integration outputs'
  // five, plus one for the geometric coverage) and are copied back.
  for (int i = 0; i < 6; ++i) {
    if (!user[i]) continue;
    if (is_device_pointer(user[i])) {
      dev[i] = user[i];
      continue;
    }
    DeviceBuffer<float>& tmp = i < 5 ? s->out_f32[i] : s->clip_geometric;
    tmp.ensure(static_cast<size_t>(s->npix), "hipMalloc(clip output)");
    dev[i] = tmp.ptr;
  }
  a.master = dev[0];
  a.weight = dev[1];
  a.coverage = dev[2];
  a.low = dev[3];
  a.high = dev[4];
  a.geometric = dev[5];
  if (tiles) {
    s->clip_tile_map.ensure(static_cast<size_t>(s->npix), "hipMalloc(tile-local multiplier map)");
    gh::launch_tile_map(s->clip_tile_map.ptr, reinterpret_cast<const int*>(ext_at),
                        reinterpret_cast<const float*>(mult_at), static_cast<int>(ntiles), tiles->max_area,
                        a.width, a.npix, st);
    a.tile_map = s->clip_tile_map.ptr;
    a.target = target_at;
    gh::launch_clip_integrate_tiles(a, st);
  } else {
    const int sampler = prm.sampler == GH_CLIP_STACK ? 0 : (prm.sampler == GH_INTERP_BILINEAR ? 1 : 2);
    gh::launch_clip_integrate(a, sampler, st);
  }
  s->mark_tables(st);
  bool any_host = false;
  for (int i = 0; i < 6; ++i) {
    if (!user[i] || dev[i] == user[i]) continue;
    any_host = true;
    check(hipMemcpyAsync(user[i], dev[i], static_cast<size_t>(s->npix) * sizeof(float), hipMemcpyDeviceToHost, st),
          "hipMemcpy(clip output)");
  }
  s->mark_caller(st);
  if (any_host || !stream_ptr) check(hipStreamSynchronize(st), "clip integration");
}

}  // namespace

extern "C" {

int gh_integrate_clip(gh_stack* s, const float* weights, const double* inverses, const gh_clip_params* params,
                      float* master, float* weight, float* coverage, float* low, float* high, float* geometric,
                      void* stream_ptr) {
  return guarded([&] {
    if (!params) throw gh::ValueError("clip parameters pointer is NULL");
    float* const user[6] = {master, weight, coverage, low, high, geometric};
    integrate_clip(s, weights, inverses, *params, nullptr, user, stream_ptr);
  });
}

int gh_integrate_tile_local(gh_stack* s, const float* weights, const uint8_t* target_mask,
                            const int32_t* tile_extents, const float* tile_multipliers, int64_t tile_count,
                            const gh_clip_params* params, float* master, float* weight, float* coverage, float* low,
                            float* high, void* stream_ptr) {
  return guarded([&] {
    if (!params) throw gh::ValueError("clip parameters pointer is NULL");
    TilePolicy t{target_mask, tile_extents, tile_multipliers, tile_count, 0};
    float* const user[6] = {master, weight, coverage, low, high, nullptr};
    integrate_clip(s, weights, nullptr, *params, &t, user, stream_ptr);
  });
}

}  // extern "C"

extern "C" {

int gh_estimate_translation(gh_stack* s, int64_t reference_index, int64_t moving_index, int32_t subpixel,
                            int32_t radius_x, int32_t radius_y, float center_dx, float center_dy, float step,
                            int32_t sample_stride, float* dx, float* dy, float* score, int64_t* candidates) {
  return guarded([&] {
    const char* what = subpixel ? "resident subpixel translation search" : "resident translation search";
    s->require_loaded(reference_index, what);
    s->require_loaded(moving_index, what);
    if (subpixel) {
      if (radius_x < 0 || radius_y < 0) throw gh::ValueError("radius_steps must be non-negative");
      if (!(step > 0.0f)) throw gh::ValueError("step must be positive");
    } else if (radius_x < 0 || radius_y < 0) {
      throw gh::ValueError("max shifts must be non-negative");
    }
    if (sample_stride <= 0) throw gh::ValueError("sample_stride must be positive");
    if (!dx || !dy || !score) throw gh::ValueError("output pointers must not be NULL");
    const int64_t n = (2 * static_cast<int64_t>(radius_x) + 1) * (2 * static_cast<int64_t>(radius_y) + 1);
    if (n > (int64_t(1) << 24)) throw gh::ValueError("too many translation candidates");
    gh::NccArgs a{};
    a.reference = s->slot[reference_index];
    a.moving = s->slot[moving_index];
    a.width = static_cast<int>(s->width);
    a.height = static_cast<int>(s->height);
    a.stride = sample_stride;
    a.radius_x = radius_x;
    a.radius_y = radius_y;
    a.center_dx = center_dx;
    a.center_dy = center_dy;
    a.step = step;
    const int chunks = gh::ncc_chunks(a);
    const size_t part = static_cast<size_t>(n) * static_cast<size_t>(chunks) * 6;
    s->ncc_ws.ensure(part + static_cast<size_t>(n), "hipMalloc(translation search)");
    s->wait_caller();
    gh::launch_ncc_search(a, subpixel != 0, s->ncc_ws.ptr, s->ncc_ws.ptr + part, s->stream);
    std::vector<double> scores(static_cast<size_t>(n));
    check(hipMemcpyAsync(scores.data(), s->ncc_ws.ptr + part, static_cast<size_t>(n) * sizeof(double),
                         hipMemcpyDeviceToHost, s->stream),
          "hipMemcpy(translation scores)");
    check(hipStreamSynchronize(s->stream), "translation search");
    // first best candidate in row-major candidate order
    int64_t best = 0;
    for (int64_t i = 1; i < n; ++i)
      if (scores[i] > scores[best]) best = i;
    const int cw = 2 * radius_x + 1;
    const int ox = static_cast<int>(best % cw) - radius_x, oy = static_cast<int>(best / cw) - radius_y;
    if (subpixel) {
      *dx = center_dx + static_cast<float>(ox) * step;
      *dy = center_dy + static_cast<float>(oy) * step;
    } else {
      *dx = static_cast<float>(ox);
      *dy = static_cast<float>(oy);
    }
    *score = static_cast<float>(scores[best]);
    if (candidates) *candidates = n;
  });
}

}  // extern "C"
