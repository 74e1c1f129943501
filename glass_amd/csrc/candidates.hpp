// Star candidate catalogues (GPU ranking of 3x3 local maxima, per-cell top-k,
// window centroids); the greedy separation filter runs on the host over the
// ranked prefix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "common.hpp"

namespace gh {

constexpr int kCentroidMaxRadius = 8;
constexpr int kCentroidMaxSamples = (2 * kCentroidMaxRadius + 1) * (2 * kCentroidMaxRadius + 1);

// keys[slot] = rank key of every local maximum (slot < capacity); *count
// counts all of them (must start at 0).
void launch_local_max_keys(const float* frame, int height, int width, float threshold, unsigned long long* keys,
                           unsigned capacity, unsigned* count, hipStream_t stream);
size_t sort_keys_temp_bytes(unsigned n);
// Ascending key sort = rank order; position_only sorts by pixel index (raster
// order).
void sort_keys(void* temp, size_t temp_bytes, const unsigned long long* in, unsigned long long* out, unsigned n,
               bool position_only, hipStream_t stream);
// (x, y, flux) of keys[0..n).
void launch_decode_keys(const unsigned long long* keys, unsigned n, const float* frame, int width, float* xs, float* ys,
                        float* fluxes, hipStream_t stream);

struct GridSelectScratch {
  unsigned *cell, *pos, *cell_sorted, *pos_sorted, *start;  // n, n, n, n, cells
  unsigned char* keep;                                      // n
  void* temp;
  size_t temp_bytes;                                        // >= grid_select_temp_bytes
};
size_t grid_select_temp_bytes(unsigned n, unsigned cells);
// out = the ranked keys that are among the first per_cell of their grid cell,
// in rank order; *out_count their number.
void grid_select(const unsigned long long* ranked, unsigned n, int width, int height, int cols, int rows,
                 unsigned per_cell, const GridSelectScratch& w, unsigned long long* out, unsigned* out_count,
                 hipStream_t stream);
// scores[i] = #finite offsets within squared distance tol2 of offset i (-1
// when offset i is not finite).
void launch_offset_scores(const float* dx, const float* dy, int n, float tol2, int* scores, hipStream_t stream);
// Alignment metrics of translated matrix candidates: partials[(c * chunks +
// k) * 8 + j], j = sums of r, m, r^2, m^2, r*m, (m-r)^2, |m-r| and the count
// over sample rows [k * rows_per_chunk, ...) of candidate c.
struct AlignArgs {
  const float* reference;
  const float* const* moving;  // [candidates] device table: the frame each candidate samples
  const float* inverses;       // [candidates][9] float32
  int width, height, stride, sample_w, sample_h, rows_per_chunk, chunks, candidates;
};
void launch_alignment_partials(const AlignArgs& a, double* partials, hipStream_t stream);
void launch_refine_centroids(const float* frame, int width, int height, float* xs, float* ys, const float* fluxes,
                             unsigned char* status, int count, int radius, float background, hipStream_t stream);

}  // namespace gh
