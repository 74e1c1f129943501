// Star candidate catalogues of one resident frame
// (cpp/cuda/star_detect_kernels.cu:41-744): 3x3 local maxima above a
// threshold, ranked by flux (descending), then y, then x (ascending).
//
// The reference keeps its top-N catalogues with a spin lock per catalogue (or
// per grid cell) that every candidate thread takes in turn.  Here every
// candidate writes one 64-bit rank key with a wave-aggregated atomic append,
// and one radix sort of the keys gives the whole ranking: key order is the
// reference's order (flux descending, then pixel index = y * width + x
// ascending), so a prefix of the sorted keys is its top-N catalogue.  The
// per-cell top-k of the grid form is a stable sort of the ranked list by cell
// followed by a rank-within-cell cut.

#include <hipcub/hipcub.hpp>

#include "candidates.hpp"

namespace gh {
namespace {

// Order key of a candidate: the float's bits made monotone (negatives
// complemented, positives with the sign bit set), complemented again for
// descending flux, above the pixel index.  -0 is keyed as +0, as the
// reference's comparisons treat the two as one flux.
__device__ __forceinline__ unsigned long long candidate_key(float v, unsigned pix) {
  unsigned u = __float_as_uint(v + 0.0f);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return (static_cast<unsigned long long>(~u) << 32) | pix;
}

// is_local_maximum (star_detect_kernels.cu:41-69): interior pixel, finite,
// above the threshold and not below any of its 8 neighbours (a NaN neighbour
// does not disqualify it).
__global__ __launch_bounds__(256) void local_max_keys_kernel(const float* __restrict__ frame, int height, int width,
                                                            float threshold, unsigned long long* keys,
                                                            unsigned capacity, unsigned* count) {
  const long long inner_w = width - 2, inner = inner_w * static_cast<long long>(height - 2);
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i - threadIdx.x < inner;
       i += stride) {
    bool hit = false;
    long long pix = 0;
    float v = 0.0f;
    if (i < inner) {
      const int y = static_cast<int>(i / inner_w) + 1, x = static_cast<int>(i % inner_w) + 1;
      pix = static_cast<long long>(y) * width + x;
      v = frame[pix];
      hit = isfinite(v) && v > threshold;
      if (hit) {
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            if (dx || dy) hit = hit && !(v < frame[pix + static_cast<long long>(dy) * width + dx]);
          }
        }
      }
    }
    const unsigned long long mask = __ballot(hit);
    if (mask) {
      const int lane = threadIdx.x & 63;
      const int leader = __ffsll(static_cast<long long>(mask)) - 1;
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(count, static_cast<unsigned>(__popcll(mask)));
      base = __shfl(base, leader, 64);
      if (hit) {
        const unsigned slot = base + static_cast<unsigned>(__popcll(mask & ((1ull << lane) - 1ull)));
        if (slot < capacity) keys[slot] = candidate_key(v, static_cast<unsigned>(pix));
      }
    }
  }
}

__global__ void decode_keys_kernel(const unsigned long long* __restrict__ keys, unsigned n,
                                   const float* __restrict__ frame, int width, float* xs, float* ys, float* fluxes) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const unsigned pix = static_cast<unsigned>(keys[j]);
  const unsigned y = pix / static_cast<unsigned>(width);
  xs[j] = static_cast<float>(pix - y * static_cast<unsigned>(width));
  ys[j] = static_cast<float>(y);
  fluxes[j] = frame[pix];
}

// Grid cell of each ranked candidate (star_detect_kernels.cu:257-259) and its
// rank position.
__global__ void cell_of_kernel(const unsigned long long* __restrict__ keys, unsigned n, int width, int height,
                               int cols, int rows, unsigned* cell, unsigned* pos) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const unsigned pix = static_cast<unsigned>(keys[j]);
  const long long y = pix / static_cast<unsigned>(width), x = pix - y * width;
  const long long cx = min((x * cols) / width, static_cast<long long>(cols - 1));
  const long long cy = min((y * rows) / height, static_cast<long long>(rows - 1));
  cell[j] = static_cast<unsigned>(cy * cols + cx);
  pos[j] = j;
}

__global__ void cell_start_kernel(const unsigned* __restrict__ cell, unsigned n, unsigned* start) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  if (j == 0 || cell[j] != cell[j - 1]) start[cell[j]] = j;
}

// cell / pos sorted by cell, rank order kept within a cell: the first k of
// each cell are its top k.
__global__ void cell_cut_kernel(const unsigned* __restrict__ cell, const unsigned* __restrict__ pos, unsigned n,
                                const unsigned* __restrict__ start, unsigned k, unsigned char* keep) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  keep[pos[j]] = (j - start[cell[j]]) < k ? 1 : 0;
}

// star_refine_centroids_kernel (star_detect_kernels.cu:650-742), one 64-lane
// block per candidate: background = the window's median of finite samples
// (the mean of the middle pair in float32) unless the override is finite;
// weights max(v - background, 0) > 0; FP64 sums in raster order on lane 0.
__global__ __launch_bounds__(64) void refine_centroids_kernel(const float* __restrict__ frame, int width, int height,
                                                              float* xs, float* ys, const float* __restrict__ fluxes,
                                                              unsigned char* status, int radius,
                                                              float background_override) {
  __shared__ float win[kCentroidMaxSamples];
  __shared__ float middle[2];
  __shared__ int finite_count;
  const int i = blockIdx.x, lane = threadIdx.x;
  const float flux = fluxes[i];
  const int cx = static_cast<int>(floorf(xs[i] + 0.5f)), cy = static_cast<int>(floorf(ys[i] + 0.5f));
  if (!isfinite(flux) || flux <= -1.0e30f || cx < 0 || cx >= width || cy < 0 || cy >= height) {
    if (lane == 0) status[i] = 0;
    return;
  }
  const int r = min(max(radius, 1), kCentroidMaxRadius);
  const int x0 = max(0, cx - r), x1 = min(width - 1, cx + r);
  const int y0 = max(0, cy - r), y1 = min(height - 1, cy + r);
  float background = background_override;
  if (!isfinite(background)) {
    if (lane == 0) {
      int n = 0;
      for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) {
          const float v = frame[static_cast<long long>(y) * width + x];
          if (isfinite(v)) win[n++] = v;
        }
      finite_count = n;
    }
    __syncthreads();
    const int n = finite_count;
    if (n <= 0) {
      if (lane == 0) status[i] = 0;
      return;
    }
    // the sample at sorted position p (stable order: equal values by window
    // position) has rank #(smaller) + #(equal and earlier)
    const int hi = n / 2, lo = (n & 1) ? -1 : n / 2 - 1;
    for (int a = lane; a < n; a += 64) {
      const float v = win[a];
      int rank = 0;
      for (int b = 0; b < n; ++b) {
        const float u = win[b];
        rank += (u < v || (u == v && b < a)) ? 1 : 0;
      }
      if (rank == hi) middle[1] = v;
      if (rank == lo) middle[0] = v;
    }
    __syncthreads();
    background = (n & 1) ? middle[1] : 0.5f * (middle[0] + middle[1]);
  }
  if (lane != 0) return;
  double sw = 0.0, sx = 0.0, sy = 0.0;
  for (int y = y0; y <= y1; ++y)
    for (int x = x0; x <= x1; ++x) {
      const float v = frame[static_cast<long long>(y) * width + x];
      if (!isfinite(v)) continue;
      const float wf = fmaxf(v - background, 0.0f);
      if (wf <= 0.0f) continue;
      const double w = static_cast<double>(wf);
      sw += w;
      sx += static_cast<double>(x) * w;
      sy += static_cast<double>(y) * w;
    }
  if (sw <= 0.0) {
    status[i] = 0;
    return;
  }
  xs[i] = static_cast<float>(sx / sw);
  ys[i] = static_cast<float>(sy / sw);
  status[i] = 1;
}

// catalog_offset_score_kernel (registration_kernels.cu:306-335): for each
// finite offset, the number of finite offsets (itself included) within
// float32 squared distance tol2; -1 for a non-finite one.  Offsets are staged
// through LDS 256 at a time.
__global__ __launch_bounds__(256) void offset_scores_kernel(const float* __restrict__ dx, const float* __restrict__ dy,
                                                            int n, float tol2, int* scores) {
  __shared__ float sx[256], sy[256];
  const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  const float mx = i < n ? dx[i] : NAN, my = i < n ? dy[i] : NAN;
  int score = 0;
  for (int base = 0; base < n; base += 256) {
    const int j = base + static_cast<int>(threadIdx.x);
    sx[threadIdx.x] = j < n ? dx[j] : NAN;
    sy[threadIdx.x] = j < n ? dy[j] : NAN;
    __syncthreads();
    const int m = min(256, n - base);
    for (int t = 0; t < m; ++t) {
      const float ox = sx[t], oy = sy[t];
      if (!isfinite(ox) || !isfinite(oy)) continue;
      const float ex = ox - mx, ey = oy - my;
      const float e2x = ex * ex, e2y = ey * ey;
      score += (e2x + e2y <= tol2) ? 1 : 0;
    }
    __syncthreads();
  }
  if (i < n) scores[i] = (isfinite(mx) && isfinite(my)) ? score : -1;
}

unsigned blocks_for(long long n, unsigned threads) {
  return static_cast<unsigned>(std::max<long long>(1, (n + threads - 1) / threads));
}

int bits_for(unsigned long long v) {
  int b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

}  // namespace

void launch_local_max_keys(const float* frame, int height, int width, float threshold, unsigned long long* keys,
                           unsigned capacity, unsigned* count, hipStream_t stream) {
  const long long inner = static_cast<long long>(width - 2) * (height - 2);
  const unsigned blocks = std::min<unsigned>(blocks_for(inner, 256), 8192u);
  local_max_keys_kernel<<<blocks, 256, 0, stream>>>(frame, height, width, threshold, keys, capacity, count);
  check(hipGetLastError(), "local maximum launch");
}

size_t sort_keys_temp_bytes(unsigned n) {
  size_t bytes = 0;
  check(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, static_cast<const unsigned long long*>(nullptr),
                                          static_cast<unsigned long long*>(nullptr), n),
        "radix sort size");
  return bytes;
}

void sort_keys(void* temp, size_t temp_bytes, const unsigned long long* in, unsigned long long* out, unsigned n,
               bool position_only, hipStream_t stream) {
  check(hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, in, out, n, 0, position_only ? 32 : 64, stream),
        "radix sort");
}

void launch_decode_keys(const unsigned long long* keys, unsigned n, const float* frame, int width, float* xs, float* ys,
                        float* fluxes, hipStream_t stream) {
  if (n == 0) return;
  decode_keys_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(keys, n, frame, width, xs, ys, fluxes);
  check(hipGetLastError(), "candidate decode launch");
}

size_t grid_select_temp_bytes(unsigned n, unsigned cells) {
  size_t pairs = 0, select = 0;
  check(hipcub::DeviceRadixSort::SortPairs(nullptr, pairs, static_cast<const unsigned*>(nullptr),
                                           static_cast<unsigned*>(nullptr), static_cast<const unsigned*>(nullptr),
                                           static_cast<unsigned*>(nullptr), n, 0, bits_for(cells)),
        "cell sort size");
  check(hipcub::DeviceSelect::Flagged(nullptr, select, static_cast<const unsigned long long*>(nullptr),
                                      static_cast<const unsigned char*>(nullptr),
                                      static_cast<unsigned long long*>(nullptr), static_cast<unsigned*>(nullptr), n),
        "cell select size");
  return std::max(pairs, select);
}

void grid_select(const unsigned long long* ranked, unsigned n, int width, int height, int cols, int rows,
                 unsigned per_cell, const GridSelectScratch& w, unsigned long long* out, unsigned* out_count,
                 hipStream_t stream) {
  const unsigned cells = static_cast<unsigned>(cols) * static_cast<unsigned>(rows);
  cell_of_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(ranked, n, width, height, cols, rows, w.cell, w.pos);
  check(hipGetLastError(), "cell launch");
  size_t bytes = w.temp_bytes;
  check(hipcub::DeviceRadixSort::SortPairs(w.temp, bytes, w.cell, w.cell_sorted, w.pos, w.pos_sorted, n, 0,
                                           bits_for(cells), stream),
        "cell sort");
  cell_start_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(w.cell_sorted, n, w.start);
  check(hipGetLastError(), "cell start launch");
  cell_cut_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(w.cell_sorted, w.pos_sorted, n, w.start, per_cell, w.keep);
  check(hipGetLastError(), "cell cut launch");
  bytes = w.temp_bytes;
  check(hipcub::DeviceSelect::Flagged(w.temp, bytes, ranked, w.keep, out, out_count, n, stream), "cell select");
}

void launch_offset_scores(const float* dx, const float* dy, int n, float tol2, int* scores, hipStream_t stream) {
  if (n <= 0) return;
  offset_scores_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(dx, dy, n, tol2, scores);
  check(hipGetLastError(), "offset vote launch");
}

void launch_refine_centroids(const float* frame, int width, int height, float* xs, float* ys, const float* fluxes,
                             unsigned char* status, int count, int radius, float background, hipStream_t stream) {
  if (count <= 0) return;
  refine_centroids_kernel<<<count, 64, 0, stream>>>(frame, width, height, xs, ys, fluxes, status, radius,
                                                    background);
  check(hipGetLastError(), "centroid refine launch");
}

}  // namespace gh

namespace gh {
namespace {

// glass_matrix_alignment_metrics_candidates_f32_kernel
// (warp_kernels.cu:1051-1172): per candidate, over the reference's pixels on
// the stride grid, the moving frame sampled bilinearly through the float32
// inverse (pixels whose source falls outside [0, W-1] x [0, H-1], or with a
// non-finite reference or sample, are skipped) and seven FP64 sums (r, m, r^2,
// m^2, r*m, (m-r)^2, |m-r|) plus the count.  The reference gives each
// candidate one block; here a candidate's sample rows are split over
// `chunks` blocks so that a handful of candidates still fills the chip, and
// the (candidate, chunk) partials are added on the host in chunk order.
// Each candidate names its moving frame, so the candidates of many frames
// (the batch form) share one launch.
__global__ __launch_bounds__(256) void alignment_partials_kernel(AlignArgs a, double* partials) {
  __shared__ double red[4][8];
  const int cand = static_cast<int>(blockIdx.x) / a.chunks, chunk = static_cast<int>(blockIdx.x) % a.chunks;
  float inv[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) inv[k] = a.inverses[9 * cand + k];
  const int r0 = chunk * a.rows_per_chunk, r1 = min(a.sample_h, r0 + a.rows_per_chunk);
  const long long n = static_cast<long long>(max(r1 - r0, 0)) * a.sample_w;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long long i = threadIdx.x; i < n; i += 256) {
    const int sr = r0 + static_cast<int>(i / a.sample_w), sc = static_cast<int>(i % a.sample_w);
    const int x = sc * a.stride, y = sr * a.stride;
    const float fx = static_cast<float>(x), fy = static_cast<float>(y);
    const float denom = inv[6] * fx + inv[7] * fy + inv[8];
    if (fabsf(denom) <= 1.0e-12f) continue;
    const float sx = (inv[0] * fx + inv[1] * fy + inv[2]) / denom;
    const float sy = (inv[3] * fx + inv[4] * fy + inv[5]) / denom;
    if (!(sx >= 0.0f) || sx > static_cast<float>(a.width - 1) || !(sy >= 0.0f) ||
        sy > static_cast<float>(a.height - 1))
      continue;
    const int x0 = static_cast<int>(floorf(sx)), y0 = static_cast<int>(floorf(sy));
    const int x1 = x0 + 1 < a.width ? x0 + 1 : x0, y1 = y0 + 1 < a.height ? y0 + 1 : y0;
    const float tx = sx - static_cast<float>(x0), ty = sy - static_cast<float>(y0);
    const float* mv = table_at(a.moving, cand);
    const float v00 = mv[static_cast<long long>(y0) * a.width + x0], v10 = mv[static_cast<long long>(y0) * a.width + x1];
    const float v01 = mv[static_cast<long long>(y1) * a.width + x0], v11 = mv[static_cast<long long>(y1) * a.width + x1];
    const float top = v00 * (1.0f - tx) + v10 * tx;
    const float bottom = v01 * (1.0f - tx) + v11 * tx;
    const float warped = top * (1.0f - ty) + bottom * ty;
    const float ref = a.reference[static_cast<long long>(y) * a.width + x];
    if (!isfinite(ref) || !isfinite(warped)) continue;
    const double r = ref, m = warped, d = m - r;
    s[0] += r;
    s[1] += m;
    s[2] += r * r;
    s[3] += m * m;
    s[4] += r * m;
    s[5] += d * d;
    s[6] += fabs(d);
    s[7] += 1.0;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    for (int off = 32; off > 0; off >>= 1) s[k] += __shfl_down(s[k], off, 64);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 8; ++k) red[wave][k] = s[k];
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int k = threadIdx.x;
    partials[static_cast<long long>(blockIdx.x) * 8 + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
  }
}

}  // namespace

void launch_alignment_partials(const AlignArgs& a, double* partials, hipStream_t stream) {
  if (a.candidates <= 0) return;
  alignment_partials_kernel<<<static_cast<unsigned>(a.candidates) * static_cast<unsigned>(a.chunks), 256, 0,
                              stream>>>(a, partials);
  check(hipGetLastError(), "alignment metrics launch");
}

}  // namespace gh
