// Star-candidate detection kernels for gfx950 (src/glass/cpu/star_detect.py:
// 148-160): a core pixel is a candidate when it is finite, above the float32
// threshold and not below any of its 8 neighbours (a NaN neighbour fails the
// comparison, as in numpy).  Candidates are compacted with one atomic per
// wave; their windows are then gathered for the host measurement.

#include "common.hpp"
#include "stars.hpp"

namespace gh {
namespace {

__global__ __launch_bounds__(kBlock) void star_candidates_kernel(const float* __restrict__ frame, int height, int width,
                                                                 float threshold, unsigned* cand, unsigned capacity,
                                                                 unsigned* count) {
  const long long inner_w = width - 2, inner = inner_w * static_cast<long long>(height - 2);
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i - threadIdx.x < inner;
       i += stride) {
    bool hit = false;
    long long pix = 0;
    if (i < inner) {
      const int y = static_cast<int>(i / inner_w) + 1, x = static_cast<int>(i % inner_w) + 1;
      pix = static_cast<long long>(y) * width + x;
      const float v = frame[pix];
      hit = isfinite(v) && v > threshold;
      if (hit) {
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            if (dx || dy) hit = hit && (v >= frame[pix + static_cast<long long>(dy) * width + dx]);
          }
        }
      }
    }
    // One atomic per wave: lanes take consecutive slots from the wave's base.
    const unsigned long long mask = __ballot(hit);
    if (mask) {
      const int lane = threadIdx.x & 63;
      const unsigned total = __popcll(mask);
      unsigned base = 0;
      if (lane == __ffsll(static_cast<long long>(mask)) - 1) base = atomicAdd(count, total);
      base = __shfl(base, __ffsll(static_cast<long long>(mask)) - 1, 64);
      if (hit) {
        const unsigned slot = base + __popcll(mask & ((1ull << lane) - 1ull));
        if (slot < capacity) cand[slot] = static_cast<unsigned>(pix);
      }
    }
  }
}

__global__ void gather_windows_kernel(const float* __restrict__ frame, int height, int width,
                                      const unsigned* __restrict__ cand, unsigned n, int radius, float* windows) {
  const int side = 2 * radius + 1, cells = side * side;
  const long long total = static_cast<long long>(n) * cells;
  for (long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long k = t / cells;
    const int c = static_cast<int>(t - k * cells);
    const unsigned pix = cand[k];
    const int y = static_cast<int>(pix / static_cast<unsigned>(width)) + c / side - radius;
    const int x = static_cast<int>(pix % static_cast<unsigned>(width)) + c % side - radius;
    windows[t] = (y >= 0 && y < height && x >= 0 && x < width) ? frame[static_cast<long long>(y) * width + x] : 0.0f;
  }
}

}  // namespace

void launch_star_candidates(const float* frame, int height, int width, float threshold, unsigned* cand,
                            unsigned capacity, unsigned* count, hipStream_t stream) {
  const long long inner = static_cast<long long>(width - 2) * (height - 2);
  long long blocks = (inner + kBlock - 1) / kBlock;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  star_candidates_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, stream>>>(frame, height, width, threshold, cand,
                                                                               capacity, count);
  check(hipGetLastError(), "star candidate launch");
}

void launch_gather_windows(const float* frame, int height, int width, const unsigned* cand, unsigned n, int radius,
                           float* windows, hipStream_t stream) {
  if (n == 0) return;
  const long long total = static_cast<long long>(n) * (2 * radius + 1) * (2 * radius + 1);
  long long blocks = (total + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  gather_windows_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, stream>>>(frame, height, width, cand, n, radius,
                                                                             windows);
  check(hipGetLastError(), "window gather launch");
}

}  // namespace gh
