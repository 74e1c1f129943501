// Translation search by normalised cross-correlation (gfx950), the resident
// registration's integer and subpixel translation estimates
// (cpp/cuda/registration_kernels.cu:10-264, driven by
// cpp/src/native_bindings.cpp:10418-10568).
//
// For each candidate offset (dx, dy) and every sampled reference pixel
// (x, y) = (i * stride, j * stride): the moving sample is moving[y - dy][x - dx]
// (integer candidates) or the float32 bilinear sample at (x - dx, y - dy)
// (subpixel candidates, valid 0 <= s <= dim - 1, far neighbour clamped); pairs
// with both values finite enter sums of r, m, r^2, m^2, r*m and the count.
// The score is (sum rm - sum r sum m / n) / sqrt(var_r var_m) with the
// variances clamped at 0, or -FLT_MAX when n <= 1 or the denominator is 0.
// Sums are FP64 (the reference accumulates in float32): a grid of
// (candidate, row chunk) blocks writes FP64 partials, and one thread per
// candidate adds its chunks in order, so every score is deterministic.

#include <algorithm>
#include <cfloat>

#include "common.hpp"
#include "kernels.hpp"

namespace gh {
namespace {

constexpr int kNccThreads = 256;
constexpr int kNccSums = 6;

__device__ __forceinline__ double ncc_wave_sum(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

template <bool SUBPIXEL>
__global__ __launch_bounds__(kNccThreads) void ncc_partials_kernel(NccArgs a, double* partials) {
  __shared__ double part[kNccThreads / 64][kNccSums];
  const int cand = blockIdx.x, chunk = blockIdx.y, chunks = gridDim.y;
  const int cw = 2 * a.radius_x + 1;
  const int ox = cand % cw - a.radius_x, oy = cand / cw - a.radius_y;
  const int sample_w = (a.width + a.stride - 1) / a.stride;
  const int sample_h = (a.height + a.stride - 1) / a.stride;
  // this block's sampled rows
  const int r0 = static_cast<int>(static_cast<long long>(sample_h) * chunk / chunks);
  const int r1 = static_cast<int>(static_cast<long long>(sample_h) * (chunk + 1) / chunks);
  const float fdx = a.center_dx + static_cast<float>(ox) * a.step;
  const float fdy = a.center_dy + static_cast<float>(oy) * a.step;
  double s[kNccSums] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int j = r0; j < r1; ++j) {
    const int y = j * a.stride;
    for (int i = threadIdx.x; i < sample_w; i += kNccThreads) {
      const int x = i * a.stride;
      float m;
      if (!SUBPIXEL) {
        const int sx = x - ox, sy = y - oy;
        if (sx < 0 || sx >= a.width || sy < 0 || sy >= a.height) continue;
        m = ld_global(a.moving, static_cast<long long>(sy) * a.width + sx);
      } else {
        const float sx = static_cast<float>(x) - fdx, sy = static_cast<float>(y) - fdy;
        if (sx < 0.0f || sx > static_cast<float>(a.width - 1) || sy < 0.0f || sy > static_cast<float>(a.height - 1))
          continue;
        const int x0 = static_cast<int>(floorf(sx)), y0 = static_cast<int>(floorf(sy));
        const int x1 = x0 + 1 < a.width ? x0 + 1 : x0, y1 = y0 + 1 < a.height ? y0 + 1 : y0;
        const float tx = sx - static_cast<float>(x0), ty = sy - static_cast<float>(y0);
        const float v00 = ld_global(a.moving, static_cast<long long>(y0) * a.width + x0);
        const float v10 = ld_global(a.moving, static_cast<long long>(y0) * a.width + x1);
        const float v01 = ld_global(a.moving, static_cast<long long>(y1) * a.width + x0);
        const float v11 = ld_global(a.moving, static_cast<long long>(y1) * a.width + x1);
        const float top = v00 * (1.0f - tx) + v10 * tx;
        const float bottom = v01 * (1.0f - tx) + v11 * tx;
        m = top * (1.0f - ty) + bottom * ty;
      }
      const float r = ld_global(a.reference, static_cast<long long>(y) * a.width + x);
      if (!isfinite(r) || !isfinite(m)) continue;
      const double rd = static_cast<double>(r), md = static_cast<double>(m);
      s[0] += rd;
      s[1] += md;
      s[2] += rd * rd;
      s[3] += md * md;
      s[4] += rd * md;
      s[5] += 1.0;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kNccSums; ++k) {
    const double v = ncc_wave_sum(s[k]);
    if (lane == 0) part[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < kNccSums) {
    double v = part[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kNccThreads / 64; ++w) v += part[w][threadIdx.x];
    partials[(static_cast<long long>(cand) * chunks + chunk) * kNccSums + threadIdx.x] = v;
  }
}

__global__ void ncc_scores_kernel(const double* partials, int candidates, int chunks, double* scores) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= candidates) return;
  double s[kNccSums] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k = 0; k < chunks; ++k)
    for (int q = 0; q < kNccSums; ++q) s[q] += partials[(static_cast<long long>(c) * chunks + k) * kNccSums + q];
  const double n = s[5];
  double score = -static_cast<double>(FLT_MAX);
  if (n > 1.0) {
    const double num = s[4] - s[0] * s[1] / n;
    const double vr = fmax(s[2] - s[0] * s[0] / n, 0.0);
    const double vm = fmax(s[3] - s[1] * s[1] / n, 0.0);
    const double den = sqrt(vr * vm);
    if (den > 0.0) score = num / den;
  }
  scores[c] = score;
}

}  // namespace

int ncc_chunks(const NccArgs& a) {
  const int sample_h = (a.height + a.stride - 1) / a.stride;
  const int candidates = (2 * a.radius_x + 1) * (2 * a.radius_y + 1);
  // enough blocks to fill the chip, at least 16 sampled rows per block
  int chunks = std::max(1, std::min(sample_h / 16, (4096 + candidates - 1) / candidates));
  return std::min(chunks, 65535);
}

void launch_ncc_search(const NccArgs& a, bool subpixel, double* partials, double* scores, hipStream_t stream) {
  const int candidates = (2 * a.radius_x + 1) * (2 * a.radius_y + 1);
  const int chunks = ncc_chunks(a);
  const dim3 grid(static_cast<unsigned>(candidates), static_cast<unsigned>(chunks));
  if (subpixel) ncc_partials_kernel<true><<<grid, kNccThreads, 0, stream>>>(a, partials);
  else ncc_partials_kernel<false><<<grid, kNccThreads, 0, stream>>>(a, partials);
  ncc_scores_kernel<<<(candidates + 255) / 256, 256, 0, stream>>>(partials, candidates, chunks, scores);
  check(hipGetLastError(), "translation search launch");
}

}  // namespace gh
