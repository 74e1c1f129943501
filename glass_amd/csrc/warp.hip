// Affine matrix warp (bilinear / Lanczos3) of resident frames for gfx950.
//
// Semantics follow the CPU oracle src/glass/engine/warp.py:155-228:
//  * source coordinates in FP64, ((i00*x + i01*y) + i02) without contraction,
//    divided by the projective denominator unless the inverse is affine;
//  * bilinear valid 0 <= s <= dim-1, f32 lerps as warp.py:99-118;
//  * Lanczos3 valid 2 <= s < dim-3, taps floor(s)-2 .. floor(s)+3
//    (warp.py:121-152); weights sinc(d)*sinc(d/3) evaluated in f32 from one
//    sin/cos pair per axis (shifted-sinc identity, see lanczos_axis);
//  * GH_WARP_SKIP_NONFINITE_TAPS / clamp: resident semantics of
//    cpp/cuda/warp_kernels.cu:643-777 as pinned by
//    tests/test_gpu_warp_vs_cpu.py:80-128.
//
// One thread owns one output pixel and walks the frames of the batch, so the
// geometric coverage accumulator is updated with one plain read-modify-write
// per pixel per launch (no atomics).  Frames are written out of place into a
// spare slot; the host swaps slots afterwards.

#include "common.hpp"
#include "kernels.hpp"

namespace gh {
namespace {

constexpr int kTileX = 64;
constexpr int kTileY = 4;

__device__ __forceinline__ bool source_coords(const double* m, int affine, int x, int y, double& sx, double& sy) {
  const double xs = static_cast<double>(x);
  const double ys = static_cast<double>(y);
  double nx = m[0] * xs + m[1] * ys;
  nx = nx + m[2];
  double ny = m[3] * xs + m[4] * ys;
  ny = ny + m[5];
  if (affine) {
    sx = nx;
    sy = ny;
    return true;
  }
  double den = m[6] * xs + m[7] * ys;
  den = den + m[8];
  if (!(fabs(den) > 1.0e-12)) return false;
  sx = nx / den;
  sy = ny / den;
  return true;
}

// Six Lanczos3 weights for taps at floor(s) + k, k = -2..3, fractional part f.
// With d = f - k:  sin(pi d) = (-1)^k sin(pi f)  and
// sin(pi d / 3) = sin(pi f/3) cos(pi k/3) - cos(pi f/3) sin(pi k/3),
// so one sincos of pi f / 3 gives every tap:
//   w_k = 3 sin(pi d) sin(pi d / 3) / (pi^2 d^2),   w_k = 1 when |d| < thr.
__device__ __forceinline__ void lanczos_axis(float f, float thr, float (&w)[6]) {
  float s3, c3;
  sincospif(f * (1.0f / 3.0f), &s3, &c3);
  const float sf = s3 * (3.0f - 4.0f * s3 * s3);  // sin(pi f) = sin(3 * pi f / 3)
  constexpr float kHalfRoot3 = 0.86602540378443864676f;
  // cos(pi k/3), sin(pi k/3) for k = -2..3
  const float ck[6] = {-0.5f, 0.5f, 1.0f, 0.5f, -0.5f, -1.0f};
  const float sk[6] = {-kHalfRoot3, -kHalfRoot3, 0.0f, kHalfRoot3, kHalfRoot3, 0.0f};
  const float sgn[6] = {1.0f, -1.0f, 1.0f, -1.0f, 1.0f, -1.0f};
  constexpr float kScale = 3.0f / (3.14159265358979323846f * 3.14159265358979323846f);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const float d = f - static_cast<float>(i - 2);
    const float num = sgn[i] * sf * (s3 * ck[i] - c3 * sk[i]);
    w[i] = (fabsf(d) < thr) ? 1.0f : kScale * num / (d * d);
  }
}

template <bool AFFINE>
__global__ __launch_bounds__(kTileX* kTileY) void warp_bilinear_kernel(WarpBatch b, int count, WarpParams p) {
  const int x = blockIdx.x * kTileX + (threadIdx.x % kTileX);
  const int y = blockIdx.y * kTileY + (threadIdx.x / kTileX);
  if (x >= p.width || y >= p.height) return;
  const long long o = static_cast<long long>(y) * p.width + x;
  float cov = 0.0f;
  for (int f = 0; f < count; ++f) {
    double sx, sy;
    bool ok = source_coords(b.inv[f], AFFINE, x, y, sx, sy);
    ok = ok && isfinite(sx) && isfinite(sy) && sx >= 0.0 && sx <= static_cast<double>(p.width - 1) && sy >= 0.0 &&
         sy <= static_cast<double>(p.height - 1);
    float value = p.fill;
    if (ok) {
      const double fx0 = floor(sx), fy0 = floor(sy);
      const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0);
      const int x1 = min(x0 + 1, p.width - 1), y1 = min(y0 + 1, p.height - 1);
      const float tx = static_cast<float>(sx - fx0);
      const float ty = static_cast<float>(sy - fy0);
      const float* src = b.src[f];
      const float a00 = src[static_cast<long long>(y0) * p.width + x0];
      const float a01 = src[static_cast<long long>(y0) * p.width + x1];
      const float a10 = src[static_cast<long long>(y1) * p.width + x0];
      const float a11 = src[static_cast<long long>(y1) * p.width + x1];
      const float top = a00 * (1.0f - tx) + a01 * tx;
      const float bottom = a10 * (1.0f - tx) + a11 * tx;
      value = top * (1.0f - ty) + bottom * ty;
      cov += 1.0f;
    }
    b.dst[f][o] = value;
    if (!isfinite(value) && __ldg(b.nonfinite[f]) == 0u) atomicOr(b.nonfinite[f], 1u);
  }
  if (p.coverage) p.coverage[o] = p.coverage[o] + cov;
}

template <bool AFFINE>
__global__ __launch_bounds__(kTileX* kTileY) void warp_lanczos3_kernel(WarpBatch b, int count, WarpParams p) {
  const int x = blockIdx.x * kTileX + (threadIdx.x % kTileX);
  const int y = blockIdx.y * kTileY + (threadIdx.x / kTileX);
  if (x >= p.width || y >= p.height) return;
  const long long o = static_cast<long long>(y) * p.width + x;
  const float thr = p.skip_nonfinite ? 1.0e-6f : 1.0e-8f;
  float cov = 0.0f;
  for (int f = 0; f < count; ++f) {
    double sx, sy;
    bool ok = source_coords(b.inv[f], AFFINE, x, y, sx, sy);
    ok = ok && isfinite(sx) && isfinite(sy) && sx >= 2.0 && sx < static_cast<double>(p.width - 3) && sy >= 2.0 &&
         sy < static_cast<double>(p.height - 3);
    float value = p.fill;
    bool covered = false;
    if (ok) {
      const double fx0 = floor(sx), fy0 = floor(sy);
      const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0);
      float wx[6], wy[6];
      lanczos_axis(static_cast<float>(sx - fx0), thr, wx);
      lanczos_axis(static_cast<float>(sy - fy0), thr, wy);
      const float* src = b.src[f] + static_cast<long long>(y0 - 2) * p.width + (x0 - 2);
      float acc = 0.0f, wsum = 0.0f;
      float vmin = INFINITY, vmax = -INFINITY;
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const float* row = src + static_cast<long long>(j) * p.width;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          const float v = row[i];
          const float w = wy[j] * wx[i];
          if (p.skip_nonfinite) {
            if (isfinite(v)) {
              acc = fmaf(v, w, acc);
              wsum += w;
              vmin = fminf(vmin, v);
              vmax = fmaxf(vmax, v);
            }
          } else if (wx[i] != 0.0f && wy[j] != 0.0f) {
            acc = fmaf(v, w, acc);
            wsum += w;
          }
        }
      }
      if (p.skip_nonfinite) {
        if (fabsf(wsum) > 1.0e-12f) {
          value = acc / wsum;
          if (p.clamp >= 0.0f && vmin <= vmax) {
            const float r = vmax - vmin;
            value = fminf(vmax + p.clamp * r, fmaxf(vmin - p.clamp * r, value));
          }
          covered = true;
        }
      } else {
        value = (fabsf(wsum) > 1.0e-12f) ? acc / wsum : 0.0f;
        covered = true;
      }
    }
    if (covered) cov += 1.0f;
    b.dst[f][o] = value;
    if (!isfinite(value) && __ldg(b.nonfinite[f]) == 0u) atomicOr(b.nonfinite[f], 1u);
  }
  if (p.coverage) p.coverage[o] = p.coverage[o] + cov;
}

}  // namespace

void launch_warp(const WarpBatch& batch, int count, const WarpParams& p, hipStream_t stream) {
  const dim3 grid((p.width + kTileX - 1) / kTileX, (p.height + kTileY - 1) / kTileY);
  const dim3 block(kTileX * kTileY);
  if (p.interpolation == 1) {
    if (p.affine) warp_bilinear_kernel<true><<<grid, block, 0, stream>>>(batch, count, p);
    else warp_bilinear_kernel<false><<<grid, block, 0, stream>>>(batch, count, p);
  } else {
    if (p.affine) warp_lanczos3_kernel<true><<<grid, block, 0, stream>>>(batch, count, p);
    else warp_lanczos3_kernel<false><<<grid, block, 0, stream>>>(batch, count, p);
  }
  check(hipGetLastError(), "warp kernel launch");
}

}  // namespace gh
