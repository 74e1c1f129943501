// Lanczos3 affine warp with its adjoint, for photometric registration
// refinement on gfx950.
//
// The warp is the CPU-path Lanczos3 of src/glass/engine/warp.py:121-228 for an
// affine source map s = B (x, y, 1) (B: 2x3, reference pixel -> moving pixel,
// FP64 in the CPU rounding order): taps floor(s)-2 .. floor(s)+3, weights
// L(d) = sinc(d) sinc(d/3) normalised by their sum, valid 2 <= s < dim-3,
// `fill` elsewhere.  There is no reference implementation of the derivatives
// (they are pinned by finite differences and closed forms); here:
//
//   out   = sum_j wy_j sum_i wx_i S_ji,       w_k = L(d_k) / sum L
//   dw_k/df = (L'(d_k) - w_k sum L') / sum L,  d_k = f - (k - 2)
//   d out / d s_x = sum_j wy_j sum_i dwx_i S_ji   (and likewise for s_y)
//   d out / d B   = (ds_x * (x, y, 1), ds_y * (x, y, 1))
//   d out / d S_ji = wx_i wy_j                   (the pixel adjoint)
//
// Backward recomputes each pixel's stencil (no stored per-pixel weights); the
// pixel adjoint scatters with f32 atomics, parameter gradients and the
// Gauss-Newton normal equations reduce through FP64 per-block partials summed
// in a fixed order.

#include "adjoint.hpp"
#include "common.hpp"

namespace gh {
namespace {

constexpr int kAdjBlock = 256;

// One axis: normalised weights w[6] and dw/df for fractional offset f.
// sin(pi d_k) = (-1)^k sin(pi f); sin(pi d_k / 3) = s3 ck - c3 sk with
// (ck, sk) = (cos, sin)(pi (k-2) / 3); cos likewise.  One sincos of pi f / 3
// gives sin / cos of pi f by the triple-angle identities, and the divisions
// are hardware reciprocals (about 1 ulp): these weights feed derivatives and
// the Gauss-Newton normal equations, pinned by finite differences.
__device__ __forceinline__ float rcp_approx(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ void lanczos_weights_derivs(float f, float (&w)[6], float (&dw)[6]) {
  float s3, c3;
  sincospif(f * (1.0f / 3.0f), &s3, &c3);
  const float s1 = s3 * (3.0f - 4.0f * s3 * s3);  // sin(3a) = 3 sin a - 4 sin^3 a
  const float c1 = c3 * (4.0f * c3 * c3 - 3.0f);  // cos(3a) = 4 cos^3 a - 3 cos a
  constexpr float kHalfRoot3 = 0.86602540378443864676f;
  constexpr float kPi = 3.14159265358979323846f;
  const float ck[6] = {-0.5f, 0.5f, 1.0f, 0.5f, -0.5f, -1.0f};
  const float sk[6] = {-kHalfRoot3, -kHalfRoot3, 0.0f, kHalfRoot3, kHalfRoot3, 0.0f};
  float l[6], dl[6], sum = 0.0f, dsum = 0.0f;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const float d = f - static_cast<float>(k - 2);
    const float sgn = (k & 1) ? -1.0f : 1.0f;
    const float sa = sgn * s1, ca = sgn * c1;  // sin, cos of pi d
    const float sb = s3 * ck[k] - c3 * sk[k];  // sin(pi d / 3)
    const float cb = c3 * ck[k] + s3 * sk[k];  // cos(pi d / 3)
    if (fabsf(d) < 1.0e-4f) {
      // L = 1 - (10/9)(pi d)^2 / 6 + ..., L' = -(10/27) pi^2 d
      l[k] = 1.0f;
      dl[k] = -(10.0f / 27.0f) * kPi * kPi * d;
    } else {
      const float rd = rcp_approx(d);
      const float inv = (1.0f / (kPi * kPi)) * rd * rd;
      l[k] = 3.0f * sa * sb * inv;
      dl[k] = 3.0f * inv * (kPi * ca * sb + (kPi / 3.0f) * sa * cb) - 2.0f * l[k] * rd;
    }
    sum += l[k];
    dsum += dl[k];
  }
  const float rs = rcp_approx(sum);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    w[k] = l[k] * rs;
    dw[k] = (dl[k] - w[k] * dsum) * rs;
  }
}

struct Stencil {
  bool ok;
  int x0, y0;
  float wx[6], wy[6], dwx[6], dwy[6];
};

__device__ __forceinline__ Stencil stencil_at(const AffineMap& m, int x, int y, int h, int w) {
  Stencil st;
  const double xs = static_cast<double>(x), ys = static_cast<double>(y);
  double sx = m.b[0] * xs + m.b[1] * ys;
  sx = sx + m.b[2];
  double sy = m.b[3] * xs + m.b[4] * ys;
  sy = sy + m.b[5];
  st.ok = sx >= 2.0 && sx < static_cast<double>(w - 3) && sy >= 2.0 && sy < static_cast<double>(h - 3);
  if (!st.ok) return st;
  const double fx0 = floor(sx), fy0 = floor(sy);
  st.x0 = static_cast<int>(fx0);
  st.y0 = static_cast<int>(fy0);
  lanczos_weights_derivs(static_cast<float>(sx - fx0), st.wx, st.dwx);
  lanczos_weights_derivs(static_cast<float>(sy - fy0), st.wy, st.dwy);
  return st;
}

// value, d value / d s_x, d value / d s_y from the 36 taps.
__device__ __forceinline__ void evaluate(const float* src, int w, const Stencil& st, float& v, float& gx, float& gy) {
  const float* t = src + static_cast<long long>(st.y0 - 2) * w + (st.x0 - 2);
  v = gx = gy = 0.0f;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const float* r = t + static_cast<long long>(j) * w;
    float h = 0.0f, hd = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      h = fmaf(r[i], st.wx[i], h);
      hd = fmaf(r[i], st.dwx[i], hd);
    }
    v = fmaf(h, st.wy[j], v);
    gx = fmaf(hd, st.wy[j], gx);
    gy = fmaf(h, st.dwy[j], gy);
  }
}

template <int K>
__device__ __forceinline__ void block_reduce_store(double (&acc)[K], double* partials) {
  __shared__ double red[kAdjBlock / 64][K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    acc[k] = v;
  }
  const int wave = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[wave][k] = acc[k];
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int wv = 0; wv < kAdjBlock / 64; ++wv) s += red[wv][threadIdx.x];
    partials[static_cast<long long>(blockIdx.x) * K + threadIdx.x] = s;
  }
}

__global__ __launch_bounds__(kAdjBlock) void forward_kernel(const float* __restrict__ src, int h, int w, AffineMap m,
                                                           float fill, float* __restrict__ out) {
  // Blocks stride over rows, threads over columns: no per-pixel division.
  for (int y = blockIdx.x; y < h; y += gridDim.x)
  for (int x = threadIdx.x; x < w; x += kAdjBlock) {
    const long long p = static_cast<long long>(y) * w + x;
    const Stencil st = stencil_at(m, x, y, h, w);
    float v = fill;
    if (st.ok) {
      float gx, gy;
      evaluate(src, w, st, v, gx, gy);
    }
    out[p] = v;
  }
}

__global__ __launch_bounds__(kAdjBlock) void backward_kernel(const float* __restrict__ src, int h, int w, AffineMap m,
                                                            const float* __restrict__ grad_out, float* grad_src,
                                                            double* partials) {
  double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int y = blockIdx.x; y < h; y += gridDim.x)
  for (int x = threadIdx.x; x < w; x += kAdjBlock) {
    const long long p = static_cast<long long>(y) * w + x;
    const float g = grad_out[p];
    if (g == 0.0f) continue;
    const Stencil st = stencil_at(m, x, y, h, w);
    if (!st.ok) continue;
    if (grad_src) {
      float* t = grad_src + static_cast<long long>(st.y0 - 2) * w + (st.x0 - 2);
#pragma unroll
      for (int j = 0; j < 6; ++j) {
#pragma unroll
        for (int i = 0; i < 6; ++i) atomicAdd(t + static_cast<long long>(j) * w + i, g * (st.wx[i] * st.wy[j]));
      }
    }
    if (partials) {
      float v, gx, gy;
      evaluate(src, w, st, v, gx, gy);
      const double dx = static_cast<double>(g) * gx, dy = static_cast<double>(g) * gy;
      acc[0] += dx * x;
      acc[1] += dx * y;
      acc[2] += dx;
      acc[3] += dy * x;
      acc[4] += dy * y;
      acc[5] += dy;
    }
  }
  if (partials) block_reduce_store<6>(acc, partials);
}

// Gauss-Newton normal equations of sum_p (warp(moving; B)(p) - reference(p))^2
// over the pixels inside the footprint with finite values: J^T J (upper
// triangle, 21), J^T r (6), sum r^2, count -> 29 sums per block.  Pixel
// coordinates enter the Jacobian centred on (cx, cy) for conditioning.
__global__ __launch_bounds__(kAdjBlock) void normal_equations_kernel(const float* __restrict__ ref,
                                                                    const float* __restrict__ mov, int h, int w,
                                                                    AffineMap m, double cx, double cy,
                                                                    double* partials) {
  double acc[kNormalSums];
#pragma unroll
  for (int k = 0; k < kNormalSums; ++k) acc[k] = 0.0;
  // Blocks stride over rows, threads over columns: no per-pixel division.
  for (int y = blockIdx.x; y < h; y += gridDim.x)
  for (int x = threadIdx.x; x < w; x += kAdjBlock) {
    const long long p = static_cast<long long>(y) * w + x;
    const Stencil st = stencil_at(m, x, y, h, w);
    if (!st.ok) continue;
    const float r0 = ref[p];
    float v, gx, gy;
    evaluate(mov, w, st, v, gx, gy);
    if (!isfinite(r0) || !isfinite(v) || !isfinite(gx) || !isfinite(gy)) continue;
    const double xc = static_cast<double>(x) - cx, yc = static_cast<double>(y) - cy;
    const double j[6] = {gx * xc, gx * yc, static_cast<double>(gx), gy * xc, gy * yc, static_cast<double>(gy)};
    const double r = static_cast<double>(v) - static_cast<double>(r0);
    // fused multiply-adds: one rounding per term (the sums are pinned by
    // tolerances, not bits)
    int t = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
#pragma unroll
      for (int b = a; b < 6; ++b) {
        acc[t] = fma(j[a], j[b], acc[t]);
        ++t;
      }
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) acc[21 + a] = fma(j[a], r, acc[21 + a]);
    acc[27] = fma(r, r, acc[27]);
    acc[28] += 1.0;
  }
  block_reduce_store<kNormalSums>(acc, partials);
}

// Column c of the per-block partials (blocks rows of k sums) in block c:
// thread t sums rows t, t + 256, ..., then a fixed-order tree in LDS, so the
// result is deterministic and the sums take a few load latencies, not one per
// row.
constexpr int kReduceThreads = 256;
__global__ __launch_bounds__(kReduceThreads) void reduce_partials_kernel(const double* partials, int blocks, int k,
                                                                         double* out) {
  __shared__ double part[kReduceThreads];
  const int c = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < blocks; b += kReduceThreads) v += partials[static_cast<long long>(b) * k + c];
  part[threadIdx.x] = v;
  __syncthreads();
  for (int width = kReduceThreads / 2; width > 0; width >>= 1) {
    if (static_cast<int>(threadIdx.x) < width) part[threadIdx.x] += part[threadIdx.x + width];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = part[0];
}

unsigned adjoint_blocks(long long n) {
  long long b = (n + kAdjBlock - 1) / kAdjBlock;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return static_cast<unsigned>(b);
}

}  // namespace

size_t adjoint_workspace_doubles(long long npix) {
  return static_cast<size_t>(adjoint_blocks(npix)) * kNormalSums + kNormalSums;
}

void launch_affine_forward(const float* src, int h, int w, const AffineMap& m, float fill, float* out,
                           hipStream_t stream) {
  forward_kernel<<<adjoint_blocks(static_cast<long long>(h) * w), kAdjBlock, 0, stream>>>(src, h, w, m, fill, out);
  check(hipGetLastError(), "affine warp forward launch");
}

void launch_affine_backward(const float* src, int h, int w, const AffineMap& m, const float* grad_out,
                            float* grad_src, double* workspace, double* grad_map, hipStream_t stream) {
  const unsigned blocks = adjoint_blocks(static_cast<long long>(h) * w);
  backward_kernel<<<blocks, kAdjBlock, 0, stream>>>(src, h, w, m, grad_out, grad_src, grad_map ? workspace : nullptr);
  if (grad_map) reduce_partials_kernel<<<6, kReduceThreads, 0, stream>>>(workspace, blocks, 6, grad_map);
  check(hipGetLastError(), "affine warp backward launch");
}

void launch_normal_equations(const float* ref, const float* mov, int h, int w, const AffineMap& m, double cx,
                             double cy, double* workspace, double* sums, hipStream_t stream) {
  const unsigned blocks = adjoint_blocks(static_cast<long long>(h) * w);
  normal_equations_kernel<<<blocks, kAdjBlock, 0, stream>>>(ref, mov, h, w, m, cx, cy, workspace);
  reduce_partials_kernel<<<kNormalSums, kReduceThreads, 0, stream>>>(workspace, blocks, kNormalSums, sums);
  check(hipGetLastError(), "normal equations launch");
}

}  // namespace gh
