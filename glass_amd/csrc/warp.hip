// Affine matrix warp (bilinear / Lanczos3) of resident frames for gfx950.
//
// Semantics follow the CPU oracle src/glass/engine/warp.py:155-228:
//  * source coordinates in FP64, ((i00*x + i01*y) + i02) without contraction,
//    divided by the projective denominator unless the inverse is affine;
//  * bilinear valid 0 <= s <= dim-1, f32 lerps as warp.py:99-118;
//  * Lanczos3 valid 2 <= s < dim-3, taps floor(s)-2 .. floor(s)+3
//    (warp.py:121-152); weights sinc(d)*sinc(d/3), normalised by their sum;
//  * GH_WARP_SKIP_NONFINITE_TAPS / clamp: resident semantics of
//    cpp/cuda/warp_kernels.cu:643-777 as pinned by
//    tests/test_gpu_warp_vs_cpu.py:80-128.
//
// Affine Lanczos3 warps run warp_lanczos3_tiled_kernel (LDS-staged source
// boxes, tabulated normalised weights); bilinear and projective warps run the
// direct kernels, where one thread owns one output pixel and evaluates the
// weights from one sin/cos pair per axis (shifted-sinc identity, see
// lanczos_axis).  Every kernel walks the frames of the batch per output
// pixel, so the geometric coverage accumulator is updated with one plain
// read-modify-write per pixel per launch (no atomics).  Frames are written out
// of place into a spare slot; the host swaps slots afterwards.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <utility>

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

// One float at byte offset `off` of a raw buffer (out-of-range offsets read 0).
__device__ __forceinline__ float ld_rsrc_b32(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
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
    // hardware reciprocal (about 1 ulp) instead of a correctly rounded divide
    w[i] = (fabsf(d) < thr) ? 1.0f : (kScale * num) * __builtin_amdgcn_rcpf(d * d);
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


// Lanczos3 sample of one pixel, covering every semantic variant (CPU: no tap
// skipping, threshold 1e-8; resident: non-finite taps skipped, threshold
// 1e-6, optional clamp).  `src` points at the window's top-left tap (row
// floor(sy) - 2, column floor(sx) - 2) of rows `stride` floats apart: the
// frame in global memory, or a staged copy in LDS (same taps, same
// arithmetic, same value).  `covered` reports whether the pixel counts toward
// geometric coverage.
// CLAMP: -1 follows p.clamp at run time; 0 / 1 fix it (p.clamp < 0 / >= 0)
// at compile time, so the unclamped form carries no per-tap extremes.
template <typename P, int CLAMP = -1>
__device__ __forceinline__ float lanczos_direct_window(P src, int stride, double sx, double sy, double fx0, double fy0,
                                                       const WarpParams& p, float thr, bool& covered) {
  float wx[6], wy[6];
  lanczos_axis(static_cast<float>(sx - fx0), thr, wx);
  lanczos_axis(static_cast<float>(sy - fy0), thr, wy);
  float acc = 0.0f, wsum = 0.0f;
  float vmin = INFINITY, vmax = -INFINITY;
  // CPU semantics take a tap when its FP64 weight is non-zero
  // (warp.py:121-152).  In FP64, sin(pi d) never rounds to 0 for an integer
  // d != 0, so that weight is 0 only at |d| >= 3: the last tap (d = f - 3) of
  // an axis whose fraction f is exactly 0.  (The f32 weights here are exactly
  // 0 at every integer tap; testing them would stop NaN the reference keeps.)
  const bool last_x = sx != fx0, last_y = sy != fy0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const P row = src + static_cast<long long>(j) * stride;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const float v = row[i];
      const float w = wy[j] * wx[i];
      if (p.skip_nonfinite) {
        if (isfinite(v)) {
          acc = fmaf(v, w, acc);
          wsum += w;
          if constexpr (CLAMP != 0) {
            vmin = fminf(vmin, v);
            vmax = fmaxf(vmax, v);
          }
        }
      } else if ((i < 5 || last_x) && (j < 5 || last_y)) {
        acc = fmaf(v, w, acc);
        wsum += w;
      }
    }
  }
  covered = false;
  float value = p.fill;
  if (p.skip_nonfinite) {
    if (fabsf(wsum) > 1.0e-12f) {
      value = acc / wsum;
      if (CLAMP != 0 && p.clamp >= 0.0f && vmin <= vmax) {
        const float r = vmax - vmin;
        value = fminf(vmax + p.clamp * r, fmaxf(vmin - p.clamp * r, value));
      }
      covered = true;
    }
  } else {
    value = (fabsf(wsum) > 1.0e-12f) ? acc / wsum : 0.0f;
    covered = true;
  }
  return value;
}

// The sample read straight from the frame.
__device__ __forceinline__ float lanczos_direct(const float* frame, int width, double sx, double sy,
                                                const WarpParams& p, float thr, bool& covered) {
  const double fx0 = floor(sx), fy0 = floor(sy);
  const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0);
  const global_cf32* src = (const global_cf32*)(frame + static_cast<long long>(y0 - 2) * width + (x0 - 2));
  return lanczos_direct_window(src, width, sx, sy, fx0, fy0, p, thr, covered);
}

__device__ __forceinline__ bool lanczos_footprint(double sx, double sy, int width, int height) {
  return isfinite(sx) && isfinite(sy) && sx >= 2.0 && sx < static_cast<double>(width - 3) && sy >= 2.0 &&
         sy < static_cast<double>(height - 3);
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
    const bool ok = source_coords(b.inv[f], AFFINE, x, y, sx, sy) && lanczos_footprint(sx, sy, p.width, p.height);
    float value = p.fill;
    if (ok) {
      bool covered;
      value = lanczos_direct(b.src[f], p.width, sx, sy, p, thr, covered);
      if (covered) cov += 1.0f;
    }
    b.dst[f][o] = value;
    if (!isfinite(value) && __ldg(b.nonfinite[f]) == 0u) atomicOr(b.nonfinite[f], 1u);
  }
  if (p.coverage) p.coverage[o] = p.coverage[o] + cov;
}

// ------------------------------------------------------------------------
// Tiled Lanczos3 for affine maps.  A block of kWaves waves owns a kOutW x
// kOutH output tile and walks the frames of the batch.  For a frame whose
// samples are all finite (and no clamp is requested) the tile's source
// bounding box, with the +-3 tap halo, is copied into LDS by pipelined
// coalesced loads; every output pixel then takes its 36 taps from LDS (through
// a per-lane register window, see strip_window_rows) and its 12 tap weights
// from an LDS table of per-interval quadratics of the *normalised* Lanczos3
// weights, so there is no per-pixel transcendental, division or weight-sum
// normalisation.  Source coordinates, the footprint
// test and floor() stay FP64 and identical to the CPU path.  Frames holding
// non-finite samples (or clamped resident warps) take lanczos_direct per pixel
// inside the same launch; boxes too large for LDS read taps from the frame.
//
// Thread layout: the block's waves split the tile into kColGroups column
// groups of 64 and row bands of kLaneRows; lane tx of wave ty owns column
// ox0 + 64 (ty % kColGroups) + tx of the kLaneRows rows starting at
// oy0 + kLaneRows (ty / kColGroups).  Its per-launch coverage counts
// (<= kWarpBatch = 8 each) are packed 4 bits per row in one u64.
//
// Shape parameters (swept on MI355X: 4 waves, 128 x 32 outputs and a 29 KB
// box buffer beat 2/8 waves, 16/24-row tiles and double-buffered asynchronous
// global->LDS staging; the macros exist for such sweeps):
//   GH_WARP_WAVES     waves per block
//   GH_WARP_OUT_H     output tile rows (a multiple of the wave count)
//   GH_WARP_MAX_TILE  floats of the LDS box buffer
#ifndef GH_WARP_WAVES
#define GH_WARP_WAVES 4
#endif
#ifndef GH_WARP_OUT_H
#define GH_WARP_OUT_H 32
#endif
#ifndef GH_WARP_MAX_TILE
#define GH_WARP_MAX_TILE 7424
#endif
constexpr int kWaves = GH_WARP_WAVES;
constexpr int kTiledThreads = 64 * kWaves;
constexpr int kOutW = 128;
constexpr int kOutH = GH_WARP_OUT_H;
constexpr int kColGroups = kOutW / 64;
constexpr int kLaneRows = kOutH * kColGroups / kWaves;
constexpr int kMaxTile = GH_WARP_MAX_TILE;
static_assert(kWaves % kColGroups == 0 && kOutH % (kWaves / kColGroups) == 0, "waves tile the output evenly");
constexpr int kLutVec = kLanczosLutStride / 4;
constexpr int kLutRows = kLanczosLutIntervals + 1;
static_assert(kLanczosLutStride % 4 == 0, "table intervals are whole float4s");
static_assert(kLutVec == 5, "an interval's 18 coefficients are four float4s and a float2");
// LDS table: four planes of kLutRows float4s, then one plane of kLutRows
// float2s (floats 16 and 17; the host table's floats 18 and 19 are padding).
constexpr int kLutLdsVec4 = 4 * kLutRows + (kLutRows + 1) / 2;
static_assert(kLaneRows * 4 <= 64 && kWarpBatch <= 15, "coverage counts are packed 4 bits per row");

// The 18 quadratic coefficients of table interval i.  The LDS table is
// stored as planes (plane v < 4 holds floats 4v .. 4v+3 of every interval as
// float4s, the last plane floats 16 and 17 as float2s), so lanes reading
// distinct intervals touch consecutive slots; row kLanczosLutIntervals serves
// f == 1 exactly.
__device__ __forceinline__ void lut_coefficients(const float4* lut, int i, float (&c)[18]) {
  // One address register: interval i's float4s lie at byte 16 i of their
  // planes and its float2 at byte 8 i = (16 i) / 2 of the last plane (the
  // plane bases are immediate offsets).
  unsigned b = static_cast<unsigned>(i) << 4;
  asm("" : "+v"(b));
  const char* base = reinterpret_cast<const char*>(lut);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const float4 q = *reinterpret_cast<const float4*>(base + b + v * kLutRows * 16);
    c[4 * v + 0] = q.x;
    c[4 * v + 1] = q.y;
    c[4 * v + 2] = q.z;
    c[4 * v + 3] = q.w;
  }
  const float2 q = *reinterpret_cast<const float2*>(base + 4 * kLutRows * 16 + (b >> 1));
  c[16] = q.x;
  c[17] = q.y;
}

// Six normalised tap weights for fractional offset f in [0, 1].
__device__ __forceinline__ void lut_weights(const float4* lut, float f, float (&w)[6]) {
  const float t = f * static_cast<float>(kLanczosLutIntervals);
  const int i = static_cast<int>(t);
  const float u = __builtin_amdgcn_fractf(t);
  float c[18];
  lut_coefficients(lut, i, c);
#pragma unroll
  for (int k = 0; k < 6; ++k) w[k] = fmaf(u, fmaf(u, c[3 * k + 2], c[3 * k + 1]), c[3 * k]);
}

// T: an LDS pointer (staged tile) or a global-address-space frame pointer.
template <typename T>
__device__ __forceinline__ float lanczos_taps(T t, int stride, const float (&wx)[6], const float (&wy)[6]) {
  float acc = 0.0f;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const auto r = t + j * stride;
    float h = r[0] * wx[0];
#pragma unroll
    for (int i = 1; i < 6; ++i) h = fmaf(r[i], wx[i], h);
    acc = j == 0 ? h * wy[0] : fmaf(h, wy[j], acc);
  }
  return acc;
}

struct TileFrame {
  const float* src;
  float* dst;
  double m0, m1, m2, m3, m4, m5;
  int bx0, by0, bw, bh;  // staged box origin, row stride and rows
  bool fast;             // table weights (no clamp; finite source unless CPU semantics)
  bool use_tile;         // fast and the box fits one LDS buffer
  bool recheck;          // fast over a source holding non-finite samples
};

// Frame f's matrix, path choice and source bounding box for the output tile
// [ox0, ox1] x [oy0, oy1].  An affine map takes its extremes at the corners;
// interior points computed in FP64 stay within ~1e-11 px of the corners' hull,
// so one pixel beyond the +-3 tap reach covers every tap.
// One frame of a tile's work: source, destination, FP64 inverse (the two
// affine rows), the source's "holds a non-finite sample" flag and the
// destination's flag to raise.
struct FrameRef {
  const float* src;
  float* dst;
  double inv[6];
  unsigned src_flag;
  unsigned* dst_flag;
};

__device__ __forceinline__ TileFrame tile_frame(const FrameRef& ref, bool skip_nonfinite, bool clamp_on, int ox0,
                                                int oy0, int ox1, int oy1, int W, int H) {
  TileFrame fr;
  fr.src = ref.src;
  fr.dst = ref.dst;
  const double* m = ref.inv;
  fr.m0 = m[0]; fr.m1 = m[1]; fr.m2 = m[2];
  fr.m3 = m[3]; fr.m4 = m[4]; fr.m5 = m[5];
  // CPU semantics over a non-finite source still take the table path; the
  // pixels it cannot reproduce are redone by recheck_column.
  fr.fast = !clamp_on && (ref.src_flag == 0u || !skip_nonfinite);
  fr.recheck = ref.src_flag != 0u;
  double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double sx, sy;
    source_coords(m, 1, (c & 1) ? ox1 : ox0, (c & 2) ? oy1 : oy0, sx, sy);
    mnx = fmin(mnx, sx);
    mxx = fmax(mxx, sx);
    mny = fmin(mny, sy);
    mxy = fmax(mxy, sy);
  }
  const bool box_finite = fabs(mnx) < 1.0e9 && fabs(mxx) < 1.0e9 && fabs(mny) < 1.0e9 && fabs(mxy) < 1.0e9;
  fr.bx0 = fr.by0 = fr.bw = fr.bh = 0;
  if (box_finite) {
    fr.bx0 = max(static_cast<int>(floor(mnx)) - 3, 0);
    fr.by0 = max(static_cast<int>(floor(mny)) - 3, 0);
    fr.bw = min(static_cast<int>(floor(mxx)) + 4, W - 1) - fr.bx0 + 1;
    fr.bh = min(static_cast<int>(floor(mxy)) + 4, H - 1) - fr.by0 + 1;
  }
  // (boxes narrower than the 6-tap window only arise when no pixel of the
  // tile is inside the footprint; they take the global-tap form)
  fr.use_tile = fr.fast && box_finite && fr.bw >= 6 && fr.bh >= 6 && fr.bw * fr.bh <= kMaxTile;
  return fr;
}

// Copies the frame's box [by0, by0 + bh) x [bx0, bx0 + bw) into tile[] (row
// stride bw) with plain loads: element e = row * bw + col goes to tile[e];
// thread t takes e = t, t + kTiledThreads, ... with 8 loads in flight per
// round.  Loads past the end read the box origin instead (in bounds) and are
// not stored.  The caller synchronises before and after.
__device__ __forceinline__ void stage_box(const TileFrame& fr, int W, float* tile) {
  constexpr int kUnroll = 8;
  const int total = fr.bw * fr.bh;
  const int dq = kTiledThreads / fr.bw, dr = kTiledThreads - dq * fr.bw;
  const global_cf32* origin = (const global_cf32*)(fr.src + static_cast<long long>(fr.by0) * W + fr.bx0);
  int e = threadIdx.x;
  int row = e / fr.bw, col = e - row * fr.bw;
  for (; e < total; e += kTiledThreads * kUnroll) {
    float v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const bool in = e + kTiledThreads * u < total;
      v[u] = *(in ? origin + static_cast<long long>(row) * W + col : origin);
      col += dr;
      row += dq;
      if (col >= fr.bw) {
        col -= fr.bw;
        row += 1;
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (e + kTiledThreads * u < total) tile[e + kTiledThreads * u] = v[u];
    }
  }
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [0, N).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Lanczos3 sum over a 6 x 6 tap block held in registers: rows R0 .. R0+5 and
// columns C0 .. C0+5 of the window.
// (Compile-time loops throughout: the window must stay in registers, and
// SROA runs before loop unrolling.)
template <int R0, int C0, int NR, int NC>
__device__ __forceinline__ float window_taps(const float (&win)[NR][NC], const float (&wx)[6], const float (&wy)[6]) {
  float acc = 0.0f;
  static_for<6>([&](auto J) __attribute__((always_inline)) {
    constexpr int j = decltype(J)::value;
    float h = win[R0 + j][C0] * wx[0];
    static_for<5>([&](auto I) __attribute__((always_inline)) {
      constexpr int i = decltype(I)::value + 1;
      h = fmaf(win[R0 + j][C0 + i], wx[i], h);
    });
    acc = j == 0 ? h * wy[0] : fmaf(h, wy[j], acc);
  });
  return acc;
}

// One vertical strip of kStripRows output pixels of column x, rows y_first ..,
// taps from the LDS box through a register window.  Consecutive output rows of
// a near-identity map read source windows one row apart (and at most one
// column apart), so the lane keeps a kWinRows x kWinCols window of the box in
// registers and loads one new row per pixel instead of 36 taps.  At each row a
// wave-wide vote checks that every lane inside the footprint finds its taps at
// window row r and the same column offset (0 or 1); otherwise that row reads
// its taps from LDS directly.  Pixels outside the footprint write `fill`
// (their LDS-fallback taps are read at an in-box position).  Row r's coverage
// count goes to bits 4 r
// of cov32 (the caller shifts the strip into its lane's 64-bit counts).
// Returns whether any written value is non-finite.
constexpr int kStripRows = 8;
constexpr int kWinRows = kStripRows + 5;
constexpr int kWinCols = 7;
static_assert(kLaneRows % kStripRows == 0, "a lane's rows split into whole strips");

__device__ __forceinline__ bool strip_window_rows(const TileFrame& fr, const float* tile, const float4* lut, int x,
                                                  int y_first, int W, int H, float fill, unsigned& cov32) {
  const double xd = static_cast<double>(x);
  const double px = fr.m0 * xd, py = fr.m3 * xd;
  const double xmax = static_cast<double>(W - 3), ymax = static_cast<double>(H - 3);
  const int base = -(fr.by0 + 2) * fr.bw - (fr.bx0 + 2);
  // Window origin (box-relative) from the strip's first pixel; the spare
  // column goes on the side the columns drift to (sx moves by m1 per row).
  const double sx0 = fmin(fmax((px + fr.m1 * static_cast<double>(y_first)) + fr.m2, -1.0e6), 1.0e6);
  const double sy0 = fmin(fmax((py + fr.m4 * static_cast<double>(y_first)) + fr.m5, -1.0e6), 1.0e6);
  const int lead = fr.m1 < 0.0 ? 3 : 2;
  const int xb = min(max(static_cast<int>(floor(sx0)) - lead - fr.bx0, 0), fr.bw - kWinCols);
  const int yb = min(max(static_cast<int>(floor(sy0)) - 2 - fr.by0, 0), fr.bh - kWinRows);
  int woff = yb * fr.bw + xb;
  // Row weights: y's fraction moves by m4 - 1 per row (about 1.5e-4 for a
  // near-identity map), so a lane keeps the y coefficients of its strip's
  // first interval and evaluates that quadratic at u = t - i_y for the later
  // rows while u stays within [-0.25, 1.25] (error <= 2.2e-7, against 1e-7
  // inside the interval) and the source row advances by one per output row.
  // Otherwise the lane re-anchors on the row's own interval.  The x fraction
  // moves by m1 per row (a rotation), so x weights are looked up per row.
  int iy_lut = min(static_cast<int>(static_cast<float>(__builtin_amdgcn_fract(sy0)) *
                                    static_cast<float>(kLanczosLutIntervals)),
                   kLanczosLutIntervals - 1);
  iy_lut = max(iy_lut, 0);
  float cy[18];
  lut_coefficients(lut, iy_lut, cy);
  float win[kWinRows][kWinCols];
  static_for<5 * kWinCols>([&](auto E) __attribute__((always_inline)) {
    constexpr int j = decltype(E)::value / kWinCols, c = decltype(E)::value % kWinCols;
    win[j][c] = tile[woff + j * fr.bw + c];
  });
  bool bad = false;
  // 32-bit element offset of (x, y_first); row r adds r * W.
  const unsigned o_first = __umul24(static_cast<unsigned>(y_first), static_cast<unsigned>(W)) + static_cast<unsigned>(x);
  static_for<kStripRows>([&](auto R) __attribute__((always_inline)) {
    constexpr int r = decltype(R)::value;
    // The row's offset passes through an opaque move so its loads stay in this
    // step: hoisted, all kWinRows rows would be live at once (occupancy 3).
    asm volatile("" : "+v"(woff));
    static_for<kWinCols>([&](auto C) __attribute__((always_inline)) {
      constexpr int c = decltype(C)::value;
      win[r + 5][c] = tile[woff + (r + 5) * fr.bw + c];
    });
    const int y = y_first + r;
    if (y >= H) return;
    const double sx = (px + fr.m1 * static_cast<double>(y)) + fr.m2;
    const double sy = (py + fr.m4 * static_cast<double>(y)) + fr.m5;
    const bool ok = x < W && sx >= 2.0 && sx < xmax && sy >= 2.0 && sy < ymax;
    // In the footprint s >= 2, so truncation is floor and fract(s) is exactly
    // s - floor(s).  Outside it the weights, the window vote and the value are
    // ignored; only the LDS fallback below must stay inside the box.
    float wx[6], wy[6];
    lut_weights(lut, static_cast<float>(__builtin_amdgcn_fract(sx)), wx);
    int ix, iy;  // the hardware conversion saturates out-of-range coordinates
    asm("v_cvt_i32_f64 %0, %1" : "=v"(ix) : "v"(sx));
    asm("v_cvt_i32_f64 %0, %1" : "=v"(iy) : "v"(sy));
    const int o = ix - 2 - fr.bx0 - xb;
    const bool row_ok = iy - 2 - fr.by0 - yb == r;
    const float ty = static_cast<float>(__builtin_amdgcn_fract(sy)) * static_cast<float>(kLanczosLutIntervals);
    float uy = ty - static_cast<float>(iy_lut);
    if (!(row_ok && uy >= -0.25f && uy <= 1.25f)) {
      iy_lut = min(static_cast<int>(ty), kLanczosLutIntervals - 1);
      uy = ty - static_cast<float>(iy_lut);
      lut_coefficients(lut, iy_lut, cy);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) wy[k] = fmaf(uy, fmaf(uy, cy[3 * k + 2], cy[3 * k + 1]), cy[3 * k]);
    float value;
    if (__all(!ok || (row_ok && o == 0))) {
      value = window_taps<r, 0>(win, wx, wy);
    } else if (__all(!ok || (row_ok && o == 1))) {
      value = window_taps<r, 1>(win, wx, wy);
    } else {
      const int ixc = ok ? ix : fr.bx0 + 2, iyc = ok ? iy : fr.by0 + 2;
      value = lanczos_taps(tile + (iyc * fr.bw + ixc + base), fr.bw, wx, wy);
    }
    value = ok ? value : fill;
    if (x < W) {
      cov32 += ok ? (1u << (4 * r)) : 0u;
      st_global_u32(fr.dst, o_first + static_cast<unsigned>(r * W), value);
      bad |= !isfinite(value);
    }
  });
  return bad;
}

// Pixels of one frame with table weights and taps read straight from the
// frame (source box too large for LDS).
__device__ __forceinline__ bool column_global_rows(const TileFrame& fr, const float4* lut, int x, int y_first, int W,
                                                   int H, float fill, unsigned long long& cov) {
  bool bad = false;
  const double xmax = static_cast<double>(W - 3), ymax = static_cast<double>(H - 3);
  if (x >= W) return false;
#pragma unroll 1
  for (int k = 0; k < kLaneRows; ++k) {
    const int y = y_first + k;
    if (y >= H) break;
    const double sx = (fr.m0 * static_cast<double>(x) + fr.m1 * static_cast<double>(y)) + fr.m2;
    const double sy = (fr.m3 * static_cast<double>(x) + fr.m4 * static_cast<double>(y)) + fr.m5;
    float value = fill;
    if (sx >= 2.0 && sx < xmax && sy >= 2.0 && sy < ymax) {
      const double fx0 = floor(sx), fy0 = floor(sy);
      const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0);
      float wx[6], wy[6];
      lut_weights(lut, static_cast<float>(sx - fx0), wx);
      lut_weights(lut, static_cast<float>(sy - fy0), wy);
      value = lanczos_taps((const global_cf32*)(fr.src + static_cast<long long>(y0 - 2) * W + (x0 - 2)), W, wx, wy);
      cov += 1ull << (4 * k);
    }
    st_global(fr.dst, static_cast<long long>(y) * W + x, value);
    bad |= !isfinite(value);
  }
  return bad;
}

// CPU-semantics frames whose source holds non-finite samples run the table
// paths too.  A table value is finite exactly when its 36 taps are (a NaN tap,
// or an infinite one, makes the weighted sum non-finite whatever its weight),
// so only in-footprint pixels whose stored value is non-finite are recomputed,
// per tap with zero-weight taps skipped as warp.py:121-152 does.  The lane
// re-reads the column it has just written (its own stores, program order).
// Returns whether any final value of the column is non-finite.
__device__ __forceinline__ bool recheck_column(const TileFrame& fr, int x, int y_first, int W, int H) {
  if (x >= W) return false;
  WarpParams q{};
  q.skip_nonfinite = 0;
  q.clamp = -1.0f;
  const double xmax = static_cast<double>(W - 3), ymax = static_cast<double>(H - 3);
  bool bad = false;
#pragma unroll 1
  for (int k = 0; k < kLaneRows; ++k) {
    const int y = y_first + k;
    if (y >= H) break;
    const long long o = static_cast<long long>(y) * W + x;
    float v = ld_global(fr.dst, o);
    if (!isfinite(v)) {
      // Same expression and footprint test as the table paths.
      const double sx = (fr.m0 * static_cast<double>(x) + fr.m1 * static_cast<double>(y)) + fr.m2;
      const double sy = (fr.m3 * static_cast<double>(x) + fr.m4 * static_cast<double>(y)) + fr.m5;
      if (sx >= 2.0 && sx < xmax && sy >= 2.0 && sy < ymax) {
        bool covered;
        v = lanczos_direct(fr.src, W, sx, sy, q, 1.0e-8f, covered);
        st_global(fr.dst, o, v);
      }
    }
    bad |= !isfinite(v);
  }
  return bad;
}

// Pixels of one frame through lanczos_direct (non-finite sources, clamp).
__device__ __forceinline__ bool column_direct_rows(const TileFrame& fr, const WarpParams& p, int x, int y_first,
                                                   unsigned long long& cov) {
  bool bad = false;
  const float thr = p.skip_nonfinite ? 1.0e-6f : 1.0e-8f;
  const double m[6] = {fr.m0, fr.m1, fr.m2, fr.m3, fr.m4, fr.m5};
  if (x >= p.width) return false;
#pragma unroll 1
  for (int k = 0; k < kLaneRows; ++k) {
    const int y = y_first + k;
    if (y >= p.height) break;
    double sx, sy;
    source_coords(m, 1, x, y, sx, sy);
    float value = p.fill;
    if (lanczos_footprint(sx, sy, p.width, p.height)) {
      bool covered;
      value = lanczos_direct(fr.src, p.width, sx, sy, p, thr, covered);
      if (covered) cov += 1ull << (4 * k);
    }
    st_global(fr.dst, static_cast<long long>(y) * p.width + x, value);
    bad |= !isfinite(value);
  }
  return bad;
}

// The tiled Lanczos3 work of one output tile over `count` frames (get(f)
// yields frame f's FrameRef).  Uses the block's LDS table and box buffer; every
// thread of the block must call it.  ATOMIC_COVERAGE: other blocks may add to
// the same tile's coverage concurrently (batched launches).
template <bool ATOMIC_COVERAGE, typename GetFrame>
__device__ __forceinline__ void lanczos_tile(const GetFrame& get, int count, const WarpParams& p, int ox0, int oy0,
                                             const float4* lut, float* tile) {
  const int tx = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  const int x = ox0 + 64 * (ty % kColGroups) + tx;
  const int y_first = oy0 + kLaneRows * (ty / kColGroups);
  const int W = p.width, H = p.height;
  const bool clamp_on = p.skip_nonfinite && p.clamp >= 0.0f;
  const int ox1 = min(ox0 + kOutW, W) - 1, oy1 = min(oy0 + kOutH, H) - 1;
  unsigned long long cov = 0;

  for (int f = 0; f < count; ++f) {
    const FrameRef ref = get(f);
    const TileFrame cur = tile_frame(ref, p.skip_nonfinite != 0, clamp_on, ox0, oy0, ox1, oy1, W, H);
    __syncthreads();  // previous frame's taps consumed (first pass: table stored)
    if (cur.use_tile) stage_box(cur, W, tile);
    __syncthreads();
    bool bad = false;
    if (cur.use_tile) {
#pragma unroll 1
      for (int s = 0; s < kLaneRows / kStripRows; ++s) {
        unsigned cov32 = 0u;
        bad |= strip_window_rows(cur, tile, lut, x, y_first + s * kStripRows, W, H, p.fill, cov32);
        cov += static_cast<unsigned long long>(cov32) << (4 * kStripRows * s);
      }
    } else if (cur.fast) {
      bad = column_global_rows(cur, lut, x, y_first, W, H, p.fill, cov);
    } else {
      bad = column_direct_rows(cur, p, x, y_first, cov);
    }
    if (cur.fast && cur.recheck && bad) bad = recheck_column(cur, x, y_first, W, H);
    if (__any(bad) && tx == 0) atomicOr(ref.dst_flag, 1u);
  }
  if (p.coverage && x < W) {
#pragma unroll 1
    for (int k = 0; k < kLaneRows; ++k) {
      const int y = y_first + k;
      if (y >= H) break;
      const long long o = static_cast<long long>(y) * W + x;
      const float add = static_cast<float>((cov >> (4 * k)) & 15ull);
      if constexpr (ATOMIC_COVERAGE) {
        if (add != 0.0f) atomicAdd(p.coverage + o, add);
      } else {
        st_global(p.coverage, o, ld_global(p.coverage, o) + add);
      }
    }
  }
}

// Bilinear over one output tile (the tiled kernels' thread layout, no LDS),
// for affine maps of frames under 2^30 pixels.  The arithmetic is the CPU
// path's (src/glass/engine/warp.py:99-118: FP64 coordinates, f32 lerps), so
// the output is the same bit for bit.  A lane computes four rows at a time
// and issues their 16 taps before any lerp, so 16 gathers are in flight per
// lane rather than 4.  Taps are buffer loads at 32-bit byte offsets; a pixel
// outside the footprint reads pixel (0, 0) and discards it.
// ATOMIC_COVERAGE: other blocks may add to the same tile's coverage.
template <bool ATOMIC_COVERAGE, typename GetFrame>
__device__ __forceinline__ void bilinear_tile(const GetFrame& get, int count, const WarpParams& p, int ox0, int oy0) {
  const int tx = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  const int x = ox0 + 64 * (ty % kColGroups) + tx;
  const int y_first = oy0 + kLaneRows * (ty / kColGroups);
  const int W = p.width, H = p.height;
  if (x >= W) return;
  unsigned long long cov = 0;
  const double xs = static_cast<double>(x);
  const double xmax = static_cast<double>(W - 1), ymax = static_cast<double>(H - 1);
  // frames hold fewer than 2^30 pixels (checked by the host): byte offsets fit 32 bits
  const unsigned bytes = static_cast<unsigned>(W) * static_cast<unsigned>(H) * 4u;
  constexpr int kGroup = 4;
  static_assert(kLaneRows % kGroup == 0, "a lane's rows split into groups");
  for (int f = 0; f < count; ++f) {
    const FrameRef ref = get(f);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(ref.src), 0,
                                                                        static_cast<int>(bytes), 0x00020000);
    const double px = ref.inv[0] * xs, py = ref.inv[3] * xs;
    bool bad = false;
#pragma unroll 1
    for (int k0 = 0; k0 < kLaneRows; k0 += kGroup) {
      float t[kGroup][4], fx[kGroup], fy[kGroup];
      bool ok[kGroup];
#pragma unroll
      for (int g = 0; g < kGroup; ++g) {
        const int y = y_first + k0 + g;
        const double ys = static_cast<double>(y);
        const double sx = (px + ref.inv[1] * ys) + ref.inv[2];
        const double sy = (py + ref.inv[4] * ys) + ref.inv[5];
        ok[g] = y < H && isfinite(sx) && isfinite(sy) && sx >= 0.0 && sx <= xmax && sy >= 0.0 && sy <= ymax;
        const double fx0 = floor(sx), fy0 = floor(sy);
        int ix, iy;  // the hardware conversion saturates (ignored outside the footprint)
        asm("v_cvt_i32_f64 %0, %1" : "=v"(ix) : "v"(fx0));
        asm("v_cvt_i32_f64 %0, %1" : "=v"(iy) : "v"(fy0));
        const int x0 = ok[g] ? ix : 0, y0 = ok[g] ? iy : 0;
        const unsigned dx = x0 + 1 < W ? 4u : 0u;
        const unsigned dy = y0 + 1 < H ? static_cast<unsigned>(W) * 4u : 0u;
        const unsigned o00 = (static_cast<unsigned>(y0) * static_cast<unsigned>(W) + static_cast<unsigned>(x0)) * 4u;
        t[g][0] = ld_rsrc_b32(rs, o00);
        t[g][1] = ld_rsrc_b32(rs, o00 + dx);
        t[g][2] = ld_rsrc_b32(rs, o00 + dy);
        t[g][3] = ld_rsrc_b32(rs, o00 + dy + dx);
        fx[g] = static_cast<float>(sx - fx0);
        fy[g] = static_cast<float>(sy - fy0);
      }
#pragma unroll
      for (int g = 0; g < kGroup; ++g) {
        const int y = y_first + k0 + g;
        float value = p.fill;
        if (ok[g]) {
          const float top = t[g][0] * (1.0f - fx[g]) + t[g][1] * fx[g];
          const float bottom = t[g][2] * (1.0f - fx[g]) + t[g][3] * fx[g];
          value = top * (1.0f - fy[g]) + bottom * fy[g];
          cov += 1ull << (4 * (k0 + g));
        }
        if (y < H) {
          st_global(ref.dst, static_cast<long long>(y) * W + x, value);
          bad |= !isfinite(value);
        }
      }
    }
    if (__any(bad) && tx == 0) atomicOr(ref.dst_flag, 1u);
  }
  if (p.coverage) {
#pragma unroll 1
    for (int k = 0; k < kLaneRows; ++k) {
      const int y = y_first + k;
      if (y >= H) break;
      const float add = static_cast<float>((cov >> (4 * k)) & 15ull);
      const long long o = static_cast<long long>(y) * W + x;
      if constexpr (ATOMIC_COVERAGE) {
        if (add != 0.0f) atomicAdd(p.coverage + o, add);
      } else {
        st_global(p.coverage, o, ld_global(p.coverage, o) + add);
      }
    }
  }
}

// Affine bilinear warp of a launch's frames over 128 x 32 output tiles.
__global__ __launch_bounds__(kTiledThreads) void warp_bilinear_tiled_kernel(WarpBatch b, int count, WarpParams p,
                                                                            int tiles_x) {
  const unsigned nb = gridDim.x, bid = blockIdx.x;
  const unsigned xcd = bid & 7u, q = nb >> 3, r = nb & 7u;
  const unsigned logical = xcd * q + min(xcd, r) + (bid >> 3);
  const int ox0 = static_cast<int>(logical % tiles_x) * kOutW;
  const int oy0 = static_cast<int>(logical / tiles_x) * kOutH;
  auto get = [&](int f) {
    const double* m = b.inv[f];
    return FrameRef{b.src[f], b.dst[f], {m[0], m[1], m[2], m[3], m[4], m[5]}, b.src_flags[f], b.nonfinite[f]};
  };
  bilinear_tile<false>(get, count, p, ox0, oy0);
}

__device__ __forceinline__ void load_lut(const float4* lut_g, float4* lut) {
  // host layout [interval][vec] -> LDS planes [vec][interval]
  for (int i = threadIdx.x; i < kLutRows * kLutVec; i += kTiledThreads) {
    const int row = i / kLutVec, v = i % kLutVec;
    const float4 q = lut_g[i];
    if (v < 4) lut[v * kLutRows + row] = q;
    else reinterpret_cast<float2*>(lut + 4 * kLutRows)[row] = make_float2(q.x, q.y);
  }
}

__global__ __launch_bounds__(kTiledThreads) __attribute__((amdgpu_waves_per_eu(3))) void warp_lanczos3_tiled_kernel(WarpBatch b, int count, WarpParams p,
                                                                             int tiles_x) {
  __shared__ float4 lut[kLutLdsVec4];
  __shared__ float tile[kMaxTile];
  load_lut(p.lanczos_lut, lut);
  // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
  // contiguous run of tiles so vertically adjacent tiles (which share source
  // halo rows) meet in the same L2.
  const unsigned nb = gridDim.x, bid = blockIdx.x;
  const unsigned xcd = bid & 7u, q = nb >> 3, r = nb & 7u;
  const unsigned logical = xcd * q + min(xcd, r) + (bid >> 3);
  const int ox0 = static_cast<int>(logical % tiles_x) * kOutW;
  const int oy0 = static_cast<int>(logical / tiles_x) * kOutH;
  auto get = [&](int f) {
    const double* m = b.inv[f];
    return FrameRef{b.src[f], b.dst[f], {m[0], m[1], m[2], m[3], m[4], m[5]}, b.src_flags[f], b.nonfinite[f]};
  };
  lanczos_tile<false>(get, count, p, ox0, oy0, lut, tile);
}

// ------------------------------------------------------------------------
// Streamed Lanczos3 for near-identity affine maps (registration drift: a
// small rotation and a scale close to one, so |m1| is small and m4 is close
// to 1).  Lane tx of wave ty owns output column ox0 + 64 (ty % kSColGroups) +
// tx over kSLaneRows rows, walked in strips of kSStrip rows.  Within a strip
// every output row moves the source row by exactly one (checked per strip),
// so the lane keeps a register window of the source columns it needs and
// reads one new window row per output row, from a window origin fixed for the
// strip.  Fractions advance by the uniform per-row drifts (m1 for x, m4 - 1
// for y, in table units) in f32 from the strip's first row, computed in FP64
// exactly as the CPU path does; the footprint test is exact FP64 at the
// strip's two end rows (the map is linear along the column, so the rows
// between lie between them).  Where some lane's x fraction wraps within the
// strip the window holds seven columns and each lane selects its six per row
// from the column sums.  So a pixel costs its filter (36 + 6 FMAs, 24 for
// the weights) and a few instructions of coordinates, with no FP64 and no
// per-pixel address arithmetic.
//
// The window rows come from the tile's source box, staged into LDS by loads
// that write LDS directly (no vector register, no vector instruction per
// element); a frame whose box does not fit streams them from the frame with
// buffer loads whose row offset is a scalar.  A strip where a lane straddles
// the footprint edge (or lies within one pixel of it in x), where a source row
// is skipped or repeated, or whose y table interval would drift out of
// [-0.25, 1.25] of the strip's anchor runs row by row with FP64 coordinates.
// Frames holding non-finite samples, and clamped resident warps, take the
// per-row paths of the tiled kernel for every pixel.
//
// Compared with the rounded FP64 coordinates of the CPU path, the f32
// fraction differs by about 1e-7 px; a floor decision can only differ where
// the fraction is that close to an integer, where both tap windows give the
// same value to the table's accuracy (the weights are continuous there and
// the source is finite).
#ifndef GH_WARP_STREAM_STRIP
#define GH_WARP_STREAM_STRIP 8
#endif
#ifndef GH_WARP_STREAM_ROWS
#define GH_WARP_STREAM_ROWS 16
#endif
#ifndef GH_WARP_STREAM_AHEAD
#define GH_WARP_STREAM_AHEAD 1
#endif
#ifndef GH_WARP_STREAM_PIN
#define GH_WARP_STREAM_PIN 1
#endif
constexpr int kSStrip = GH_WARP_STREAM_STRIP;
constexpr int kSLaneRows = GH_WARP_STREAM_ROWS;
constexpr int kSAhead = GH_WARP_STREAM_AHEAD;
constexpr int kSWaves = 4;
constexpr int kSThreads = 64 * kSWaves;
constexpr int kSOutW = 128;
constexpr int kSColGroups = kSOutW / 64;
constexpr int kSOutH = kSLaneRows * kSWaves / kSColGroups;
static_assert(kSLaneRows % kSStrip == 0 && kSLaneRows % kLaneRows == 0, "lane rows split into strips and groups");
static_assert(kSStrip <= 16, "a strip's coverage counts fit one 64-bit word");
static_assert(kSThreads == kTiledThreads, "the table loader strides by the block size");

typedef __amdgpu_buffer_rsrc_t frame_rsrc_t;

__device__ __forceinline__ float ld_rsrc(frame_rsrc_t r, unsigned voff, unsigned soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// Per-frame uniform values of the streamed path.
struct StreamFrame {
  frame_rsrc_t src;
  frame_rsrc_t dst_rs;  // the output frame, for stores at 32-bit lane offsets
  float* dst;
  double m1, m2, m4, m5;
  float dtx, dty;  // per-row drift of the x / y table coordinate
  float thr;       // 7-column mode: lanes whose x table coordinate is >= thr use columns 1..6
  int neg;         // m1 < 0: the 7-column window starts one column further left
};

// Staged source box of one output tile and frame (LDS, row stride
// kBoxStride floats): rows by0 .. by0 + bh - 1, columns bx0 .. bx0 + bw - 1.
#ifndef GH_WARP_BOX_ROWS
#define GH_WARP_BOX_ROWS 40
#endif
#ifndef GH_WARP_BOX_AHEAD
#define GH_WARP_BOX_AHEAD 1
#endif
#ifndef GH_WARP_LUT_AHEAD
#define GH_WARP_LUT_AHEAD 1
#endif
#ifndef GH_WARP_BOX_STRIDE
#define GH_WARP_BOX_STRIDE 138
#endif
constexpr int kBoxStride = GH_WARP_BOX_STRIDE;
static_assert(kBoxStride > 128 && kBoxStride <= 192, "a box row is two 64-lane segments and a partial third");
constexpr int kBoxRows = GH_WARP_BOX_ROWS;
constexpr int kBoxFloats = kBoxStride * kBoxRows;
// Two boxes (the next frame's is staged while this frame's is read); with
// the table they fit three blocks into the CU's 160 KB of LDS.
constexpr int kStreamLdsFloats = 2 * kBoxFloats > kMaxTile ? 2 * kBoxFloats : kMaxTile;
// LDS is allocated in 1280-byte granules (measured: a 54480-byte block runs
// two per CU, a 53448-byte one three).
constexpr int kLdsGranule = 1280;
static_assert(((kLutLdsVec4 * 16 + kStreamLdsFloats * 4 + kLdsGranule - 1) / kLdsGranule) * kLdsGranule * 3 <=
                  160 * 1024,
              "three blocks of the streamed kernel per CU");
static_assert(kSOutW == kOutW && kSOutH == kOutH, "both affine Lanczos3 kernels use the same output tiles");
struct BoxGeom {
  int bx0, by0, bw, bh;
  bool ok;      // staged (finite source, no clamp, and the box fits)
  bool inside;  // every pixel of the tile maps inside the footprint (strip tests skipped)
};

// Window rows of a strip: loaded from the frame with buffer loads into
// registers (lane byte offset voff, row j at soffset j * row_bytes), or read
// from the staged LDS box (lane element offset base, row j at j * kBoxStride).
struct RowsFromFrame {
  frame_rsrc_t rs;
  unsigned voff, row_bytes;
  static constexpr int kAhead = kSAhead;
  static constexpr bool kLutAhead = false;
  __device__ __forceinline__ float at(int j, int c) const { return ld_rsrc(rs, voff + 4u * c, row_bytes * j); }
  __device__ __forceinline__ void pin() { asm volatile("" : "+v"(voff)); }
};
typedef __attribute__((address_space(3))) const float lds_f32;
struct RowsFromBox {
  lds_f32* row0;  // the lane's window row 0, column 0 in the box
  static constexpr int kAhead = GH_WARP_BOX_AHEAD;
  static constexpr bool kLutAhead = GH_WARP_LUT_AHEAD;
  __device__ __forceinline__ float at(int j, int c) const {
    lds_f32* p = row0 + j * kBoxStride;
    asm("" : "+v"(p));  // one LDS byte address per row; its columns are immediate offsets
    return p[c];
  }
  __device__ __forceinline__ void pin() { asm volatile("" : "+v"(row0)); }
};

// s_waitcnt vmcnt(0) as a builtin (a wait the compiler accounts for), at the
// ends of the streamed kernel's paths that read the frame with vector loads
// (rows from the frame, exact rows, non-finite frames).  Without it the
// compiler's wait bookkeeping carries those loads around the strip and frame
// loops and places waits into the staged-box strips, where they would also
// wait for the next frame's box loads in flight (a wave's vector memory
// operations complete in order).
__device__ __forceinline__ void drain_vector_memory() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// The 18 quadratic coefficients of interval i and the six weights at u.
__device__ __forceinline__ void lut_eval(const float4* lut, int i, float u, float (&w)[6]) {
  float c[18];
  lut_coefficients(lut, i, c);
#pragma unroll
  for (int k = 0; k < 6; ++k) w[k] = fmaf(u, fmaf(u, c[3 * k + 2], c[3 * k + 1]), c[3 * k]);
}

// One strip on the streamed path; NC = 6 (no lane's x fraction wraps) or 7.
// Row r of the strip is output row y0 + r; window row j is source row
// iy0 - 2 + j of the lane.  The column sums run down the window first, so a
// 7-column strip gives the same values as a 6-column one at the same
// fractions.
template <int S, int NC, typename Rows>
__device__ __forceinline__ bool stream_strip_rows(const StreamFrame& fr, const float4* lut, Rows src, float tx0,
                                                  float uy0, const float (&cy)[18], bool live, bool present, int x,
                                                  int y0, int W, float fill) {
  constexpr int kAhead = Rows::kAhead;
  float win[S + 5][NC];
  auto load_row = [&](auto J) __attribute__((always_inline)) {
    constexpr int j = decltype(J)::value;
    static_for<NC>([&](auto C) __attribute__((always_inline)) {
      constexpr int c = decltype(C)::value;
      win[j][c] = src.at(j, c);
    });
  };
  static_for<5 + kAhead>([&](auto J) __attribute__((always_inline)) { load_row(J); });
  float nf = 0.0f;
  const unsigned lane_byte = static_cast<unsigned>(x) * 4u;  // frames of < 2^31 bytes
  const unsigned row0_byte = static_cast<unsigned>(y0) * static_cast<unsigned>(W) * 4u;  // wave-uniform
  // x table interval, position in it and column choice of row r
  auto x_coord = [&](int r, int& ix, float& u, bool& sel) __attribute__((always_inline)) {
    const float txr = fmaf(static_cast<float>(r), fr.dtx, tx0);
    sel = false;
    if constexpr (NC == 6) {
      ix = static_cast<int>(txr);
      u = __builtin_amdgcn_fractf(txr);
    } else {
      // |txr| < 2^8 here; floor and conversion in one instruction
      int fl;
      asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(fl) : "v"(txr));
      ix = fl & (kLanczosLutIntervals - 1);
      u = __builtin_amdgcn_fractf(txr);
      sel = txr >= fr.thr;
    }
  };
  // From the box, the x coefficients of the next row are read one row ahead
  // (registers allow it there).
  constexpr bool kLutAhead = Rows::kLutAhead;
  int ix_n = 0;
  float u_n = 0.0f;
  bool sel_n = false;
  float cx_n[18];
  if constexpr (kLutAhead) {
    x_coord(0, ix_n, u_n, sel_n);
    lut_coefficients(lut, ix_n, cx_n);
  }
  static_for<S>([&](auto R) __attribute__((always_inline)) {
    constexpr int r = decltype(R)::value;
#if GH_WARP_STREAM_PIN
    // keeps the next row's loads in this step (hoisted, every window row of
    // the strip would be live at once)
    src.pin();
#endif
    if constexpr (r + 5 + kAhead < S + 5) load_row(std::integral_constant<int, r + 5 + kAhead>{});
    float wx[6], wy[6];
    float u;
    bool sel;
    if constexpr (kLutAhead) {
      float cx[18];
#pragma unroll
      for (int k = 0; k < 18; ++k) cx[k] = cx_n[k];
      u = u_n;
      sel = sel_n;
      if constexpr (r + 1 < S) {
        x_coord(r + 1, ix_n, u_n, sel_n);
        lut_coefficients(lut, ix_n, cx_n);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) wx[k] = fmaf(u, fmaf(u, cx[3 * k + 2], cx[3 * k + 1]), cx[3 * k]);
    } else {
      int ix;
      x_coord(r, ix, u, sel);
      lut_eval(lut, ix, u, wx);
    }
    const float uy = fmaf(static_cast<float>(r), fr.dty, uy0);
#pragma unroll
    for (int k = 0; k < 6; ++k) wy[k] = fmaf(uy, fmaf(uy, cy[3 * k + 2], cy[3 * k + 1]), cy[3 * k]);
    float v[NC];
    static_for<NC>([&](auto C) __attribute__((always_inline)) {
      constexpr int c = decltype(C)::value;
      float a = win[r][c] * wy[0];
      static_for<5>([&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value + 1;
        a = fmaf(win[r + j][c], wy[j], a);
      });
      v[c] = a;
    });
    float h[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      if constexpr (NC == 6) h[i] = v[i];
      else h[i] = sel ? v[i + 1] : v[i];
    }
    float value = h[0] * wx[0];
#pragma unroll
    for (int i = 1; i < 6; ++i) value = fmaf(h[i], wx[i], value);
    value = live ? value : fill;
    // 0 * value is NaN exactly when value is not finite: a vector-only test
    nf = fmaf(value, 0.0f, nf);
    if (present)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, value), fr.dst_rs, lane_byte,
                                            row0_byte + static_cast<unsigned>(r * W) * 4u, 0);
  });
  if constexpr (!Rows::kLutAhead) drain_vector_memory();  // rows read from the frame
  return present && !isfinite(nf);
}

// One strip row by row: FP64 coordinates as the CPU path, table weights, taps
// read from the frame (finite source) or, BOX, from the staged box when they
// lie inside it.  Row r's coverage count goes to bits 4 r of cov.
template <int S, bool BOX>
__device__ __forceinline__ bool stream_strip_exact(const FrameRef& ref, const float4* lut, double px, double py,
                                                   int x, int y0, int W, int H, float fill, unsigned long long& cov,
                                                   const float* box, const BoxGeom& g) {
  if (x >= W) return false;
  const double xmax = static_cast<double>(W - 3), ymax = static_cast<double>(H - 3);
  bool bad = false;
#pragma unroll 1
  for (int r = 0; r < S; ++r) {
    const int y = y0 + r;
    if (y >= H) break;
    const double sx = (px + ref.inv[1] * static_cast<double>(y)) + ref.inv[2];
    const double sy = (py + ref.inv[4] * static_cast<double>(y)) + ref.inv[5];
    float value = fill;
    if (sx >= 2.0 && sx < xmax && sy >= 2.0 && sy < ymax) {
      const double fx0 = floor(sx), fy0 = floor(sy);
      const int x0 = static_cast<int>(fx0), yy0 = static_cast<int>(fy0);
      float wx[6], wy[6];
      lut_weights(lut, static_cast<float>(sx - fx0), wx);
      lut_weights(lut, static_cast<float>(sy - fy0), wy);
      bool from_box = BOX;
      if constexpr (BOX) {
        from_box = x0 - 2 >= g.bx0 && x0 + 3 < g.bx0 + g.bw && yy0 - 2 >= g.by0 && yy0 + 3 < g.by0 + g.bh;
      }
      if (from_box) {
        value = lanczos_taps(box + ((yy0 - 2 - g.by0) * kBoxStride + (x0 - 2 - g.bx0)), kBoxStride, wx, wy);
      } else {
        value = lanczos_taps((const global_cf32*)(ref.src + static_cast<long long>(yy0 - 2) * W + (x0 - 2)), W, wx,
                             wy);
        drain_vector_memory();
      }
      cov += 1ull << (4 * r);
    }
    st_global(ref.dst, static_cast<long long>(y) * W + x, value);
    bad |= !isfinite(value);
  }
  return bad;
}

// One strip of a finite-source frame: decides (wave-wide) between the
// streamed rows and the exact rows.  BOX: the window rows come from the
// staged box g (each lane's window must lie inside it).
template <int S, bool BOX>
__device__ __forceinline__ bool stream_strip(const FrameRef& ref, const StreamFrame& fr, const float4* lut, double px,
                                             double py, int x, int y0, int W, int H, float fill,
                                             unsigned long long& cov, const float* box, const BoxGeom& g) {
  const bool present = x < W;
  const double yd0 = static_cast<double>(y0), yde = static_cast<double>(y0 + S - 1);
  const double sx0 = (px + fr.m1 * yd0) + fr.m2, sy0 = (py + fr.m4 * yd0) + fr.m5;
  const double sxe = (px + fr.m1 * yde) + fr.m2, sye = (py + fr.m4 * yde) + fr.m5;
  // Exact footprint decisions need a margin far above the FP64 rounding of
  // the coordinates (~1e-12 px); x keeps one more pixel so that the 7-column
  // window stays inside the row.  Tiles wholly inside skip the tests.
  bool inside = true, outside = false;
  if (!g.inside) {
    const double lox = fmin(sx0, sxe), hix = fmax(sx0, sxe), loy = fmin(sy0, sye), hiy = fmax(sy0, sye);
    constexpr double kEdge = 1.0e-7;
    const double xmax = static_cast<double>(W - 3), ymax = static_cast<double>(H - 3);
    inside = lox >= 3.0 && hix < xmax - 1.0 && loy >= 2.0 + kEdge && hiy < ymax - kEdge;
    outside = hix < 2.0 - kEdge || lox >= xmax + kEdge || hiy < 2.0 - kEdge || loy >= ymax + kEdge;
  }
  // In the footprint s >= 2, so truncation is floor and fract(s) = s - floor(s)
  // (the hardware conversion saturates coordinates far outside).
  int ix0, iy0, iye;
  asm("v_cvt_i32_f64 %0, %1" : "=v"(ix0) : "v"(sx0));
  asm("v_cvt_i32_f64 %0, %1" : "=v"(iy0) : "v"(sy0));
  asm("v_cvt_i32_f64 %0, %1" : "=v"(iye) : "v"(sye));
  const float tx0 = static_cast<float>(__builtin_amdgcn_fract(sx0)) * static_cast<float>(kLanczosLutIntervals);
  const float ty0 = static_cast<float>(__builtin_amdgcn_fract(sy0)) * static_cast<float>(kLanczosLutIntervals);
  // y coefficients anchored on the interval at the strip's middle row.
  const float tym = fmaf(0.5f * static_cast<float>(S - 1), fr.dty, ty0);
  const int iyl = min(max(static_cast<int>(floorf(tym)), 0), kLanczosLutIntervals - 1);
  const float uy0 = ty0 - static_cast<float>(iyl);
  const float uye = fmaf(static_cast<float>(S - 1), fr.dty, uy0);
  const float txe = fmaf(static_cast<float>(S - 1), fr.dtx, tx0);
  const bool rows_ok = iye - iy0 == S - 1;
  const bool anchor_ok = fminf(uy0, uye) >= -0.25f && fmaxf(uy0, uye) <= 1.25f;
  // At most one wrap of the x fraction per strip (the 7-column window).
  const bool wrap_ok = txe >= -static_cast<float>(kLanczosLutIntervals) &&
                       txe < 2.0f * static_cast<float>(kLanczosLutIntervals);
  // Box mode: the window's columns ix0 - 3 .. ix0 + 4 and rows iy0 - 2 ..
  // iy0 + S + 2 lie inside the staged box.
  bool in_box = true;
  if constexpr (BOX) {
    in_box = ix0 - 3 >= g.bx0 && ix0 + 4 < g.bx0 + g.bw && iy0 - 2 >= g.by0 && iy0 + S + 2 < g.by0 + g.bh;
  }
  const bool live = present && inside;
  const bool lane_ok = !present || outside || (inside && rows_ok && anchor_ok && wrap_ok && in_box);
  if (y0 + S <= H && __all(lane_ok)) {
    float cy[18];
    lut_coefficients(lut, iyl, cy);
    const bool cross = live && (txe < 0.0f || txe >= static_cast<float>(kLanczosLutIntervals));
    const bool seven = __any(cross);
    const int cb = ix0 - 2 - (seven ? fr.neg : 0);
    bool bad;
    if constexpr (BOX) {
      const RowsFromBox rows{(lds_f32*)box + (live ? (iy0 - 2 - g.by0) * kBoxStride + (cb - g.bx0) : 0)};
      bad = seven ? stream_strip_rows<S, 7>(fr, lut, rows, tx0, uy0, cy, live, present, x, y0, W, fill)
                  : stream_strip_rows<S, 6>(fr, lut, rows, tx0, uy0, cy, live, present, x, y0, W, fill);
    } else {
      const RowsFromFrame rows{fr.src, live ? static_cast<unsigned>(((iy0 - 2) * W + cb) * 4) : 0u,
                               static_cast<unsigned>(W) * 4u};
      bad = seven ? stream_strip_rows<S, 7>(fr, lut, rows, tx0, uy0, cy, live, present, x, y0, W, fill)
                  : stream_strip_rows<S, 6>(fr, lut, rows, tx0, uy0, cy, live, present, x, y0, W, fill);
    }
    if (live) cov += 0x1111111111111111ull >> (64 - 4 * S);
    return bad;
  }
  return stream_strip_exact<S, BOX>(ref, lut, px, py, x, y0, W, H, fill, cov, box, g);
}

// Source box of output tile [ox0, ox1] x [oy0, oy1] for frame `ref`: the
// affine map takes its extremes at the corners, and the box holds the
// 7-column window (columns floor(sx) - 3 .. + 4) and the six tap rows
// (floor(sy) - 2 .. + 3) of each.  An interior pixel whose FP64 coordinate
// rounds past a corner's can need a row or column more: the streamed strips
// check their windows against the box, and the exact rows read such taps
// from the frame.  Every value is wave-uniform.
__device__ __forceinline__ BoxGeom box_geometry(const FrameRef& ref, bool eligible, int ox0, int oy0, int ox1,
                                                int oy1, int W, int H) {
  BoxGeom g{0, 0, 0, 0, false, false};
  if (!eligible) return g;
  const double m0 = ref.inv[0], m1 = ref.inv[1], m2 = ref.inv[2], m3 = ref.inv[3], m4 = ref.inv[4], m5 = ref.inv[5];
  // Corners chosen by the coefficients' sign bits: scalar tests of the
  // uniform values instead of FP64 compares and 64-bit selects (a zero
  // coefficient gives the same product at either corner).
  auto nonneg = [](double m) { return __builtin_bit_cast(long long, m) >= 0; };
  const int cx0 = nonneg(m0) ? ox0 : ox1, cx1 = nonneg(m0) ? ox1 : ox0;
  const int cy0 = nonneg(m1) ? oy0 : oy1, cy1 = nonneg(m1) ? oy1 : oy0;
  const int dx0 = nonneg(m3) ? ox0 : ox1, dx1 = nonneg(m3) ? ox1 : ox0;
  const int dy0 = nonneg(m4) ? oy0 : oy1, dy1 = nonneg(m4) ? oy1 : oy0;
  const double sxl = (m0 * static_cast<double>(cx0) + m1 * static_cast<double>(cy0)) + m2;
  const double sxh = (m0 * static_cast<double>(cx1) + m1 * static_cast<double>(cy1)) + m2;
  const double syl = (m3 * static_cast<double>(dx0) + m4 * static_cast<double>(dy0)) + m5;
  const double syh = (m3 * static_cast<double>(dx1) + m4 * static_cast<double>(dy1)) + m5;
  int a, b, c, d;
  asm("v_cvt_i32_f64 %0, %1" : "=v"(a) : "v"(floor(sxl)));
  asm("v_cvt_i32_f64 %0, %1" : "=v"(b) : "v"(floor(sxh)));
  asm("v_cvt_i32_f64 %0, %1" : "=v"(c) : "v"(floor(syl)));
  asm("v_cvt_i32_f64 %0, %1" : "=v"(d) : "v"(floor(syh)));
  a = __builtin_amdgcn_readfirstlane(a);
  b = __builtin_amdgcn_readfirstlane(b);
  c = __builtin_amdgcn_readfirstlane(c);
  d = __builtin_amdgcn_readfirstlane(d);
  const int bx0 = max(a - 3, 0), bx1 = min(b + 4, W - 1), by0 = max(c - 2, 0), by1 = min(d + 3, H - 1);
  // Every pixel of the tile lies inside the footprint with the streamed
  // strips' margins: their tests at the corner extremes, which bound every
  // pixel's FP64 coordinate (the rounded products and sums are monotonic).
  constexpr double kEdge = 1.0e-7;
  g.inside = sxl >= 3.0 && sxh < static_cast<double>(W - 3) - 1.0 && syl >= 2.0 + kEdge &&
             syh < static_cast<double>(H - 3) - kEdge;
  g.bx0 = __builtin_amdgcn_readfirstlane(bx0);
  g.by0 = __builtin_amdgcn_readfirstlane(by0);
  g.bw = __builtin_amdgcn_readfirstlane(bx1 - bx0 + 1);
  g.bh = __builtin_amdgcn_readfirstlane(by1 - by0 + 1);
  g.ok = g.bw >= 1 && g.bh >= 1 && g.bw <= kBoxStride && g.bh <= kBoxRows;
  return g;
}

// Buffer resource over `bytes` bytes from `base` (raw buffer: stride 0,
// byte offsets), for the LDS loads below.
typedef int rsrc_words_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ rsrc_words_t rsrc_words(const float* base, int bytes) {
  const unsigned long long a = reinterpret_cast<unsigned long long>(base);
  rsrc_words_t r;
  r.x = __builtin_amdgcn_readfirstlane(static_cast<int>(a));
  r.y = __builtin_amdgcn_readfirstlane(static_cast<int>(a >> 32));
  r.z = bytes;
  r.w = 0x00020000;
  return r;
}

// One dword per lane from byte offset voff + soff of the buffer into LDS at
// lds_byte + 4 * lane.  Written as inline assembly so that the compiler does
// not order its own LDS reads behind these loads (it cannot tell the box
// being read from the box being filled); nothing is returned in a register,
// and the caller waits (s_waitcnt vmcnt(0) and a barrier) before reading the
// data.  Waits the compiler places for its own loads stay correct: vector
// memory operations of a wave complete in order.  M0 is a reserved register
// (the compiler warns that the clobber is not tracked); nothing the compiler
// generates in the kernels that stage boxes reads or writes M0 (their ISA
// holds M0 only inside these statements).
__device__ __forceinline__ void load_dword_to_lds(rsrc_words_t rs, unsigned lds_byte, unsigned voff, unsigned soff) {
  asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tbuffer_load_dword %1, %2, %3 offen lds"
               :
               : "s"(lds_byte), "v"(voff), "s"(rs), "s"(soff)
               : "memory", "m0");
}

// Issues the box of frame `ref` into LDS (no wait here).  Wave w loads rows
// w, w + 4, ...; each row is up to three segments of 64 lanes (the third 16
// lanes wide).  Lanes past the box width load its last column again, so every
// address lies inside the frame.
__device__ __forceinline__ void stage_box_lds(const FrameRef& ref, const BoxGeom& g, int W, int bytes, float* box) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  // The lane index recomputed here from an opaque zero: hoisted to the kernel
  // entry, the staging's lane offsets are kept across the whole frame loop
  // and spilled, and their reloads would sit before every staging.
  unsigned zero = 0u;
  asm volatile("" : "+s"(zero));
  const int lane = static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, zero)));
  const rsrc_words_t rs = rsrc_words(ref.src, bytes);
  const unsigned v0 = static_cast<unsigned>(min(lane, g.bw - 1) * 4);
  const unsigned v1 = static_cast<unsigned>(min(64 + lane, g.bw - 1) * 4);
  const unsigned v2 = static_cast<unsigned>(min(128 + lane, g.bw - 1) * 4);
  typedef __attribute__((address_space(3))) float lds_f32;
  const unsigned box_byte = static_cast<unsigned>(reinterpret_cast<unsigned long long>((lds_f32*)box));
#pragma unroll 1
  for (int row = wave; row < g.bh; row += kSWaves) {
    const unsigned soff = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(((g.by0 + row) * W + g.bx0) * 4));
    const unsigned d = __builtin_amdgcn_readfirstlane(box_byte + static_cast<unsigned>(row * kBoxStride * 4));
    load_dword_to_lds(rs, d, v0, soff);
    if (g.bw > 64) load_dword_to_lds(rs, d + 256u, v1, soff);
    if (g.bw > 128 && lane < kBoxStride - 128) load_dword_to_lds(rs, d + 512u, v2, soff);
  }
}

// Coordinates and drifts of a finite-source frame on the streamed path.
__device__ __forceinline__ StreamFrame stream_frame(const FrameRef& ref, int bytes) {
  StreamFrame fr;
  fr.src = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(ref.src), 0, bytes, 0x00020000);
  fr.dst_rs = __builtin_amdgcn_make_buffer_rsrc(ref.dst, 0, bytes, 0x00020000);
  fr.dst = ref.dst;
  fr.m1 = ref.inv[1];
  fr.m2 = ref.inv[2];
  fr.m4 = ref.inv[4];
  fr.m5 = ref.inv[5];
  fr.dtx = static_cast<float>(ref.inv[1]) * static_cast<float>(kLanczosLutIntervals);
  fr.dty = static_cast<float>(ref.inv[4] - 1.0) * static_cast<float>(kLanczosLutIntervals);
  fr.neg = ref.inv[1] < 0.0 ? 1 : 0;
  fr.thr = fr.neg ? 0.0f : static_cast<float>(kLanczosLutIntervals);
  return fr;
}

// The streamed Lanczos3 work of one output tile over `count` frames; every
// thread of the block calls it.  A finite-source frame whose box fits (no
// clamp) is staged into LDS by loads that write LDS directly, one frame
// ahead into the second of two boxes (`boxes`: 2 * kBoxFloats floats), and
// the waves read their window rows from the box.  138 x 40 boxes keep three
// blocks per CU (a 1-degree rotation of a 128 x 32 tile needs up to 136 x 40).
// Other finite-source frames stream their window rows from the frame; the
// rest take the per-row paths of the tiled kernel.
template <bool ATOMIC_COVERAGE, typename GetFrame>
__device__ __forceinline__ void lanczos_stream_tile(const GetFrame& get, int count, const WarpParams& p, int ox0,
                                                    int oy0, const float4* lut, float* boxes) {
  const int tx = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  const int x = ox0 + 64 * (ty % kSColGroups) + tx;
  // wave-uniform (row bases and store addresses stay scalar)
  const int y_first = __builtin_amdgcn_readfirstlane(oy0 + kSLaneRows * (ty / kSColGroups));
  const int W = p.width, H = p.height;
  const int ox1 = min(ox0 + kSOutW, W) - 1, oy1 = min(oy0 + kSOutH, H) - 1;
  const bool clamp_on = p.skip_nonfinite && p.clamp >= 0.0f;
  constexpr int kGroups = kSLaneRows / kLaneRows;
  constexpr int kStripsPerGroup = kLaneRows / kSStrip;
  unsigned long long cov[kGroups];
#pragma unroll
  for (int g = 0; g < kGroups; ++g) cov[g] = 0ull;
  const int bytes = W * H * 4;  // < 2^31 (checked by the host)
  const double xd = static_cast<double>(x);

  // Frame f's box is boxes[f & 1]; frame f + 1's is staged while frame f is
  // computed (or right after it, see below).  One barrier per frame: after it
  // every wave's loads of frame f's box have landed (each wave waits for its
  // own first) and every wave is done with frame f - 1, whose buffer frame
  // f + 1's box then takes.
  auto geometry = [&](const FrameRef& r) {
    return box_geometry(r, !clamp_on && r.src_flag == 0u, ox0, oy0, ox1, oy1, W, H);
  };
  __syncthreads();  // the buffers' previous users (an earlier item of a batched launch) are done
  BoxGeom geo = geometry(get(0));
  if (geo.ok) stage_box_lds(get(0), geo, W, bytes, boxes);
  for (int f = 0; f < count; ++f) {
    const FrameRef ref = get(f);
    const bool finite = !clamp_on && ref.src_flag == 0u;
    BoxGeom next{0, 0, 0, 0, false, false};
    if (f + 1 < count) next = geometry(get(f + 1));
    if (geo.ok || next.ok) {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
    }
    // The next box streams in during this frame only when this frame's
    // strips read nothing but their box (it fits and the tile lies inside
    // the footprint): a vector load of the frame issued meanwhile would wait
    // for the staging loads.  Otherwise it is staged after this frame.
    const bool ahead = geo.ok && geo.inside;
    if (next.ok && ahead) stage_box_lds(get(f + 1), next, W, bytes, boxes + ((f + 1) & 1) * kBoxFloats);
    const float* box = boxes + (f & 1) * kBoxFloats;
    bool bad = false;
    if (finite) {
      const StreamFrame fr = stream_frame(ref, bytes);
      const double px = ref.inv[0] * xd, py = ref.inv[3] * xd;
#pragma unroll
      for (int g = 0; g < kGroups; ++g) {
#pragma unroll 1
        for (int s = 0; s < kStripsPerGroup; ++s) {
          unsigned long long c = 0ull;
          const int y0 = y_first + g * kLaneRows + s * kSStrip;
          if (geo.ok)
            bad |= stream_strip<kSStrip, true>(ref, fr, lut, px, py, x, y0, W, H, p.fill, c, box, geo);
          else
            bad |= stream_strip<kSStrip, false>(ref, fr, lut, px, py, x, y0, W, H, p.fill, c, box, geo);
          cov[g] += c << (4 * kSStrip * s);
        }
      }
    } else {
      // Non-finite source (CPU semantics: table rows, then the non-finite
      // results recomputed tap by tap) or resident semantics with skipping /
      // clamp (direct rows).
      TileFrame tf{};
      tf.src = ref.src;
      tf.dst = ref.dst;
      tf.m0 = ref.inv[0]; tf.m1 = ref.inv[1]; tf.m2 = ref.inv[2];
      tf.m3 = ref.inv[3]; tf.m4 = ref.inv[4]; tf.m5 = ref.inv[5];
      tf.fast = !clamp_on && !p.skip_nonfinite;
      tf.recheck = true;
#pragma unroll
      for (int g = 0; g < kGroups; ++g) {
        const int yg = y_first + g * kLaneRows;
        bool b = tf.fast ? column_global_rows(tf, lut, x, yg, W, H, p.fill, cov[g])
                         : column_direct_rows(tf, p, x, yg, cov[g]);
        if (tf.fast && b) b = recheck_column(tf, x, yg, W, H);
        bad |= b;
      }
      drain_vector_memory();
    }
    if (__any(bad) && tx == 0) atomicOr(ref.dst_flag, 1u);
    if (next.ok && !ahead) stage_box_lds(get(f + 1), next, W, bytes, boxes + ((f + 1) & 1) * kBoxFloats);
    geo = next;
  }
  if (p.coverage && x < W) {
#pragma unroll
    for (int g = 0; g < kGroups; ++g) {
#pragma unroll 1
      for (int k = 0; k < kLaneRows; ++k) {
        const int y = y_first + g * kLaneRows + k;
        if (y >= H) break;
        const float add = static_cast<float>((cov[g] >> (4 * k)) & 15ull);
        const long long o = static_cast<long long>(y) * W + x;
        if constexpr (ATOMIC_COVERAGE) {
          if (add != 0.0f) atomicAdd(p.coverage + o, add);
        } else {
          st_global(p.coverage, o, ld_global(p.coverage, o) + add);
        }
      }
    }
  }
}

#ifndef GH_WARP_STREAM_WAVES_PER_EU
#define GH_WARP_STREAM_WAVES_PER_EU 3
#endif
__global__ __launch_bounds__(kSThreads) __attribute__((amdgpu_waves_per_eu(GH_WARP_STREAM_WAVES_PER_EU))) void
warp_lanczos3_stream_kernel(WarpBatch b, int count, WarpParams p, int tiles_x) {
  __shared__ float4 lut[kLutLdsVec4];
  __shared__ float boxes[2 * kBoxFloats];
  load_lut(p.lanczos_lut, lut);
  const unsigned nb = gridDim.x, bid = blockIdx.x;
  const unsigned xcd = bid & 7u, q = nb >> 3, r = nb & 7u;
  const unsigned logical = xcd * q + min(xcd, r) + (bid >> 3);
  const int ox0 = static_cast<int>(logical % tiles_x) * kSOutW;
  const int oy0 = static_cast<int>(logical / tiles_x) * kSOutH;
  // The source flags are read by scalar loads: a vector load issued while a
  // box is being staged would wait for the staging loads (vector memory
  // operations of a wave complete in order).
  typedef __attribute__((address_space(4))) const unsigned const_u32;
  auto get = [&](int f) {
    const double* m = b.inv[f];
    return FrameRef{b.src[f], b.dst[f], {m[0], m[1], m[2], m[3], m[4], m[5]}, ((const const_u32*)b.src_flags)[f],
                    b.nonfinite[f]};
  };
  __syncthreads();  // the table
  lanczos_stream_tile<false>(get, count, p, ox0, oy0, lut, boxes);
}

// ------------------------------------------------- batched multi-job warp
// Kernels over the work items of many jobs: item = (job, group of up to
// kWarpBatch frames, output tile).  Blocks stride over the item list (one
// item per block when the grid covers it); item_job[] names each item's job,
// so a block re-reads the job descriptor only when the job changes.
template <bool LANCZOS>
__device__ __forceinline__ void warp_batched_items(const BatchWarpJob* __restrict__ jobs,
                                                   const unsigned short* __restrict__ item_job,
                                                   const BatchWarpFrame* __restrict__ frames, long long total_items,
                                                   int order_stride, const float4* lut, float* tile) {
  int cur = -1;
  BatchWarpJob job{};
  for (long long it = blockIdx.x; it < total_items; it += gridDim.x) {
    // Optional interleaving: visit items with a stride coprime to the total so
    // concurrent blocks spread over jobs.
    const long long item = order_stride > 1 ? (it * order_stride) % total_items : it;
    const int j = __builtin_amdgcn_readfirstlane(static_cast<int>(item_job[item]));
    if (j != cur) {
      job = load_uniform(jobs + j);
      cur = j;
    }
    const long long local = item - job.first_item;
    const int batch = static_cast<int>(local / job.tiles), t = static_cast<int>(local % job.tiles);
    const int f0 = job.first_frame + batch * job.frames_per_item;
    const int count = min(job.frames_per_item, job.first_frame + job.n_frames - f0);
    const int ox0 = (t % job.tiles_x) * kOutW, oy0 = (t / job.tiles_x) * kOutH;
    WarpParams p{};
    p.height = job.height;
    p.width = job.width;
    p.interpolation = job.interpolation;
    p.fill = job.fill;
    p.clamp = job.clamp;
    p.skip_nonfinite = job.skip_nonfinite;
    p.affine = 1;
    p.coverage = job.coverage;
    // Descriptors and source flags by scalar loads (see the streamed kernel).
    typedef __attribute__((address_space(4))) const unsigned const_u32;
    auto get = [&](int f) {
      const BatchWarpFrame fr = load_scalar(frames + (f0 + f));
      return FrameRef{fr.src, fr.dst, {fr.inv[0], fr.inv[1], fr.inv[2], fr.inv[3], fr.inv[4], fr.inv[5]},
                      *(const const_u32*)fr.src_flag, fr.dst_flag};
    };
    if constexpr (LANCZOS) {
      // the same kernel choice as a per-job launch of this 8-frame group
      const int streamed = __builtin_amdgcn_readfirstlane(*(const global_u32*)&frames[f0].stream);
      if (streamed) lanczos_stream_tile<true>(get, count, p, ox0, oy0, lut, tile);
      else lanczos_tile<true>(get, count, p, ox0, oy0, lut, tile);
    } else {
      bilinear_tile<true>(get, count, p, ox0, oy0);
    }
  }
}

// The Lanczos3 jobs' items and the bilinear jobs' items run as two launches:
// the bilinear taps are a few gathers per pixel, and a kernel holding the
// Lanczos3 paths' registers and LDS would run them at a third of the
// occupancy they need to hide the gathers' latency.
__global__ __launch_bounds__(kTiledThreads) __attribute__((amdgpu_waves_per_eu(3))) void warp_batched_kernel(
    const BatchWarpJob* __restrict__ jobs, const unsigned short* __restrict__ item_job,
    const BatchWarpFrame* __restrict__ frames, long long total_items, const float4* lut_g, int order_stride) {
  __shared__ float4 lut[kLutLdsVec4];
  __shared__ float tile[kStreamLdsFloats];
  load_lut(lut_g, lut);
  warp_batched_items<true>(jobs, item_job, frames, total_items, order_stride, lut, tile);
}

__global__ __launch_bounds__(kTiledThreads) void warp_batched_bilinear_kernel(
    const BatchWarpJob* __restrict__ jobs, const unsigned short* __restrict__ item_job,
    const BatchWarpFrame* __restrict__ frames, long long total_items, int order_stride) {
  warp_batched_items<false>(jobs, item_job, frames, total_items, order_stride, nullptr, nullptr);
}

__global__ void snapshot_warp_flags_kernel(WarpBatch b, int count, unsigned* snap) {
  const int j = threadIdx.x;
  if (j < count) {
    snap[j] = *b.nonfinite[j];
    *b.nonfinite[j] = 0u;
  }
}

// ------------------------------------------------------------------------
// Mean / sigma-clip integration that samples every frame itself: from the
// resident stack as stored (SAMPLER 0) or through each frame's matrix with
// bilinear (1) or Lanczos3 (2) taps, without writing a warped stack.  Per
// pixel, over the frames of positive finite weight in order
// (cpp/cuda/integration_kernels.cu:1743-1900, :3840-4123):
//  * mean (mode 0): sum w*x and sum w over finite samples, master = sum / sum w;
//  * sigma clip (mode 1): mean and population std of the finite samples,
//    thresholds mean -/+ k*std; with winsorize, the samples clipped to them give
//    a new mean and std and new thresholds; samples outside are rejected (low /
//    high counts) and the weighted mean of the rest is the master.
// All sums are float32 in frame order, as in the reference kernels.  Each
// frame is sampled once (into registers or LDS, see STORE) instead of once per
// pass as the reference does; beyond the stores' capacity every pass samples
// again.
// A sample outside the footprint is neither counted nor covered; one inside
// it adds to the geometric coverage and counts only when finite.
template <int SAMPLER>
__device__ __forceinline__ float clip_sample(const ClipArgs& a, int k, int x, int y, bool& footprint) {
  const float* frame = table_at(a.frames, k);
  if (SAMPLER == 0) {
    footprint = true;
    return ld_global(frame, static_cast<long long>(y) * a.width + x);
  }
  const double* m = a.inverses + 9 * k;
  double sx, sy;
  footprint = source_coords(m, a.affine, x, y, sx, sy) && isfinite(sx) && isfinite(sy);
  if (SAMPLER == 1) {
    footprint = footprint && sx >= 0.0 && sx <= static_cast<double>(a.width - 1) && sy >= 0.0 &&
                sy <= static_cast<double>(a.height - 1);
    if (!footprint) return NAN;
    const double fx0 = floor(sx), fy0 = floor(sy);
    const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0);
    const int x1 = min(x0 + 1, a.width - 1), y1 = min(y0 + 1, a.height - 1);
    const float tx = static_cast<float>(sx - fx0), ty = static_cast<float>(sy - fy0);
    const float a00 = ld_global(frame, static_cast<long long>(y0) * a.width + x0);
    const float a01 = ld_global(frame, static_cast<long long>(y0) * a.width + x1);
    const float a10 = ld_global(frame, static_cast<long long>(y1) * a.width + x0);
    const float a11 = ld_global(frame, static_cast<long long>(y1) * a.width + x1);
    const float top = a00 * (1.0f - tx) + a01 * tx;
    const float bottom = a10 * (1.0f - tx) + a11 * tx;
    return top * (1.0f - ty) + bottom * ty;
  }
  footprint = footprint && lanczos_footprint(sx, sy, a.width, a.height);
  if (!footprint) return NAN;
  WarpParams p{};
  p.fill = NAN;
  p.clamp = a.clamp;
  p.skip_nonfinite = 1;
  bool covered;
  const float v = lanczos_direct(frame, a.width, sx, sy, p, 1.0e-6f, covered);
  footprint = covered;
  return v;
}

// Sample storage: STORE 0 registers (CAP compile-time slots, sampler 0),
// 1 LDS (n slots per thread in dynamic shared memory, 128-thread blocks up to
// 128 frames and 64-thread blocks up to 256, at most 64 KB per block: the
// warped samplers' per-frame code is too large to unroll CAP times), 2 none
// (every pass samples again; the single-pass weighted mean always, and the
// sigma clip where LDS occupancy would cost more than sampling again).

// TILE: tile-local weights.  Table entry k marked in a.target has weight
// weights[k] * tile_map[pixel] at the pixel; an entry whose weight there is
// not positive and finite does not count at that pixel (a NaN sample).
template <int SAMPLER, int CAP, int STORE, bool TILE = false>
__global__ __launch_bounds__(256) void clip_integrate_kernel(ClipArgs a) {
  extern __shared__ float clip_lds[];
  const int nthreads = STORE == 1 ? static_cast<int>(blockDim.x) : 256;
  const long long p = static_cast<long long>(blockIdx.x) * nthreads + threadIdx.x;
  const bool live = p < a.npix;
  const long long pp = live ? p : 0;
  const int y = static_cast<int>(pp / a.width), x = static_cast<int>(pp - static_cast<long long>(y) * a.width);
  float v[CAP > 0 ? CAP : 1];
  float geometric = 0.0f;
  float tile_multiplier = 1.0f;
  if constexpr (TILE) tile_multiplier = ld_global(a.tile_map, pp);
  auto weight_of = [&](int k) -> float {
    float w = ld_global(a.weights, k);
    if constexpr (TILE) {
      if (a.target[k] != 0) w *= tile_multiplier;
    }
    return w;
  };
  auto sample = [&](int k, float& out) -> bool {
    bool fp;
    out = clip_sample<SAMPLER>(a, k, x, y, fp);
    if constexpr (TILE) {
      const float w = weight_of(k);
      if (!(w > 0.0f) || !isfinite(w)) out = NAN;
    }
    return fp;
  };
  if constexpr (STORE == 0) {
#pragma unroll
    for (int k = 0; k < CAP; ++k) {
      if (k < a.n) {
        float s;
        const bool fp = sample(k, s);
        geometric += fp ? 1.0f : 0.0f;
        v[k] = (fp && isfinite(s)) ? s : NAN;
      }
    }
  } else {
    for (int k = 0; k < a.n; ++k) {
      float s;
      const bool fp = sample(k, s);
      geometric += fp ? 1.0f : 0.0f;
      if constexpr (STORE == 1) clip_lds[k * nthreads + threadIdx.x] = (fp && isfinite(s)) ? s : NAN;
    }
  }
  if (!live) return;  // no barrier below: each thread reads only its own slots
  // value k of the pass (NaN when it does not count)
  auto value = [&](int k) -> float {
    if constexpr (STORE == 0) {
      return v[k];
    } else if constexpr (STORE == 1) {
      return clip_lds[k * nthreads + threadIdx.x];
    } else {
      float s;
      const bool fp = sample(k, s);
      return (fp && isfinite(s)) ? s : NAN;
    }
  };
  constexpr int kUnroll = STORE == 0 ? CAP : 1;
  float master = 0.0f, wsum = 0.0f, coverage = 0.0f, low = 0.0f, high = 0.0f;
  if (a.mode == 0) {
    float sum = 0.0f;
#pragma unroll kUnroll
    for (int k = 0; k < (STORE == 0 ? CAP : a.n); ++k) {
      if (k >= a.n) break;
      const float s = value(k);
      if (!isfinite(s)) continue;
      const float w = weight_of(k);
      sum += s * w;
      wsum += w;
      coverage += 1.0f;
    }
    master = wsum > 0.0f ? sum / wsum : 0.0f;
  } else {
    float mean = 0.0f, count = 0.0f;
#pragma unroll kUnroll
    for (int k = 0; k < (STORE == 0 ? CAP : a.n); ++k) {
      if (k >= a.n) break;
      const float s = value(k);
      if (isfinite(s)) {
        mean += s;
        count += 1.0f;
      }
    }
    if (count > 0.0f) {
      mean /= count;
      float var = 0.0f;
#pragma unroll kUnroll
      for (int k = 0; k < (STORE == 0 ? CAP : a.n); ++k) {
        if (k >= a.n) break;
        const float s = value(k);
        if (isfinite(s)) {
          const float d = s - mean;
          var += d * d;
        }
      }
      float center = mean, scale = sqrtf(var / count);
      float lo = center - a.low_sigma * scale, hi = center + a.high_sigma * scale;
      if (a.winsorize) {
        float wmean = 0.0f;
#pragma unroll kUnroll
        for (int k = 0; k < (STORE == 0 ? CAP : a.n); ++k) {
          if (k >= a.n) break;
          float s = value(k);
          if (!isfinite(s)) continue;
          s = s < lo ? lo : (s > hi ? hi : s);
          wmean += s;
        }
        wmean /= count;
        float wvar = 0.0f;
#pragma unroll kUnroll
        for (int k = 0; k < (STORE == 0 ? CAP : a.n); ++k) {
          if (k >= a.n) break;
          float s = value(k);
          if (!isfinite(s)) continue;
          s = s < lo ? lo : (s > hi ? hi : s);
          const float d = s - wmean;
          wvar += d * d;
        }
        center = wmean;
        scale = sqrtf(wvar / count);
        lo = center - a.low_sigma * scale;
        hi = center + a.high_sigma * scale;
      }
      float sum = 0.0f;
#pragma unroll kUnroll
      for (int k = 0; k < (STORE == 0 ? CAP : a.n); ++k) {
        if (k >= a.n) break;
        const float s = value(k);
        if (!isfinite(s)) continue;
        if (s < lo) {
          low += 1.0f;
          continue;
        }
        if (s > hi) {
          high += 1.0f;
          continue;
        }
        const float w = weight_of(k);
        sum += s * w;
        wsum += w;
        coverage += 1.0f;
      }
      master = wsum > 0.0f ? sum / wsum : 0.0f;
    }
  }
  if (a.master) st_global(a.master, p, master);
  if (a.weight) st_global(a.weight, p, wsum);
  if (a.coverage) st_global(a.coverage, p, coverage);
  if (a.low) st_global(a.low, p, low);
  if (a.high) st_global(a.high, p, high);
  if (a.geometric) st_global(a.geometric, p, geometric);
}

// ------------------------------------------------------------------------
// The fused Lanczos3 integrations of affine maps over 128 x 16 output tiles
// (the same statistics as clip_integrate_kernel<2, ...>, sample for sample
// and sum for sum).  Per frame the tile's source box is staged into LDS
// (box_geometry / stage_box_lds of the streamed warp) and every lane samples
// its column of 8 rows from it with the resident tap rules
// (lanczos_direct_window: the same taps and arithmetic as from the frame,
// which a window outside the box, or a frame whose box does not fit, reads
// instead).  The per-pixel state of the lane's 8 rows stays in registers; the
// sigma clip takes its passes (mean, variance, winsorized mean and variance,
// final) over the frames again, sampling each time, where the per-pixel
// kernel gathers 36 taps per sample from global memory and keeps the samples
// in LDS.  MODE 0: weighted mean, 1: sigma clip.
#ifndef GH_CLIP_ROWS
#define GH_CLIP_ROWS 8
#endif
constexpr int kClipRows = GH_CLIP_ROWS;
constexpr int kClipTileH = kClipRows * (kSThreads / 64 / kSColGroups);

template <int MODE, bool CLAMP>
__global__ __launch_bounds__(kSThreads) __attribute__((amdgpu_waves_per_eu(2, MODE == 0 ? 4 : 3))) void clip_lanczos_tile_kernel(
    ClipArgs a, int tiles_x) {
  __shared__ float box[kBoxFloats];
  const unsigned nb = gridDim.x, bid = blockIdx.x;
  const unsigned xcd = bid & 7u, q = nb >> 3, rr = nb & 7u;
  const unsigned logical = xcd * q + min(xcd, rr) + (bid >> 3);
  const int ox0 = static_cast<int>(logical % tiles_x) * kSOutW;
  const int oy0 = static_cast<int>(logical / tiles_x) * kClipTileH;
  const int W = a.width, H = a.height;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int x = ox0 + 64 * (ty % kSColGroups) + tx;
  const int y_first = oy0 + kClipRows * (ty / kSColGroups);
  const int ox1 = min(ox0 + kSOutW, W) - 1, oy1 = min(oy0 + kClipTileH, H) - 1;
  const int bytes = W * H * 4;  // frames of < 2^31 bytes (checked by the host)
  WarpParams p{};
  p.fill = NAN;
  p.clamp = a.clamp;
  p.skip_nonfinite = 1;
  typedef __attribute__((address_space(4))) const double const_f64;

  // The sample of frame k at row r of this lane's column (NaN when it does
  // not count); `covered` reports its geometric coverage.
  auto sample = [&](const BoxGeom& g, const FrameRef& ref, int r, bool& covered) -> float {
    covered = false;
    const int y = y_first + r;
    if (x >= W || y >= H) return NAN;
    double sx, sy;
    source_coords(ref.inv, 1, x, y, sx, sy);
    if (!lanczos_footprint(sx, sy, W, H)) return NAN;
    const double fx0 = floor(sx), fy0 = floor(sy);
    const int x0 = static_cast<int>(fx0), y0 = static_cast<int>(fy0);
    float v;
    if (g.ok && x0 - 2 >= g.bx0 && x0 + 3 < g.bx0 + g.bw && y0 - 2 >= g.by0 && y0 + 3 < g.by0 + g.bh) {
      v = lanczos_direct_window<const float*, CLAMP ? 1 : 0>(box + ((y0 - 2 - g.by0) * kBoxStride + (x0 - 2 - g.bx0)),
                                                             kBoxStride, sx, sy, fx0, fy0, p, 1.0e-6f, covered);
    } else {
      v = lanczos_direct_window<const global_cf32*, CLAMP ? 1 : 0>(
          (const global_cf32*)(ref.src + static_cast<long long>(y0 - 2) * W + (x0 - 2)), W, sx, sy, fx0, fy0, p,
          1.0e-6f, covered);
    }
    return (covered && isfinite(v)) ? v : NAN;
  };
  // One pass over the frames in order: body(k, weight, row, sample, covered).
  auto pass = [&](auto&& body) {
#pragma unroll 1
    for (int k = 0; k < a.n; ++k) {
      const double* m = a.inverses + 9 * k;
      FrameRef ref{};
      ref.src = table_at(a.frames, k);
      for (int i = 0; i < 6; ++i) ref.inv[i] = ((const const_f64*)m)[i];
      const float w = ((const const_f32*)a.weights)[k];
      const BoxGeom g = box_geometry(ref, true, ox0, oy0, ox1, oy1, W, H);
      if (g.ok) {
        __syncthreads();  // every wave is done with the previous frame's box
        stage_box_lds(ref, g, W, bytes, box);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();  // every wave's part of the box has landed
      }
      // compile-time rows: the per-pixel arrays stay in registers
      static_for<kClipRows>([&](auto R) __attribute__((always_inline)) {
        constexpr int r = decltype(R)::value;
        bool covered;
        const float v = sample(g, ref, r, covered);
        body(w, r, v, covered);
      });
    }
  };

  // The pass bodies add +0.0 instead of skipping a sample (branch-free, so
  // the unrolled rows keep compile-time indices); no sum here is ever -0.0,
  // so the sums are the per-pixel kernel's bit for bit.
  float geometric[kClipRows], s0[kClipRows], s1[kClipRows], s2[kClipRows];
#pragma unroll
  for (int r = 0; r < kClipRows; ++r) geometric[r] = s0[r] = s1[r] = s2[r] = 0.0f;
  float master[kClipRows], wsum[kClipRows], coverage[kClipRows], low[kClipRows], high[kClipRows];
  if constexpr (MODE == 0) {
    // s0 = sum w x, s1 = sum w, s2 = count
    pass([&](float w, int r, float v, bool covered) {
      const bool fin = isfinite(v);
      geometric[r] += covered ? 1.0f : 0.0f;
      s0[r] += fin ? v * w : 0.0f;
      s1[r] += fin ? w : 0.0f;
      s2[r] += fin ? 1.0f : 0.0f;
    });
#pragma unroll
    for (int r = 0; r < kClipRows; ++r) {
      master[r] = s1[r] > 0.0f ? s0[r] / s1[r] : 0.0f;
      wsum[r] = s1[r];
      coverage[r] = s2[r];
      low[r] = high[r] = 0.0f;
    }
  } else {
    // mean and count (s0, s1), then the variance (s2)
    pass([&](float, int r, float v, bool covered) {
      const bool fin = isfinite(v);
      geometric[r] += covered ? 1.0f : 0.0f;
      s0[r] += fin ? v : 0.0f;
      s1[r] += fin ? 1.0f : 0.0f;
    });
    float center[kClipRows], lo[kClipRows], hi[kClipRows];
#pragma unroll
    for (int r = 0; r < kClipRows; ++r) center[r] = s1[r] > 0.0f ? s0[r] / s1[r] : 0.0f;
    pass([&](float, int r, float v, bool) {
      const float d = v - center[r];
      s2[r] += isfinite(v) ? d * d : 0.0f;
    });
#pragma unroll
    for (int r = 0; r < kClipRows; ++r) {
      const float scale = s1[r] > 0.0f ? sqrtf(s2[r] / s1[r]) : 0.0f;
      lo[r] = center[r] - a.low_sigma * scale;
      hi[r] = center[r] + a.high_sigma * scale;
    }
    if (a.winsorize) {
#pragma unroll
      for (int r = 0; r < kClipRows; ++r) s0[r] = s2[r] = 0.0f;
      pass([&](float, int r, float v, bool) {
        const float c = v < lo[r] ? lo[r] : (v > hi[r] ? hi[r] : v);
        s0[r] += isfinite(v) ? c : 0.0f;
      });
#pragma unroll
      for (int r = 0; r < kClipRows; ++r) center[r] = s1[r] > 0.0f ? s0[r] / s1[r] : 0.0f;
      pass([&](float, int r, float v, bool) {
        const float c = v < lo[r] ? lo[r] : (v > hi[r] ? hi[r] : v);
        const float d = c - center[r];
        s2[r] += isfinite(v) ? d * d : 0.0f;
      });
#pragma unroll
      for (int r = 0; r < kClipRows; ++r) {
        const float scale = s1[r] > 0.0f ? sqrtf(s2[r] / s1[r]) : 0.0f;
        lo[r] = center[r] - a.low_sigma * scale;
        hi[r] = center[r] + a.high_sigma * scale;
      }
    }
#pragma unroll
    for (int r = 0; r < kClipRows; ++r) master[r] = wsum[r] = coverage[r] = low[r] = high[r] = 0.0f;
    pass([&](float w, int r, float v, bool) {
      const bool fin = isfinite(v) && s1[r] > 0.0f;
      const bool below = fin && v < lo[r];
      const bool above = fin && !below && v > hi[r];
      const bool keep = fin && !below && !above;
      low[r] += below ? 1.0f : 0.0f;
      high[r] += above ? 1.0f : 0.0f;
      master[r] += keep ? v * w : 0.0f;
      wsum[r] += keep ? w : 0.0f;
      coverage[r] += keep ? 1.0f : 0.0f;
    });
#pragma unroll
    for (int r = 0; r < kClipRows; ++r) master[r] = wsum[r] > 0.0f ? master[r] / wsum[r] : 0.0f;
  }
  if (x >= W) return;
#pragma unroll
  for (int r = 0; r < kClipRows; ++r) {
    const int y = y_first + r;
    if (y >= H) break;
    const long long o = static_cast<long long>(y) * W + x;
    if (a.master) st_global(a.master, o, master[r]);
    if (a.weight) st_global(a.weight, o, wsum[r]);
    if (a.coverage) st_global(a.coverage, o, coverage[r]);
    if (a.low) st_global(a.low, o, low[r]);
    if (a.high) st_global(a.high, o, high[r]);
    if (a.geometric) st_global(a.geometric, o, geometric[r]);
  }
}

// ------------------------------------------------------------------------
// Translations (cpp/cuda/warp_kernels.cu:103-190): the integer form copies
// in[y - dy][x - dx] (fill outside); the bilinear form samples at
// (x - dx, y - dy) computed in float32, valid 0 <= s <= dim - 1, the far
// neighbour clamped to the edge.  Out of place (the host swaps slots); the
// geometric coverage accumulator gains 1 where the pixel is inside.
template <bool BILINEAR>
__global__ __launch_bounds__(256) void translate_kernel(const float* src, float* dst, int width, int height, int idx,
                                                        int idy, float fdx, float fdy, float fill, float* coverage,
                                                        unsigned* flag) {
  bool bad = false;
  const long long n = static_cast<long long>(width) * height;
  for (long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * 256) {
    const int y = static_cast<int>(i / width), x = static_cast<int>(i - static_cast<long long>(y) * width);
    float v = fill;
    bool inside;
    if (!BILINEAR) {
      const int sx = x - idx, sy = y - idy;
      inside = sx >= 0 && sx < width && sy >= 0 && sy < height;
      if (inside) v = ld_global(src, static_cast<long long>(sy) * width + sx);
    } else {
      const float sx = static_cast<float>(x) - fdx, sy = static_cast<float>(y) - fdy;
      inside = !(sx < 0.0f || sx > static_cast<float>(width - 1) || sy < 0.0f || sy > static_cast<float>(height - 1));
      if (inside) {
        const int x0 = static_cast<int>(floorf(sx)), y0 = static_cast<int>(floorf(sy));
        const int x1 = x0 + 1 < width ? x0 + 1 : x0, y1 = y0 + 1 < height ? y0 + 1 : y0;
        const float tx = sx - static_cast<float>(x0), ty = sy - static_cast<float>(y0);
        const float a00 = ld_global(src, static_cast<long long>(y0) * width + x0);
        const float a01 = ld_global(src, static_cast<long long>(y0) * width + x1);
        const float a10 = ld_global(src, static_cast<long long>(y1) * width + x0);
        const float a11 = ld_global(src, static_cast<long long>(y1) * width + x1);
        const float top = a00 * (1.0f - tx) + a01 * tx;
        const float bottom = a10 * (1.0f - tx) + a11 * tx;
        v = top * (1.0f - ty) + bottom * ty;
      }
    }
    st_global(dst, i, v);
    bad |= !isfinite(v);
    if (coverage && inside) st_global(coverage, i, ld_global(coverage, i) + 1.0f);
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

}  // namespace

void make_lanczos_lut(float* table) {
  // Normalised weights w_k(f), FP64 with the CPU sinc definition.
  auto weights = [](double f, double (&w)[6]) {
    constexpr double kPi = 3.14159265358979323846;
    double sum = 0.0;
    for (int k = 0; k < 6; ++k) {
      const double d = std::fabs(f - static_cast<double>(k - 2));
      double v = 0.0;
      if (d < 3.0) {
        const double a = d < 1.0e-8 ? 1.0 : std::sin(kPi * d) / (kPi * d);
        const double b3 = d / 3.0 < 1.0e-8 ? 1.0 : std::sin(kPi * d / 3.0) / (kPi * d / 3.0);
        v = a * b3;
      }
      // Integer offsets: the mathematical kernel is exactly a delta.
      if (d != 0.0 && d == std::floor(d)) v = 0.0;
      w[k] = v;
      sum += v;
    }
    for (int k = 0; k < 6; ++k) w[k] /= sum;
  };
  for (int i = 0; i <= kLanczosLutIntervals; ++i) {
    double y0[6], y1[6], y2[6];
    weights((i + 0.0) / kLanczosLutIntervals, y0);
    weights((i + 0.5) / kLanczosLutIntervals, y1);
    weights((i + 1.0) / kLanczosLutIntervals, y2);
    float* e = table + i * kLanczosLutStride;
    for (int k = 0; k < 6; ++k) {
      // p(u) = y0 + b u + c u^2 through (0, y0), (1/2, y1), (1, y2)
      const double c = 2.0 * y2[k] - 4.0 * y1[k] + 2.0 * y0[k];
      const double b = y2[k] - y0[k] - c;
      e[3 * k + 0] = static_cast<float>(y0[k]);
      e[3 * k + 1] = static_cast<float>(b);
      e[3 * k + 2] = static_cast<float>(c);
    }
    for (int k = 18; k < kLanczosLutStride; ++k) e[k] = 0.0f;
    if (i == kLanczosLutIntervals) {  // f == 1 exactly: constant delta on tap 3
      for (int k = 0; k < 6; ++k) {
        e[3 * k + 0] = k == 3 ? 1.0f : 0.0f;
        e[3 * k + 1] = e[3 * k + 2] = 0.0f;
      }
    }
  }
}

void warp_tile_shape(int* tile_w, int* tile_h) {
  *tile_w = kOutW;
  *tile_h = kOutH;
}

long long warp_batched_items(const BatchWarpJob& job) {
  return static_cast<long long>((job.n_frames + job.frames_per_item - 1) / job.frames_per_item) * job.tiles;
}

unsigned warp_batched_max_blocks(bool lanczos) {
  int dev = 0, cus = 0, per_cu = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device CU count");
  if (lanczos) {
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, warp_batched_kernel, kTiledThreads, 0),
          "warp occupancy");
  } else {
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, warp_batched_bilinear_kernel, kTiledThreads, 0),
          "warp occupancy");
  }
  return static_cast<unsigned>(std::max(1, cus) * std::max(1, per_cu));
}

void launch_warp_batched(bool lanczos, const BatchWarpJob* jobs, const unsigned short* item_job,
                         const BatchWarpFrame* frames, long long total_items, const float4* lut, unsigned grid,
                         int order_stride, hipStream_t stream) {
  if (total_items <= 0) return;
  const unsigned g = static_cast<unsigned>(std::min<long long>(std::max(grid, 1u), total_items));
  if (lanczos) {
    warp_batched_kernel<<<g, kTiledThreads, 0, stream>>>(jobs, item_job, frames, total_items, lut, order_stride);
  } else {
    warp_batched_bilinear_kernel<<<g, kTiledThreads, 0, stream>>>(jobs, item_job, frames, total_items, order_stride);
  }
  check(hipGetLastError(), "batched warp launch");
}

void launch_snapshot_warp_flags(const WarpBatch& b, int count, unsigned* snap, hipStream_t stream) {
  snapshot_warp_flags_kernel<<<1, 64, 0, stream>>>(b, count, snap);
  check(hipGetLastError(), "snapshot_warp_flags launch");
}

// Kernel choice for an affine Lanczos3 group of frames (a per-job launch
// of up to kWarpBatch frames, or the same group inside a batched launch).
// The streamed kernel needs 32-bit byte offsets into a frame and room for its
// window; it pays off when every map is near the identity: |m1| small (the x
// fraction wraps at most once per strip) and |m4 - 1| small enough that the y
// table interval anchored at a strip's middle row stays within [-0.25, 1.25]
// over the strip.  GLASS_HIP_WARP=tiled / stream forces a kernel (stream only
// where its size conditions hold); the streamed kernel is correct for any
// affine map.
bool lanczos_stream_eligible(const double (*inv)[9], int count, int width, int height) {
  const long long bytes = static_cast<long long>(width) * height * 4;
  if (width < 64 || height < 64 || bytes >= (1ll << 31)) return false;
  const char* force = std::getenv("GLASS_HIP_WARP");
  if (force && std::strcmp(force, "tiled") == 0) return false;
  if (force && std::strcmp(force, "stream") == 0) return true;
  const double max_dy = 0.25 / (0.5 * (kSStrip - 1) * kLanczosLutIntervals);
  for (int f = 0; f < count; ++f) {
    if (!(std::fabs(inv[f][1]) <= 0.1 && std::fabs(inv[f][4] - 1.0) <= max_dy)) return false;
  }
  return true;
}

void launch_warp(const WarpBatch& batch, int count, const WarpParams& p, hipStream_t stream) {
  const dim3 grid((p.width + kTileX - 1) / kTileX, (p.height + kTileY - 1) / kTileY);
  const dim3 block(kTileX * kTileY);
  if (p.interpolation == 3 && p.affine && p.lanczos_lut != nullptr &&
      lanczos_stream_eligible(batch.inv, count, p.width, p.height)) {
    const int tiles_x = (p.width + kSOutW - 1) / kSOutW;
    const int tiles_y = (p.height + kSOutH - 1) / kSOutH;
    warp_lanczos3_stream_kernel<<<tiles_x * tiles_y, kSThreads, 0, stream>>>(batch, count, p, tiles_x);
  } else if (p.interpolation == 3 && p.affine && p.lanczos_lut != nullptr) {
    const int tiles_x = (p.width + kOutW - 1) / kOutW;
    const int tiles_y = (p.height + kOutH - 1) / kOutH;
    warp_lanczos3_tiled_kernel<<<tiles_x * tiles_y, kTiledThreads, 0, stream>>>(batch, count, p, tiles_x);
  } else if (p.interpolation == 1) {
    if (p.affine && static_cast<long long>(p.width) * p.height < (1ll << 30)) {
      const int tiles_x = (p.width + kOutW - 1) / kOutW;
      const int tiles_y = (p.height + kOutH - 1) / kOutH;
      warp_bilinear_tiled_kernel<<<tiles_x * tiles_y, kTiledThreads, 0, stream>>>(batch, count, p, tiles_x);
    } else if (p.affine) {
      warp_bilinear_kernel<true><<<grid, block, 0, stream>>>(batch, count, p);
    } else {
      warp_bilinear_kernel<false><<<grid, block, 0, stream>>>(batch, count, p);
    }
  } else {
    if (p.affine) warp_lanczos3_kernel<true><<<grid, block, 0, stream>>>(batch, count, p);
    else warp_lanczos3_kernel<false><<<grid, block, 0, stream>>>(batch, count, p);
  }
  check(hipGetLastError(), "warp kernel launch");
}

void launch_clip_integrate(const ClipArgs& a, int sampler, hipStream_t stream) {
  const unsigned grid256 = static_cast<unsigned>((a.npix + 255) / 256);
  const int lds_block = a.n <= 128 ? 128 : 64;
  const unsigned grid_lds = static_cast<unsigned>((a.npix + lds_block - 1) / lds_block);
  const size_t lds = static_cast<size_t>(a.n) * lds_block * sizeof(float);
#define GH_CLIP_LDS(S) clip_integrate_kernel<S, 0, 1><<<grid_lds, lds_block, lds, stream>>>(a)
#define GH_CLIP_AGAIN(S) clip_integrate_kernel<S, 0, 2><<<grid256, 256, 0, stream>>>(a)
  if (sampler == 0) {
    if (a.n <= 16) clip_integrate_kernel<0, 16, 0><<<grid256, 256, 0, stream>>>(a);
    else if (a.n <= 64) clip_integrate_kernel<0, 64, 0><<<grid256, 256, 0, stream>>>(a);
    else if (a.n <= 128) clip_integrate_kernel<0, 128, 0><<<grid256, 256, 0, stream>>>(a);
    else GH_CLIP_AGAIN(0);
  } else if (sampler == 2 && a.affine && !a.tile_map && static_cast<long long>(a.width) * a.height * 4 < (1ll << 31)) {
    // affine Lanczos3: staged source boxes per output tile
    const int tiles_x = (a.width + kSOutW - 1) / kSOutW, tiles_y = (a.height + kClipTileH - 1) / kClipTileH;
    const unsigned grid = static_cast<unsigned>(tiles_x * tiles_y);
    const bool clamp = a.clamp >= 0.0f;
    if (a.mode == 0) {
      if (clamp) clip_lanczos_tile_kernel<0, true><<<grid, kSThreads, 0, stream>>>(a, tiles_x);
      else clip_lanczos_tile_kernel<0, false><<<grid, kSThreads, 0, stream>>>(a, tiles_x);
    } else {
      if (clamp) clip_lanczos_tile_kernel<1, true><<<grid, kSThreads, 0, stream>>>(a, tiles_x);
      else clip_lanczos_tile_kernel<1, false><<<grid, kSThreads, 0, stream>>>(a, tiles_x);
    }
  } else if (a.mode == 0) {
    // the weighted mean is a single pass: nothing to store
    if (sampler == 1) GH_CLIP_AGAIN(1);
    else GH_CLIP_AGAIN(2);
  } else if (sampler == 1) {
    // bilinear samples are cheap to take again (4 taps); past 128 frames the
    // 64-thread LDS blocks lose more to occupancy than re-sampling costs
    if (a.n <= 128) GH_CLIP_LDS(1);
    else GH_CLIP_AGAIN(1);
  } else {
    if (a.n <= 256) GH_CLIP_LDS(2);
    else GH_CLIP_AGAIN(2);
  }
#undef GH_CLIP_LDS
#undef GH_CLIP_AGAIN
  check(hipGetLastError(), "clip integrate launch");
}

__global__ __launch_bounds__(256) void tile_map_kernel(float* map, const int* extents, const float* multipliers,
                                                       int tiles, int width) {
  for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
    const int x0 = extents[4 * t], y0 = extents[4 * t + 1], x1 = extents[4 * t + 2], y1 = extents[4 * t + 3];
    const long long tw = x1 - x0, area = tw * (y1 - y0);
    const float m = multipliers[t];
    for (long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; i < area;
         i += static_cast<long long>(gridDim.x) * 256) {
      const long long r = i / tw;
      st_global(map, (y0 + r) * width + x0 + (i - r * tw), m);
    }
  }
}

void launch_tile_map(float* map, const int* extents, const float* multipliers, int tiles, long long max_area,
                     int width, long long npix, hipStream_t stream) {
  launch_fill(map, npix, 1.0f, stream);
  const dim3 grid(static_cast<unsigned>(std::min<long long>((max_area + 255) / 256, 1024)),
                  static_cast<unsigned>(std::min(tiles, 65535)));
  tile_map_kernel<<<grid, 256, 0, stream>>>(map, extents, multipliers, tiles, width);
  check(hipGetLastError(), "tile map launch");
}

void launch_clip_integrate_tiles(const ClipArgs& a, hipStream_t stream) {
  const unsigned grid256 = static_cast<unsigned>((a.npix + 255) / 256);
  if (a.n <= 16) clip_integrate_kernel<0, 16, 0, true><<<grid256, 256, 0, stream>>>(a);
  else if (a.n <= 64) clip_integrate_kernel<0, 64, 0, true><<<grid256, 256, 0, stream>>>(a);
  else if (a.n <= 128) clip_integrate_kernel<0, 128, 0, true><<<grid256, 256, 0, stream>>>(a);
  else clip_integrate_kernel<0, 0, 2, true><<<grid256, 256, 0, stream>>>(a);
  check(hipGetLastError(), "tile-local integrate launch");
}

void launch_translate(const float* src, float* dst, int width, int height, bool bilinear, int idx, int idy, float fdx,
                      float fdy, float fill, float* coverage, unsigned* flag, hipStream_t stream) {
  const long long n = static_cast<long long>(width) * height;
  const unsigned grid = static_cast<unsigned>(std::min<long long>((n + 255) / 256, 8192));
  if (bilinear)
    translate_kernel<true><<<grid, 256, 0, stream>>>(src, dst, width, height, idx, idy, fdx, fdy, fill, coverage, flag);
  else
    translate_kernel<false><<<grid, 256, 0, stream>>>(src, dst, width, height, idx, idy, fdx, fdy, fill, coverage,
                                                      flag);
  check(hipGetLastError(), "translation launch");
}

}  // namespace gh
