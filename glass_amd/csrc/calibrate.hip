// Light calibration and per-frame statistics kernels for gfx950.
//
// Calibration arithmetic follows the CPU oracle src/glass/cpu/calibration.py:51-76
// operation for operation in f32 (no contraction: the library is built with
// -ffp-contract=off):  x = raw; x -= dark*s; [x -= bias]; x /= max(flat, floor);
// [x += pedestal].  Raw FITS u16 decode as cpp/cuda/calibration_kernels.cu:133-137.
//
// HBM bytes per frame-pixel: 2 (u16 raw) or 4 (f32 raw) read + 4 written, plus
// the masters (12 B/px) once per batch of up to 16 frames: each thread holds 4
// consecutive pixels' masters in registers and walks every frame of the batch
// (8 B / 16 B raw loads, 16 B stores).  Optional fused FP64 moments replace a
// separate stats pass over the stack.

#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace gh {
namespace {

constexpr int kFmtF32 = 0;
constexpr int kFmtU16 = 1;
constexpr int kFmtBE = 2;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ float calibrate_one(float v, float bias, float dark, float flat, const CalibParams& p,
                                               float dark_scale) {
  if (p.has_dark) {
    v = v - dark * dark_scale;
    if (!p.dark_includes_bias && p.has_bias) v = v - bias;
  } else if (p.has_bias) {
    v = v - bias;
  }
  if (p.has_flat) {
    // np.maximum keeps NaN; a NaN flat must not turn into the floor.
    const float d = (flat < p.flat_floor) ? p.flat_floor : flat;
    v = v / d;
  }
  if (p.pedestal != 0.0f) v = v + p.pedestal;
  return v;
}

__device__ __forceinline__ float decode_be(unsigned short word) {
  // bytes b0 b1 in memory; the little-endian load gives b0 | b1<<8.
  const unsigned b0 = word & 0xffu;
  const unsigned b1 = (word >> 8) & 0xffu;
  return static_cast<float>(((b0 << 8) | b1) ^ 0x8000u);
}

template <int FMT>
__device__ __forceinline__ void load_raw4(const void* src, long long i, float (&v)[4]) {
  if constexpr (FMT == kFmtF32) {
    const float4 q = reinterpret_cast<const float4*>(src)[i];
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  } else {
    const uint2 q = reinterpret_cast<const uint2*>(src)[i];
    const unsigned short s0 = q.x & 0xffffu, s1 = q.x >> 16, s2 = q.y & 0xffffu, s3 = q.y >> 16;
    if constexpr (FMT == kFmtU16) {
      v[0] = static_cast<float>(s0); v[1] = static_cast<float>(s1);
      v[2] = static_cast<float>(s2); v[3] = static_cast<float>(s3);
    } else {
      v[0] = decode_be(s0); v[1] = decode_be(s1); v[2] = decode_be(s2); v[3] = decode_be(s3);
    }
  }
}

template <int FMT>
__device__ __forceinline__ float load_raw1(const void* src, long long i) {
  if constexpr (FMT == kFmtF32) {
    return reinterpret_cast<const float*>(src)[i];
  } else {
    const unsigned short s = reinterpret_cast<const unsigned short*>(src)[i];
    if constexpr (FMT == kFmtU16) return static_cast<float>(s);
    return decode_be(s);
  }
}

// Batched calibration: each thread owns kCalPx consecutive pixels, loads their
// bias/dark/flat once and calibrates them in every frame of the batch, so the
// masters cost 12 B per pixel per batch rather than per frame.  Raw loads are
// issued kCalAhead frames at a time ahead of the arithmetic; calibrated frames
// are written with plain stores (GH_CAL_NT=1 selects non-temporal ones, which
// measured slower on MI355X).  STATS adds the fused per-frame FP64 moments
// (partials[k * gridDim.x + blockIdx.x]); MB is the frames per launch (the
// statistics variant keeps 4 frames of FP64 sums in registers; the host splits
// larger batches).
#ifndef GH_CAL_PX
#define GH_CAL_PX 8
#endif
#ifndef GH_CAL_AHEAD
#define GH_CAL_AHEAD 8
#endif
#ifndef GH_CAL_NT
#define GH_CAL_NT 0
#endif
constexpr int kCalPx = GH_CAL_PX;
constexpr int kCalAhead = GH_CAL_AHEAD;
static_assert(kCalPx == 4 || kCalPx == 8, "4 or 8 pixels per thread");

template <int FMT>
__device__ __forceinline__ void load_raw_px(const void* src, long long g, float (&v)[kCalPx]) {
  if constexpr (kCalPx == 4) {
    load_raw4<FMT>(src, g, v);
  } else {
    float a[4], b[4];
    load_raw4<FMT>(src, 2 * g, a);
    load_raw4<FMT>(src, 2 * g + 1, b);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[j] = a[j];
      v[4 + j] = b[j];
    }
  }
}

__device__ __forceinline__ void load_master_px(const float* m, long long g, float (&v)[kCalPx]) {
#pragma unroll
  for (int h = 0; h < kCalPx / 4; ++h) {
    const float4 q = reinterpret_cast<const float4*>(m)[g * (kCalPx / 4) + h];
    v[4 * h + 0] = q.x; v[4 * h + 1] = q.y; v[4 * h + 2] = q.z; v[4 * h + 3] = q.w;
  }
}

__device__ __forceinline__ void store_px(float* dst, long long g, const float (&v)[kCalPx]) {
#pragma unroll
  for (int h = 0; h < kCalPx / 4; ++h) {
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    f32x4* o = reinterpret_cast<f32x4*>(dst) + g * (kCalPx / 4) + h;
    const f32x4 q = {v[4 * h + 0], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]};
    if constexpr (GH_CAL_NT) {
      __builtin_nontemporal_store(q, o);
    } else {
      *o = q;
    }
  }
}

template <int FMT, bool STATS, int MB = STATS ? 4 : kCalibBatch>
__global__ __launch_bounds__(kBlock) void calibrate_batch_kernel(CalibBatch batch, int count, CalibParams p,
                                                                 long long groups, StatPartial* partials) {
  constexpr int kAhead = MB < kCalAhead ? MB : kCalAhead;
  static_assert(MB % kAhead == 0 && MB <= kCalibBatch, "batch splits into load groups");
  double sum[MB], sum_sq[MB];
  unsigned fin[MB], nonfin[MB];
#pragma unroll
  for (int k = 0; k < MB; ++k) {
    sum[k] = sum_sq[k] = 0.0;
    fin[k] = nonfin[k] = 0u;
  }
  unsigned bad = 0;  // bit k: frame k produced a non-finite value
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long g = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; g < groups; g += stride) {
    float b[kCalPx] = {}, d[kCalPx] = {}, fl[kCalPx] = {};
    if (p.has_bias) load_master_px(p.bias, g, b);
    if (p.has_dark) load_master_px(p.dark, g, d);
    if (p.has_flat) load_master_px(p.flat, g, fl);
#pragma unroll
    for (int k0 = 0; k0 < MB; k0 += kAhead) {
      if (k0 >= count) break;
      // The frames' pointers and scales are indexed through an opaque zero so
      // their kernel-argument loads stay in this group: hoisted out of the
      // pixel loop, MB frames' pointers overflow the SGPRs and are re-read
      // from VGPR lanes on every iteration.
      int zero = 0;
      asm volatile("" : "+s"(zero));
      float v[kAhead][kCalPx];
#pragma unroll
      for (int a = 0; a < kAhead; ++a) {
        if (k0 + a < count) load_raw_px<FMT>(batch.src[k0 + a + zero], g, v[a]);
      }
#pragma unroll
      for (int a = 0; a < kAhead; ++a) {
        const int k = k0 + a;
        if (k >= count) break;
        const float dark_scale = batch.dark_scale[k + zero];
#pragma unroll
        for (int j = 0; j < kCalPx; ++j) {
          v[a][j] = calibrate_one(v[a][j], b[j], d[j], fl[j], p, dark_scale);
          if (isfinite(v[a][j])) {
            if constexpr (STATS) {
              const double x = static_cast<double>(v[a][j]);
              sum[k] += x;
              sum_sq[k] += x * x;
              ++fin[k];
            }
          } else {
            bad |= 1u << k;
            if constexpr (STATS) ++nonfin[k];
          }
        }
        store_px(batch.dst[k + zero], g, v[a]);
      }
    }
  }
  // Any non-finite output marks its frame (one atomic per wave that saw one).
#pragma unroll
  for (int k = 0; k < MB; ++k) {
    if (k < count && __any((bad >> k) & 1u) && (threadIdx.x & 63) == 0) atomicOr(batch.nonfinite[k], 1u);
  }
  if constexpr (STATS) {
    __shared__ StatPartial s_part[kBlock / 64][MB];
    const int wave = threadIdx.x / 64;
#pragma unroll
    for (int k = 0; k < MB; ++k) {
      if (k >= count) break;
      const double ws = wave_sum(sum[k]), wq = wave_sum(sum_sq[k]);
      const unsigned long long wf = wave_sum(static_cast<unsigned long long>(fin[k]));
      const unsigned long long wn = wave_sum(static_cast<unsigned long long>(nonfin[k]));
      if ((threadIdx.x & 63) == 0) s_part[wave][k] = StatPartial{ws, wq, wf, wn};
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < count) {
      const int k = threadIdx.x;
      StatPartial out{0.0, 0.0, 0, 0};
      for (int w = 0; w < kBlock / 64; ++w) {
        out.sum += s_part[w][k].sum;
        out.sum_sq += s_part[w][k].sum_sq;
        out.finite += s_part[w][k].finite;
        out.nonfinite += s_part[w][k].nonfinite;
      }
      partials[static_cast<long long>(k) * gridDim.x + blockIdx.x] = out;
    }
  }
}

// One frame per blockIdx.y; serves frames whose pixel count is not a multiple
// of the batched kernel's pixels per thread (VEC = 4 or 1).
template <int FMT, int VEC>
__global__ __launch_bounds__(kBlock) void calibrate_kernel(CalibBatch batch, CalibParams p, long long npix,
                                                           StatPartial* partials) {
  const int f = blockIdx.y;
  const void* src = batch.src[f];
  float* dst = batch.dst[f];
  const float s = batch.dark_scale[f];
  double sum = 0.0, sum_sq = 0.0;
  unsigned long long finite = 0, nonfinite = 0;
  const long long groups = npix / VEC;
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long g = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; g < groups; g += stride) {
    float v[VEC];
    float b[VEC] = {}, d[VEC] = {}, fl[VEC] = {};
    if constexpr (VEC == 4) {
      load_raw4<FMT>(src, g, v);
      if (p.has_bias) {
        const float4 q = reinterpret_cast<const float4*>(p.bias)[g];
        b[0] = q.x; b[1] = q.y; b[2] = q.z; b[3] = q.w;
      }
      if (p.has_dark) {
        const float4 q = reinterpret_cast<const float4*>(p.dark)[g];
        d[0] = q.x; d[1] = q.y; d[2] = q.z; d[3] = q.w;
      }
      if (p.has_flat) {
        const float4 q = reinterpret_cast<const float4*>(p.flat)[g];
        fl[0] = q.x; fl[1] = q.y; fl[2] = q.z; fl[3] = q.w;
      }
    } else {
      v[0] = load_raw1<FMT>(src, g);
      if (p.has_bias) b[0] = p.bias[g];
      if (p.has_dark) d[0] = p.dark[g];
      if (p.has_flat) fl[0] = p.flat[g];
    }
#pragma unroll
    for (int j = 0; j < VEC; ++j) {
      v[j] = calibrate_one(v[j], b[j], d[j], fl[j], p, s);
      if (isfinite(v[j])) {
        if (p.accumulate_stats) {
          const double x = static_cast<double>(v[j]);
          sum += x;
          sum_sq += x * x;
          ++finite;
        }
      } else {
        ++nonfinite;
      }
    }
    if constexpr (VEC == 4) {
      reinterpret_cast<float4*>(dst)[g] = make_float4(v[0], v[1], v[2], v[3]);
    } else {
      dst[g] = v[0];
    }
  }
  // Any non-finite output marks the frame (one atomic per wave that saw one).
  if (__any(nonfinite != 0) && (threadIdx.x & 63) == 0) atomicOr(batch.nonfinite[f], 1u);
  if (!p.accumulate_stats) return;
  __shared__ double s_sum[kBlock / 64], s_sq[kBlock / 64];
  __shared__ unsigned long long s_fin[kBlock / 64], s_nf[kBlock / 64];
  sum = wave_sum(sum);
  sum_sq = wave_sum(sum_sq);
  finite = wave_sum(finite);
  nonfinite = wave_sum(nonfinite);
  const int wave = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    s_sum[wave] = sum; s_sq[wave] = sum_sq; s_fin[wave] = finite; s_nf[wave] = nonfinite;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    StatPartial out{0.0, 0.0, 0, 0};
    for (int w = 0; w < kBlock / 64; ++w) {
      out.sum += s_sum[w]; out.sum_sq += s_sq[w]; out.finite += s_fin[w]; out.nonfinite += s_nf[w];
    }
    partials[static_cast<long long>(f) * gridDim.x + blockIdx.x] = out;
  }
}

__global__ void clear_flags_kernel(CalibBatch batch, int count) {
  const int i = threadIdx.x;
  if (i < count) *batch.nonfinite[i] = 0u;
}

__global__ __launch_bounds__(kBlock) void stats_reduce_kernel(const StatPartial* partials, unsigned blocks_per_row,
                                                               CalibBatch batch, StatPartial* out) {
  const int row = blockIdx.x;
  double sum = 0.0, sum_sq = 0.0;
  unsigned long long finite = 0, nonfinite = 0;
  for (unsigned b = threadIdx.x; b < blocks_per_row; b += kBlock) {
    const StatPartial q = partials[static_cast<long long>(row) * blocks_per_row + b];
    sum += q.sum; sum_sq += q.sum_sq; finite += q.finite; nonfinite += q.nonfinite;
  }
  __shared__ double s_sum[kBlock / 64], s_sq[kBlock / 64];
  __shared__ unsigned long long s_fin[kBlock / 64], s_nf[kBlock / 64];
  sum = wave_sum(sum); sum_sq = wave_sum(sum_sq); finite = wave_sum(finite); nonfinite = wave_sum(nonfinite);
  const int wave = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    s_sum[wave] = sum; s_sq[wave] = sum_sq; s_fin[wave] = finite; s_nf[wave] = nonfinite;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    StatPartial r{0.0, 0.0, 0, 0};
    for (int w = 0; w < kBlock / 64; ++w) {
      r.sum += s_sum[w]; r.sum_sq += s_sq[w]; r.finite += s_fin[w]; r.nonfinite += s_nf[w];
    }
    out[batch.frame[row]] = r;
  }
}

// ------------------------------------------------------------ frame moments
// Block partial of a frame's FP64 moments over the grid-stride slice
// blockIdx.x of `blocks` blocks; the single-frame and the batched kernels share
// it, so a frame gets the same bits either way.
__device__ __forceinline__ void frame_moments_block(const float* frame, long long npix, unsigned blocks,
                                                    FrameMoments* out) {
  double sum = 0.0, sum_sq = 0.0;
  unsigned fin32 = 0, nonfin32 = 0;  // per thread at most npix / kBlock < 2^32
  float mn = INFINITY, mx = -INFINITY;
  const long long stride = static_cast<long long>(blocks) * kBlock;
  auto take = [&](float v) {
    if (isfinite(v)) {
      const double x = static_cast<double>(v);
      sum += x; sum_sq += x * x; ++fin32;
      mn = fminf(mn, v); mx = fmaxf(mx, v);
    } else {
      ++nonfin32;
    }
  };
  // Four loads in flight per thread; the values are taken in index order, as
  // the plain loop would.
  long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  for (; i + 3 * stride < npix; i += 4 * stride) {
    const float v0 = ld_global(frame, i), v1 = ld_global(frame, i + stride);
    const float v2 = ld_global(frame, i + 2 * stride), v3 = ld_global(frame, i + 3 * stride);
    take(v0); take(v1); take(v2); take(v3);
  }
  for (; i < npix; i += stride) take(ld_global(frame, i));
  unsigned long long finite = fin32, nonfinite = nonfin32;
  __shared__ FrameMoments s[kBlock / 64];
  sum = wave_sum(sum); sum_sq = wave_sum(sum_sq); finite = wave_sum(finite); nonfinite = wave_sum(nonfinite);
  mn = wave_min(mn); mx = wave_max(mx);
  const int wave = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) s[wave] = FrameMoments{sum, sum_sq, finite, nonfinite, mn, mx};
  __syncthreads();
  if (threadIdx.x == 0) {
    FrameMoments r{0.0, 0.0, 0, 0, INFINITY, -INFINITY};
    for (int w = 0; w < kBlock / 64; ++w) {
      r.sum += s[w].sum; r.sum_sq += s[w].sum_sq; r.finite += s[w].finite; r.nonfinite += s[w].nonfinite;
      r.min = fminf(r.min, s[w].min); r.max = fmaxf(r.max, s[w].max);
    }
    *out = r;
  }
}

__device__ __forceinline__ void frame_moments_reduce(const FrameMoments* partials, unsigned blocks,
                                                     FrameMoments* out) {
  double sum = 0.0, sum_sq = 0.0;
  unsigned long long finite = 0, nonfinite = 0;
  float mn = INFINITY, mx = -INFINITY;
  for (unsigned b = threadIdx.x; b < blocks; b += 64) {
    const FrameMoments q = partials[b];
    sum += q.sum; sum_sq += q.sum_sq; finite += q.finite; nonfinite += q.nonfinite;
    mn = fminf(mn, q.min); mx = fmaxf(mx, q.max);
  }
  sum = wave_sum(sum); sum_sq = wave_sum(sum_sq); finite = wave_sum(finite); nonfinite = wave_sum(nonfinite);
  mn = wave_min(mn); mx = wave_max(mx);
  if (threadIdx.x == 0) *out = FrameMoments{sum, sum_sq, finite, nonfinite, mn, mx};
}

__global__ __launch_bounds__(kBlock) void frame_moments_kernel(const float* frame, long long npix,
                                                                FrameMoments* partials) {
  frame_moments_block(frame, npix, gridDim.x, partials + blockIdx.x);
}

__global__ __launch_bounds__(64) void frame_moments_final(const FrameMoments* partials, unsigned blocks,
                                                          FrameMoments* out) {
  frame_moments_reduce(partials, blocks, out);
}

// Batched form: frame blockIdx.y of the table, partials [frame][blocks].
__global__ __launch_bounds__(kBlock) void frame_moments_batch_kernel(LnBatch b, long long npix,
                                                                      FrameMoments* partials) {
  int zero = 0;
  asm volatile("" : "+s"(zero));
  const float* frame = b.frame[blockIdx.y + zero];
  frame_moments_block(frame, npix, gridDim.x,
                      partials + static_cast<long long>(blockIdx.y) * gridDim.x + blockIdx.x);
}

__global__ __launch_bounds__(64) void frame_moments_batch_final(const FrameMoments* partials, unsigned blocks,
                                                                FrameMoments* out) {
  frame_moments_reduce(partials + static_cast<long long>(blockIdx.x) * blocks, blocks, out + blockIdx.x);
}

// ------------------------------------------------------------ radix select
// Order-preserving key of a finite float: negative values flip all bits,
// non-negative values set the sign bit.
__device__ __forceinline__ unsigned sort_key(float v) {
  const unsigned b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// state[0] = prefix, state[1] = resolved-bit mask, state[2..3] = remaining rank (u64)
__global__ __launch_bounds__(kBlock) void radix_hist_kernel(const float* frame, long long npix, int abs_mode,
                                                             float center, int shift, const unsigned* state,
                                                             unsigned* hist) {
  __shared__ unsigned local[256];
  local[threadIdx.x] = 0u;
  __syncthreads();
  const unsigned prefix = state[0];
  const unsigned mask = state[1];
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < npix; i += stride) {
    float v = frame[i];
    if (!isfinite(v)) continue;
    if (abs_mode) v = fabsf(v - center);
    const unsigned k = sort_key(v);
    if (((k ^ prefix) & mask) == 0u) atomicAdd(&local[(k >> shift) & 255u], 1u);
  }
  __syncthreads();
  const unsigned c = local[threadIdx.x];
  if (c) atomicAdd(&hist[threadIdx.x], c);
}

__global__ __launch_bounds__(256) void radix_scan_kernel(unsigned* hist, int shift, unsigned* state) {
  if (threadIdx.x != 0) return;
  unsigned long long rank = static_cast<unsigned long long>(state[2]) |
                            (static_cast<unsigned long long>(state[3]) << 32);
  unsigned long long cum = 0;
  int bucket = 255;
  for (int b = 0; b < 256; ++b) {
    const unsigned long long c = hist[b];
    if (rank < cum + c) {
      bucket = b;
      break;
    }
    cum += c;
  }
  rank -= cum;
  state[0] |= static_cast<unsigned>(bucket) << shift;
  state[1] |= 255u << shift;
  state[2] = static_cast<unsigned>(rank & 0xffffffffu);
  state[3] = static_cast<unsigned>(rank >> 32);
  for (int b = 0; b < 256; ++b) hist[b] = 0u;
}

__global__ void radix_init_kernel(unsigned* state, unsigned long long rank) {
  state[0] = 0u;
  state[1] = 0u;
  state[2] = static_cast<unsigned>(rank & 0xffffffffu);
  state[3] = static_cast<unsigned>(rank >> 32);
}

// --------------------------------------------------------------- utilities
__global__ void fill_kernel(float* dst, long long n, float value) {
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = value;
}

__global__ void add_scalar_kernel(float* dst, long long n, float value) {
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = dst[i] + value;
}

__global__ __launch_bounds__(kBlock) void normalize_kernel(float* frame, long long npix, float offset, float divisor,
                                                           unsigned* flag) {
  int bad = 0;
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < npix; i += stride) {
    const float v = (frame[i] - offset) / divisor;
    frame[i] = v;
    bad |= isfinite(v) ? 0 : 1;
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

__global__ __launch_bounds__(kBlock) void flag_nonfinite_kernel(const float* frame, long long npix, unsigned* flag) {
  int bad = 0;
  const long long stride = static_cast<long long>(gridDim.x) * kBlock;
  for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < npix; i += stride)
    bad |= isfinite(frame[i]) ? 0 : 1;
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

unsigned stream_blocks(long long n) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

}  // namespace

unsigned calibrate_blocks_per_frame(long long npix) {
  const int vec = (npix % 4 == 0) ? 4 : 1;
  long long groups = npix / vec;
  long long g = (groups + kBlock - 1) / kBlock;
  // Enough blocks per frame to fill the chip even for one frame, bounded so
  // each thread loops a few times and the partials stay small.
  if (g > 2048) g = 2048;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

void launch_calibrate(const CalibBatch& batch, int count, const CalibParams& p, long long npix,
                      StatPartial* partials, hipStream_t stream) {
  const bool vec4 = (npix % 4 == 0);
  const dim3 grid(calibrate_blocks_per_frame(npix), count);
#define GH_CAL(FMT, VEC) calibrate_kernel<FMT, VEC><<<grid, kBlock, 0, stream>>>(batch, p, npix, partials)
  // Without statistics partials the batched kernel takes one block per 256
  // pixel groups (up to 65,536 blocks): a u16 -> f32 stream measured 4.83 TB/s
  // with such grids against 4.5 TB/s with 2,048 grid-stride blocks.
  const unsigned grid_stream = static_cast<unsigned>(
      std::min<long long>(std::max<long long>((npix / kCalPx + kBlock - 1) / kBlock, 1), 65536));
#define GH_CALB(FMT, STATS) \
  calibrate_batch_kernel<FMT, STATS><<<grid_stream, kBlock, 0, stream>>>(batch, count, p, npix / kCalPx, partials)
#define GH_CALS(FMT)                                                                                      \
  calibrate_batch_kernel<FMT, true><<<grid.x, kBlock, 0, stream>>>(sub, n, p, npix / kCalPx,                   \
                                                                   partials + static_cast<long long>(k0) * grid.x)
  if (npix % kCalPx == 0) {
    if (p.accumulate_stats) {
      for (int k0 = 0; k0 < count; k0 += 4) {
        const int n = count - k0 < 4 ? count - k0 : 4;
        CalibBatch sub{};
        for (int j = 0; j < n; ++j) {
          sub.src[j] = batch.src[k0 + j];
          sub.dst[j] = batch.dst[k0 + j];
          sub.dark_scale[j] = batch.dark_scale[k0 + j];
          sub.nonfinite[j] = batch.nonfinite[k0 + j];
          sub.frame[j] = batch.frame[k0 + j];
        }
        if (p.format == kFmtF32) GH_CALS(kFmtF32);
        else if (p.format == kFmtU16) GH_CALS(kFmtU16);
        else GH_CALS(kFmtBE);
      }
    } else {
      if (p.format == kFmtF32) GH_CALB(kFmtF32, false);
      else if (p.format == kFmtU16) GH_CALB(kFmtU16, false);
      else GH_CALB(kFmtBE, false);
    }
  } else if (vec4) {
    if (p.format == kFmtF32) GH_CAL(kFmtF32, 4);
    else if (p.format == kFmtU16) GH_CAL(kFmtU16, 4);
    else GH_CAL(kFmtBE, 4);
  } else {
    if (p.format == kFmtF32) GH_CAL(kFmtF32, 1);
    else if (p.format == kFmtU16) GH_CAL(kFmtU16, 1);
    else GH_CAL(kFmtBE, 1);
  }
#undef GH_CAL
#undef GH_CALB
#undef GH_CALS
  check(hipGetLastError(), "calibrate kernel launch");
}

void launch_stats_reduce(const StatPartial* partials, const CalibBatch& batch, int rows, unsigned blocks_per_row,
                         StatPartial* out, hipStream_t stream) {
  stats_reduce_kernel<<<rows, kBlock, 0, stream>>>(partials, blocks_per_row, batch, out);
  check(hipGetLastError(), "stats reduce launch");
}

void launch_clear_flags(unsigned* /*flags*/, const CalibBatch& batch, int count, hipStream_t stream) {
  clear_flags_kernel<<<1, 64, 0, stream>>>(batch, count);
  check(hipGetLastError(), "clear flags launch");
}

unsigned frame_moments_blocks(long long npix) { return stream_blocks(npix); }

void launch_frame_moments_batch(const LnBatch& b, int n, long long npix, FrameMoments* scratch_partials,
                                unsigned blocks, FrameMoments* out, hipStream_t stream) {
  if (n <= 0) return;
  frame_moments_batch_kernel<<<dim3(blocks, static_cast<unsigned>(n)), kBlock, 0, stream>>>(b, npix,
                                                                                             scratch_partials);
  frame_moments_batch_final<<<static_cast<unsigned>(n), 64, 0, stream>>>(scratch_partials, blocks, out);
  check(hipGetLastError(), "frame moments batch launch");
}

void launch_frame_moments(const float* frame, long long npix, FrameMoments* scratch_partials, unsigned blocks,
                          FrameMoments* out, hipStream_t stream) {
  frame_moments_kernel<<<blocks, kBlock, 0, stream>>>(frame, npix, scratch_partials);
  frame_moments_final<<<1, 64, 0, stream>>>(scratch_partials, blocks, out);
  check(hipGetLastError(), "frame moments launch");
}

void launch_radix_select(const float* frame, long long npix, int abs_mode, float center,
                         unsigned long long rank, unsigned* hist, unsigned* state, hipStream_t stream) {
  zero_u32_kernel<<<1, 256, 0, stream>>>(hist, 256);
  radix_init_kernel<<<1, 1, 0, stream>>>(state, rank);
  const unsigned blocks = stream_blocks(npix);
  for (int shift = 24; shift >= 0; shift -= 8) {
    radix_hist_kernel<<<blocks, kBlock, 0, stream>>>(frame, npix, abs_mode, center, shift, state, hist);
    radix_scan_kernel<<<1, 256, 0, stream>>>(hist, shift, state);
  }
  check(hipGetLastError(), "radix select launch");
}

void launch_fill(float* dst, long long n, float value, hipStream_t stream) {
  fill_kernel<<<stream_blocks(n), kBlock, 0, stream>>>(dst, n, value);
  check(hipGetLastError(), "fill launch");
}

void launch_add_scalar(float* dst, long long n, float value, hipStream_t stream) {
  add_scalar_kernel<<<stream_blocks(n), kBlock, 0, stream>>>(dst, n, value);
  check(hipGetLastError(), "add scalar launch");
}

void launch_normalize(float* frame, long long npix, float offset, float divisor, unsigned* flag,
                      hipStream_t stream) {
  normalize_kernel<<<stream_blocks(npix), kBlock, 0, stream>>>(frame, npix, offset, divisor, flag);
  check(hipGetLastError(), "normalize launch");
}

void launch_flag_nonfinite(const float* frame, long long npix, unsigned* flag, hipStream_t stream) {
  flag_nonfinite_kernel<<<stream_blocks(npix), kBlock, 0, stream>>>(frame, npix, flag);
  check(hipGetLastError(), "flag nonfinite launch");
}

}  // namespace gh
