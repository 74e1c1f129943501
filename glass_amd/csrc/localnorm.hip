// Resident local normalisation between warp and integration (gfx950).
//
// Semantics of the reference's resident path (resident_cuda.py:12928-13100,
// coefficients resident_cuda.py:1897-1946, kernels
// cpp/cuda/local_norm_kernels.cu:6-71 and :220-304):
//  * statistics of a (source, reference) frame pair per tile of a grid, over
//    the pixels where both frames are finite: FP64 sums of s, s^2, r, r^2 and
//    the count (the host turns them into float32 means / standard deviations);
//  * application in place, x * scale[tile] + offset[tile] with the tile
//    coefficients constant over the tile (the last row / column of tiles
//    extends to the frame edge); the global model is a 1 x 1 grid.
// The product and the sum are rounded separately (the library builds with
// contraction off), as numpy does on the CPU path; the reference's CUDA kernel
// lets the compiler contract them into one fused multiply-add, which can differ
// in the last place.
//
// Both kernels are HBM-bound streams.  The statistics kernel reads each
// reference pixel once for kLnGroup sources (the reference re-reads it per
// source), and the groups of one tile are dispatched next to each other so
// their reference reads hit L2.

#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace gh {
namespace {

constexpr int kLnThreads = 256;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) f32x4 global_f32x4;
constexpr int kLnWaves = kLnThreads / 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

__device__ __forceinline__ unsigned wave_sum(unsigned v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Block b = tile * groups + g: sources g*kLnGroup .. of the batch against the
// reference over one tile (the groups of a tile are neighbours in dispatch
// order, so their reference reads share L2).  A 256-thread strided walk of
// the tile's rows (coalesced within a row), FP64 per-thread sums, then a
// fixed-order reduction (wave butterflies, then the four waves in order).
// out[(first + source) * tiles + tile]: sums[.. * 4 + {0,1,2,3}] = sum s, s^2,
// r, r^2; counts[..] = paired finite pixels.
__global__ __launch_bounds__(kLnThreads) void pair_grid_sums_kernel(LnBatch b, int n, const float* reference,
                                                                    int width, int height, int tile_w, int tile_h,
                                                                    int grid_cols, double* sums,
                                                                    unsigned long long* counts, int first) {
  __shared__ double part[kLnWaves][kLnGroup][4];
  __shared__ unsigned cpart[kLnWaves][kLnGroup];
  const int groups = (n + kLnGroup - 1) / kLnGroup;
  const int tile = static_cast<int>(blockIdx.x / static_cast<unsigned>(groups));
  const int g0 = static_cast<int>(blockIdx.x - static_cast<unsigned>(tile * groups)) * kLnGroup;
  const int tiles = static_cast<int>(gridDim.x / static_cast<unsigned>(groups));
  const int nk = min(kLnGroup, n - g0);
  int zero = 0;
  asm volatile("" : "+s"(zero));
  const float* src[kLnGroup];
#pragma unroll
  for (int j = 0; j < kLnGroup; ++j) src[j] = b.frame[min(g0 + j, n - 1) + zero];
  const int tx = tile % grid_cols, ty = tile / grid_cols;
  const int x0 = tx * tile_w, y0 = ty * tile_h;
  const int x1 = min(x0 + tile_w, width), y1 = min(y0 + tile_h, height);
  double acc[kLnGroup][4];
  unsigned cnt[kLnGroup];
#pragma unroll
  for (int j = 0; j < kLnGroup; ++j) {
    acc[j][0] = acc[j][1] = acc[j][2] = acc[j][3] = 0.0;
    cnt[j] = 0;
  }
  for (int y = y0; y < y1; ++y) {
    const unsigned row = static_cast<unsigned>(y) * static_cast<unsigned>(width);
    for (int x = x0 + static_cast<int>(threadIdx.x); x < x1; x += kLnThreads) {
      const unsigned i = row + static_cast<unsigned>(x);
      const float r = ld_global(reference, i);
      float s[kLnGroup];
#pragma unroll
      for (int j = 0; j < kLnGroup; ++j) s[j] = j < nk ? ld_global(src[j], i) : 0.0f;
      if (isfinite(r)) {
        const double rd = static_cast<double>(r);
        const double rd2 = rd * rd;
#pragma unroll
        for (int j = 0; j < kLnGroup; ++j) {
          if (isfinite(s[j])) {
            const double sd = static_cast<double>(s[j]);
            acc[j][0] += sd;
            acc[j][1] += sd * sd;
            acc[j][2] += rd;
            acc[j][3] += rd2;
            ++cnt[j];
          }
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kLnGroup; ++j) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double v = wave_sum(acc[j][k]);
      if (lane == 0) part[wave][j][k] = v;
    }
    const unsigned c = wave_sum(cnt[j]);
    if (lane == 0) cpart[wave][j] = c;
  }
  __syncthreads();
  // Thread t < nk * 5 finishes one (source, quantity) pair.
  const int t = threadIdx.x;
  if (t < nk * 5) {
    const int j = t / 5, k = t % 5;
    const long long o = static_cast<long long>(first + g0 + j) * tiles + tile;
    if (k < 4) {
      double v = part[0][j][k];
#pragma unroll
      for (int w = 1; w < kLnWaves; ++w) v += part[w][j][k];
      sums[o * 4 + k] = v;
    } else {
      unsigned long long c = 0;
#pragma unroll
      for (int w = 0; w < kLnWaves; ++w) c += cpart[w][j];
      counts[o] = c;
    }
  }
}

// Column tile of pixel x: x / tile_w through an FP64 reciprocal (exact to
// within one for any 31-bit x, then corrected), clamped to the last column.
__device__ __forceinline__ int column_tile(int x, int tile_w, double inv_tile_w, int last) {
  int c = static_cast<int>(static_cast<double>(x) * inv_tile_w);
  c -= (c * tile_w > x) ? 1 : 0;
  c += ((c + 1) * tile_w <= x) ? 1 : 0;
  return min(c, last);
}

// x * scale[tile] + offset[tile] in place for the frames of the batch
// (blockIdx.y); coefficient grids are [first + frame][grid_rows][grid_cols].
// Blocks stride over rows.  VEC: width and tile width are multiples of 4, so
// a float4 never straddles a tile and every row starts 16-byte aligned; each
// thread keeps kApplyUnroll float4 loads in flight.
constexpr int kApplyUnroll = 4;

template <bool VEC>
__global__ __launch_bounds__(kLnThreads) void apply_grid_kernel(LnBatch b, const float* scales,
                                                                const float* offsets, int width, int height,
                                                                int tile_w, int tile_h, double inv_tile_w,
                                                                int grid_cols, int grid_rows, int first) {
  int zero = 0;
  asm volatile("" : "+s"(zero));
  const int frame = blockIdx.y;
  float* dst = const_cast<float*>(b.frame[frame + zero]);
  unsigned* flag = b.flag[frame + zero];
  const long long grid = static_cast<long long>(grid_rows) * grid_cols;
  const float* sc = scales + (first + frame) * grid;
  const float* of = offsets + (first + frame) * grid;
  const int last = grid_cols - 1;
  bool bad = false;
  for (int y = blockIdx.x; y < height; y += gridDim.x) {
    const int trow = min(y / tile_h, grid_rows - 1) * grid_cols;
    const unsigned row = static_cast<unsigned>(y) * static_cast<unsigned>(width);
    if (VEC) {
      global_f32x4* p = (global_f32x4*)(dst + row);
      const int quads = width >> 2;
      for (int q0 = threadIdx.x; q0 < quads; q0 += kApplyUnroll * kLnThreads) {
        f32x4 v[kApplyUnroll];
#pragma unroll
        for (int u = 0; u < kApplyUnroll; ++u) {
          const int q = q0 + u * kLnThreads;
          if (q < quads) v[u] = p[q];
        }
#pragma unroll
        for (int u = 0; u < kApplyUnroll; ++u) {
          const int q = q0 + u * kLnThreads;
          if (q < quads) {
            const int c = trow + column_tile(q << 2, tile_w, inv_tile_w, last);
            const float s = ld_global(sc, c), o = ld_global(of, c);
            f32x4 r = v[u];
            r.x = r.x * s + o;
            r.y = r.y * s + o;
            r.z = r.z * s + o;
            r.w = r.w * s + o;
            p[q] = r;
            bad |= !(isfinite(r.x) && isfinite(r.y) && isfinite(r.z) && isfinite(r.w));
          }
        }
      }
    } else {
      for (int x = threadIdx.x; x < width; x += kLnThreads) {
        const int c = trow + column_tile(x, tile_w, inv_tile_w, last);
        const unsigned i = row + static_cast<unsigned>(x);
        const float v = ld_global(dst, i) * ld_global(sc, c) + ld_global(of, c);
        st_global_u32(dst, i, v);
        bad |= !isfinite(v);
      }
    }
  }
  // The frame's "holds a non-finite sample" flag stays exact.
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

}  // namespace

void launch_pair_grid_sums(const LnBatch& b, int n, const float* reference, int width, int height, int tile_w,
                           int tile_h, int grid_cols, int grid_rows, double* sums, unsigned long long* counts,
                           int first, hipStream_t stream) {
  if (n <= 0) return;
  const long long blocks = static_cast<long long>((n + kLnGroup - 1) / kLnGroup) * grid_cols * grid_rows;
  if (blocks > 0x7fffffffLL) throw ValueError("local normalization grid has too many tiles");
  pair_grid_sums_kernel<<<static_cast<unsigned>(blocks), kLnThreads, 0, stream>>>(b, n, reference, width, height, tile_w, tile_h, grid_cols,
                                                         sums, counts, first);
  check(hipGetLastError(), "pair grid statistics launch");
}

void launch_apply_grid(const LnBatch& b, int n, const float* scales, const float* offsets, int width, int height,
                       int tile_w, int tile_h, int grid_cols, int grid_rows, int first, hipStream_t stream) {
  if (n <= 0) return;
  const dim3 grid(static_cast<unsigned>(std::min(height, 1024)), static_cast<unsigned>(n));
  const double inv_tile_w = 1.0 / static_cast<double>(tile_w);
  if (width % 4 == 0 && tile_w % 4 == 0)
    apply_grid_kernel<true><<<grid, kLnThreads, 0, stream>>>(b, scales, offsets, width, height, tile_w, tile_h,
                                                              inv_tile_w, grid_cols, grid_rows, first);
  else
    apply_grid_kernel<false><<<grid, kLnThreads, 0, stream>>>(b, scales, offsets, width, height, tile_w, tile_h,
                                                               inv_tile_w, grid_cols, grid_rows, first);
  check(hipGetLastError(), "grid normalisation launch");
}

}  // namespace gh
