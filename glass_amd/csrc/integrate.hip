// Winsorized sigma-clip rejection + weighted-mean integration along the frame
// axis for gfx950 (one thread per output pixel).
//
// Oracle: src/glass/cpu/integration.py:36-118 (weighted_integrate_stack) with
// src/glass/engine/rejection.py:31-366 (center_and_scale); the reference's
// CPU-parity CUDA kernel is cpp/cuda/integration_kernels.cu:2292-2779.
//
// Design (MI355X):
//  * Frame-major slots, so the 64 lanes of a wave read 256 contiguous bytes of
//    each frame; the frame pointers and weights are wave-uniform scalar loads.
//  * The valid samples of a pixel (invalid ones become +inf, as the oracle's
//    np.sort(np.where(valid, x, inf))) are sorted in VGPRs by Batcher's
//    merge-exchange network, generated at compile time for the capacity CAP.
//    No per-thread local array ever reaches scratch memory.
//  * Quartile ranks depend only on the valid count; the count is almost always
//    wave-uniform, so the rank registers are fetched by a scalar binary search
//    over the count (a "waterfall" over the distinct counts in the wave).
//  * Means, variances and the final weighted mean accumulate in FP64 like the
//    oracle; the rejection-free weighted mean is accumulated in frame order
//    during the load pass, so those pixels match the oracle bit for bit.
//  * HBM bytes: 4 B per active frame-pixel read, 4-20 B per pixel written.

#include <utility>

#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

#include "integrate_pixel.hpp"

namespace gh {
namespace {


// ---------------------------------------------------- more than 256 frames
// Stacks with more positive-weight frames than the register kernels hold sort
// each pixel's samples in LDS.  A block of kBlock threads takes P = 2^log2_p
// consecutive pixels: it stages their samples with coalesced loads (P floats
// of one frame per segment), keyed as in the register kernels (invalid ->
// +inf), into P segments of fpad = next power of two >= n_active keys, and
// sorts every segment with a bitonic network run by T = kBlock / P threads.
// Segments are XOR-swizzled (key k of pixel p lives at p * fpad + (k ^ (p & 31)))
// so that both the staging writes and the sort's pair accesses spread over the
// LDS banks without padding.  The first thread of each pixel then runs the
// winsorized arithmetic sequentially over the sorted keys in the register
// kernels' order and precision.
__device__ __forceinline__ int swz(int pix, int k, int fpad) { return pix * fpad + (k ^ (pix & 31)); }

template <bool UNIT>
__global__ __launch_bounds__(kBlock) void integrate_winsorized_lds_kernel(IntegrateArgs a, int fpad, int log2_p) {
  extern __shared__ float keys[];
  const int P = 1 << log2_p;
  const int T = kBlock >> log2_p;
  const int n_act = a.n_active;
  const long long p0 = static_cast<long long>(blockIdx.x) * P;
  // Stage: element e -> pixel e % P, frame e / P (a wave reads contiguous
  // pixels of one or a few frames).
  for (int e = threadIdx.x; e < P * fpad; e += kBlock) {
    const int pix = e & (P - 1), k = e >> log2_p;
    const long long p = p0 + pix;
    float key = INFINITY;
    if (k < n_act && p < a.npix) {
      const float x = load_global(a.frame_ptr[k], p);  // lanes may differ in k
      key = isfinite(x) ? x : INFINITY;
    }
    keys[swz(pix, k, fpad)] = key;
  }
  __syncthreads();
  const int pix = threadIdx.x / T, lane = threadIdx.x - pix * T;
  const long long p = p0 + pix;
  const bool leader = lane == 0 && p < a.npix;
  // Frame-order sums of the rejection-free mean, before the keys are sorted
  // (the register kernels' load pass; invalid samples add +0.0).
  double swx = 0.0, sw = 0.0;
  int cnt = 0;
  if (leader) {
    for (int k = 0; k < n_act; ++k) {
      const float x = keys[swz(pix, k, fpad)];
      const bool ok = x < INFINITY;
      cnt += ok ? 1 : 0;
      if constexpr (UNIT) {
        swx += static_cast<double>(ok ? x : 0.0f);
      } else {
        const double w = static_cast<double>(ld_global(a.frame_w, k));
        swx += ok ? static_cast<double>(x) * w : 0.0;
        sw += ok ? w : 0.0;
      }
    }
    if constexpr (UNIT) sw = static_cast<double>(cnt);
  }
  // Bitonic sort of every segment (ascending).
  for (int size = 2; size <= fpad; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (int q = lane; q < (fpad >> 1); q += T) {
        const int i = ((q & ~(stride - 1)) << 1) | (q & (stride - 1));
        const int j = i + stride;
        const int si = swz(pix, i, fpad), sj = swz(pix, j, fpad);
        const float x = keys[si], y = keys[sj];
        const float lo = fminf(x, y), hi = fmaxf(x, y);
        const bool up = (i & size) == 0;
        keys[si] = up ? lo : hi;
        keys[sj] = up ? hi : lo;
      }
    }
  }
  __syncthreads();
  if (!leader) return;
  auto key = [&](int k) { return keys[swz(pix, k, fpad)]; };

  bool all_valid = (a.variant_ptr ? *(const global_u32*)a.variant_ptr : static_cast<unsigned>(a.variant_all_valid)) != 0u;
  if (a.tile_flags) {
    const int y = static_cast<int>(p / a.width);
    const int x = static_cast<int>(p - static_cast<long long>(y) * a.width);
    all_valid = a.tile_flags[(y / a.tile_size) * a.tiles_x + (x / a.tile_size)] != 0;
  }
  if (cnt == 0) {
    write_outputs(a, p, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  // Linear-percentile ranks (pick_ranks).
  const int last = cnt - 1;
  const int p25 = last, p50 = last * 2, p75 = last * 3;
  const int l25 = p25 >> 2, l50 = p50 >> 2, l75 = p75 >> 2;
  const float center0 = lerp_percentile(key(l50), key(min(l50 + 1, last)), static_cast<float>(p50 & 3) * 0.25f,
                                        all_valid);
  const float q25 = lerp_percentile(key(l25), key(min(l25 + 1, last)), static_cast<float>(p25 & 3) * 0.25f,
                                    all_valid);
  const float q75 = lerp_percentile(key(l75), key(min(l75 + 1, last)), static_cast<float>(p75 & 3) * 0.25f,
                                    all_valid);
  float scale0 = (q75 - q25) / 1.349f;
  const double n = static_cast<double>(cnt);
  const bool scale0_ok = all_valid ? (scale0 > 0.0f) : (scale0 > 0.0f && isfinite(scale0));
  if (!scale0_ok) scale0 = frame_axis_std(a, p, n, all_valid);
  const float lo_clip = center0 - a.low_sigma * scale0;
  const float hi_clip = center0 + a.high_sigma * scale0;
  double ws = 0.0;
  for (int k = 0; k < cnt; ++k) ws += static_cast<double>(__builtin_amdgcn_fmed3f(key(k), lo_clip, hi_clip));
  double wmean = ws / n;
  float center = static_cast<float>(wmean);
  if (!all_valid) wmean = static_cast<double>(center);
  double wss = 0.0;
  for (int k = 0; k < cnt; ++k) {
    const double d = static_cast<double>(__builtin_amdgcn_fmed3f(key(k), lo_clip, hi_clip)) - wmean;
    wss += d * d;
  }
  float scale = static_cast<float>(sqrt(wss / n));
  if (all_valid) {
    if (!(scale > 0.0f)) scale = scale0;
  } else {
    if (!(scale > 0.0f && isfinite(scale))) scale = scale0;
    center = nan_to_num(center);
    scale = nan_to_num(scale);
  }
  const float lo_t = center - a.low_sigma * scale;
  const float hi_t = center + a.high_sigma * scale;
  int n_low = 0, n_high = 0;
  const bool usable = !a.require_positive_scale || (isfinite(scale) && scale > 0.0f);
  if (usable && !(key(0) >= lo_t && key(last) <= hi_t)) {
    for (int k = 0; k < cnt; ++k) {
      const float x = key(k);
      n_low += (x < lo_t) ? 1 : 0;
      n_high += (x > hi_t) ? 1 : 0;
    }
    const int rejected = n_low + n_high;
    if (rejected > 0) {
      const int remaining = cnt - rejected;
      const float fraction = static_cast<float>(static_cast<double>(rejected) / n);
      if (!(remaining >= a.min_samples && fraction <= a.max_frac)) {
        n_low = 0;
        n_high = 0;
      }
    }
  }
  float master, weight, coverage;
  if (n_low + n_high == 0) {
    master = (sw > 0.0) ? static_cast<float>(swx / sw) : 0.0f;
    weight = static_cast<float>(sw);
    coverage = static_cast<float>(cnt);
  } else if constexpr (UNIT) {
    double s = 0.0;
    for (int k = 0; k < cnt; ++k) {
      const float x = key(k);
      s += static_cast<double>((x >= lo_t && x <= hi_t) ? x : 0.0f);
    }
    const int kept = cnt - n_low - n_high;
    master = static_cast<float>(s / static_cast<double>(kept));
    weight = static_cast<float>(kept);
    coverage = static_cast<float>(kept);
  } else {
    double s = 0.0, wsum = 0.0;
    for (int k = 0; k < n_act; ++k) {
      const float x = load_global(a.frame_ptr[k], p);
      if (isfinite(x) && x >= lo_t && x <= hi_t) {
        const double w = static_cast<double>(ld_global(a.frame_w, k));
        s += static_cast<double>(x) * w;
        wsum += w;
      }
    }
    master = (wsum > 0.0) ? static_cast<float>(s / wsum) : 0.0f;
    weight = static_cast<float>(wsum);
    coverage = static_cast<float>(cnt - n_low - n_high);
  }
  write_outputs(a, p, master, weight, coverage, static_cast<float>(n_low), static_cast<float>(n_high));
}

// Rejection-free weighted mean for any frame count (integrate_mean,
// cpp/cuda/integration_kernels.cu:1610-1656; oracle cpu/integration.py:106-111).
template <bool UNIT>
__global__ __launch_bounds__(kBlock) void integrate_mean_kernel(IntegrateArgs a) {
  const long long p = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= a.npix) return;
  double swx = 0.0, sw = 0.0;
  int cnt = 0;
  for (int k = 0; k < a.n_active; ++k) {
    const float x = load_global(table_at(a.frame_ptr, k), p);
    if (isfinite(x)) {
      ++cnt;
      if constexpr (UNIT) {
        swx += static_cast<double>(x);
        sw += 1.0;
      } else {
        const double w = static_cast<double>(ld_global(a.frame_w, k));
        swx += static_cast<double>(x) * w;
        sw += w;
      }
    }
  }
  const float master = (sw > 0.0) ? static_cast<float>(swx / sw) : 0.0f;
  write_outputs(a, p, master, static_cast<float>(sw), static_cast<float>(cnt), 0.0f, 0.0f);
}

// One wave per job: OR of the job's frame flags -> statistics variant
// (rejection.py:39-44: np.all(valid) with positive weights selects the
// all-valid arithmetic).
__global__ __launch_bounds__(64) void decide_variants_kernel(const VariantJob* __restrict__ jobs, int* out) {
  const VariantJob v = load_uniform(jobs + blockIdx.x);
  unsigned any = 0u;
  for (int k = threadIdx.x; k < v.n_flags; k += 64) any |= *(const global_u32*)v.flags[k];  // lanes differ
  const bool nonfinite = __any(any != 0u);
  if (threadIdx.x == 0) out[blockIdx.x] = (v.weights_positive && !nonfinite) ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void tile_all_finite_kernel(const float* const* frame_ptr, int n_active,
                                                                  int height, int width, int tile_size,
                                                                  unsigned char* flags) {
  // One block per tile; flags[t] = 1 when every sample of every frame is finite.
  const int tiles_x = (width + tile_size - 1) / tile_size;
  const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const int x0 = tx * tile_size, y0 = ty * tile_size;
  const int x1 = min(x0 + tile_size, width), y1 = min(y0 + tile_size, height);
  const int tw = x1 - x0;
  const long long n = static_cast<long long>(tw) * (y1 - y0);
  int bad = 0;
  for (int k = 0; k < n_active && !bad; ++k) {
    for (long long i = threadIdx.x; i < n; i += blockDim.x) {
      const int yy = y0 + static_cast<int>(i / tw), xx = x0 + static_cast<int>(i % tw);
      if (!isfinite(frame_ptr[k][static_cast<long long>(yy) * width + xx])) bad = 1;
    }
    bad = __syncthreads_or(bad);
  }
  if (threadIdx.x == 0) flags[blockIdx.x] = bad ? 0 : 1;
}


}  // namespace

int integrate_register_capacity() { return 256; }

int integrate_cap_for(int n_active) {
  for (int cap : kIntegrateCaps)
    if (n_active <= cap) return cap;
  throw ValueError("more frames than the register kernels hold");
}

void launch_integrate_cap(int cap, const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid,
                          hipStream_t stream) {
  if (!(launch_integrate_caps_small(cap, a, t, grid, stream) || launch_integrate_caps_mid(cap, a, t, grid, stream) ||
        launch_integrate_caps_upper(cap, a, t, grid, stream) || launch_integrate_caps_large(cap, a, t, grid, stream)))
    throw ValueError("no register kernel of that capacity");
  check(hipGetLastError(), "integrate kernel launch");
}
int integrate_max_frames() { return kLdsSortMaxFrames; }

void launch_decide_variants(const VariantJob* jobs, int n_jobs, int* out, hipStream_t stream) {
  if (n_jobs <= 0) return;
  decide_variants_kernel<<<n_jobs, 64, 0, stream>>>(jobs, out);
  check(hipGetLastError(), "decide variants launch");
}

void launch_integrate(const IntegrateArgs& a, hipStream_t stream) {
  if (a.rejection == 0) {
    const unsigned grid = grid_for(a.npix, kBlock);
    if (a.unit_weights) integrate_mean_kernel<true><<<grid, kBlock, 0, stream>>>(a);
    else integrate_mean_kernel<false><<<grid, kBlock, 0, stream>>>(a);
  } else if (a.n_active <= integrate_register_capacity()) {
    launch_integrate_cap(integrate_cap_for(a.n_active), a, IntegrateJobTable{nullptr, nullptr, nullptr},
                         grid_for(a.npix, kBlock), stream);
  } else if (a.n_active <= kLdsSortMaxFrames) {
    int fpad = 512;
    while (fpad < a.n_active) fpad <<= 1;
    int log2_p = 0;  // pixels per block: P * fpad keys fill kLdsSortBytes
    while ((static_cast<long long>(2) << log2_p) * fpad * 4 <= kLdsSortBytes && (2 << log2_p) <= 64) ++log2_p;
    const long long blocks = (a.npix + (1LL << log2_p) - 1) >> log2_p;
    const size_t lds = (static_cast<size_t>(fpad) << log2_p) * sizeof(float);
    if (a.unit_weights) {
      integrate_winsorized_lds_kernel<true><<<static_cast<unsigned>(blocks), kBlock, lds, stream>>>(a, fpad, log2_p);
    } else {
      integrate_winsorized_lds_kernel<false><<<static_cast<unsigned>(blocks), kBlock, lds, stream>>>(a, fpad, log2_p);
    }
  } else {
    throw ValueError("winsorized integration supports at most 16384 positive-weight frames");
  }
  check(hipGetLastError(), "integrate kernel launch");
}

void launch_tile_all_finite(const float* const* frame_ptr, int n_active, int height, int width, int tile_size,
                            unsigned char* flags, hipStream_t stream) {
  const int tiles = ((width + tile_size - 1) / tile_size) * ((height + tile_size - 1) / tile_size);
  tile_all_finite_kernel<<<tiles, kBlock, 0, stream>>>(frame_ptr, n_active, height, width, tile_size, flags);
  check(hipGetLastError(), "tile flag kernel launch");
}

}  // namespace gh
