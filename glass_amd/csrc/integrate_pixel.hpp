// Device code of the register-sort winsorized integration, shared by the
// translation units that instantiate its kernels (integrate.hip and the
// integrate_caps_*.hip capacity sets): sorting networks, rank extraction,
// scans of the sorted ends, output maps and the per-pixel routine.  See
// integrate.hip for the design and the oracle it follows.
#pragma once

#include <utility>

#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"

namespace gh {
namespace {


// ---------------------------------------------------------- sorting network
// Knuth TAOCP vol. 3, 5.2.2 Algorithm M (Batcher's merge exchange), n inputs.
constexpr int net_size(int n) {
  if (n < 2) return 0;
  int t = 0;
  while ((1 << t) < n) ++t;
  int c = 0;
  for (int p = 1 << (t - 1); p > 0; p >>= 1) {
    int q = 1 << (t - 1), r = 0, d = p;
    while (true) {
      for (int i = 0; i < n - d; ++i)
        if ((i & p) == r) ++c;
      if (q == p) break;
      d = q - p;
      q >>= 1;
      r = p;
    }
  }
  return c;
}

constexpr int net_passes(int n) {
  if (n < 2) return 0;
  int t = 0;
  while ((1 << t) < n) ++t;
  return t * (t + 1) / 2;
}

// Pruning for waves whose valid counts all lie in (CAP - kFastWindow, CAP]:
// the integration reads the sorted keys only at the quartile ranks of such a
// count, at its largest valid sample and at the ends, where the scans of the
// clip bounds walk.  A comparator neither of whose outputs reaches one of
// those slots (through the comparators after it) can be skipped; the slots
// not read may then hold their keys in any order, but the multiset of all
// slots is unchanged.  Exact slots: the bottom kExactEnd, the top
// kFastWindow + kExactEnd (the +inf keys of the invalid samples sit above the
// valid ones) and the rank slots of every count in the window.
constexpr int kFastWindow = 8;
constexpr int kExactEnd = 16;
constexpr int kExactTop = kFastWindow + kExactEnd;

constexpr bool exact_slot(int cap, int k) {
  if (k < kExactEnd || k >= cap - kExactTop) return true;
  for (int c = cap - kFastWindow + 1; c <= cap; ++c) {
    if (c < 1) continue;
    const int last = c - 1;
    if (k == last) return true;
    for (int m = 1; m <= 3; ++m) {
      const int l = (last * m) >> 2;  // the ranks pick_ranks reads
      if (k == l || k == (l + 1 < last ? l + 1 : last)) return true;
    }
  }
  return false;
}

// Algorithm M's comparators in pass order (the comparators of one pass touch
// disjoint slots, so their order within the pass is free).  Within pass p the
// comparators [begin[p], mid[p]) feed an exact slot and [mid[p], begin[p+1])
// do not.
template <int N>
struct Network {
  static constexpr int kSize = net_size(N);
  static constexpr int kPasses = net_passes(N);
  short a[kSize > 0 ? kSize : 1];
  short b[kSize > 0 ? kSize : 1];
  short begin[kPasses + 1];
  short mid[kPasses > 0 ? kPasses : 1];
};

template <int N>
constexpr Network<N> make_network() {
  Network<N> net{};
  if (N < 2) return net;
  constexpr int kSize = Network<N>::kSize;
  short pa[kSize]{}, pb[kSize]{}, pass[kSize]{};
  int t = 0;
  while ((1 << t) < N) ++t;
  int c = 0, np = 0;
  for (int p = 1 << (t - 1); p > 0; p >>= 1) {
    int q = 1 << (t - 1), r = 0, d = p;
    while (true) {
      for (int i = 0; i < N - d; ++i) {
        if ((i & p) == r) {
          pa[c] = static_cast<short>(i);
          pb[c] = static_cast<short>(i + d);
          pass[c] = static_cast<short>(np);
          ++c;
        }
      }
      ++np;
      if (q == p) break;
      d = q - p;
      q >>= 1;
      r = p;
    }
  }
  // Walk back from the exact slots: a comparator is kept when one of its
  // outputs is needed, and then both of its inputs are.
  bool need[N]{}, keep[kSize]{};
  for (int k = 0; k < N; ++k) need[k] = exact_slot(N, k);
  for (int i = kSize - 1; i >= 0; --i) {
    keep[i] = need[pa[i]] || need[pb[i]];
    if (keep[i]) need[pa[i]] = need[pb[i]] = true;
  }
  int out = 0;
  for (int p = 0; p < np; ++p) {
    net.begin[p] = static_cast<short>(out);
    for (int pruned = 0; pruned < 2; ++pruned) {
      if (pruned) net.mid[p] = static_cast<short>(out);
      for (int i = 0; i < kSize; ++i) {
        if (pass[i] == p && keep[i] == (pruned == 0)) {
          net.a[out] = pa[i];
          net.b[out] = pb[i];
          ++out;
        }
      }
    }
  }
  net.begin[np] = static_cast<short>(out);
  return net;
}

template <int N>
inline constexpr Network<N> kNetwork = make_network<N>();

// The LDS-sort path (more than 256 frames): one block's pixel segments fill
// 64 KB (two blocks per CU); fpad <= 16384 keys leaves at least one pixel.
constexpr int kLdsSortBytes = 64 * 1024;
constexpr int kLdsSortMaxFrames = kLdsSortBytes / 4;

// Frame slots are global memory.  Pointers read from the frame table are
// generic; casting them to the global address space turns flat loads (which
// also occupy the LDS/scalar counter and serialise behind every scalar load)
// into global loads that can all be in flight at once.
typedef __attribute__((address_space(1))) const float global_float;
__device__ __forceinline__ float load_global(const float* base, long long i) {
  return ((const global_float*)base)[i];  // C-style cast = addrspacecast
}
// Uniform base + 32-bit byte offset: selects the SGPR-base addressing form, so
// the ~200 loads of a pixel share one VGPR offset instead of one VGPR pair each.
__device__ __forceinline__ float load_global_off(const float* base, unsigned byte_offset) {
  return *(const global_float*)((const global_char*)base + byte_offset);
}

// Lambdas of the compile-time loops must inline, or the sample array would be
// passed by address and demoted to scratch memory.
#define GH_ALWAYS_INLINE __attribute__((always_inline))

// Compile-time loop: f(std::integral_constant<int, I>) for I in [0, N).
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// The samples of one pixel live in VGPRs.  Every access in the network and
// the linear passes uses a compile-time index (so the array is promoted to
// registers); the only runtime index is a wave-uniform rank, resolved by a
// scalar branch to its 16-element chunk and a 16-way select inside it.
template <int CAP>
struct Samples {
  static constexpr int kChunks = (CAP + 15) / 16;
  float x[CAP];

  template <int I>
  __device__ __forceinline__ float get() const {
    return x[I];
  }
  template <int I>
  __device__ __forceinline__ void set(float value) {
    x[I] = value;
  }
  // Element `idx` for a wave-uniform idx: scalar branches to the chunk and
  // to the element, then one move.  The empty volatile asm in each branch
  // body keeps the branches (a side effect cannot be speculated), or the
  // optimiser if-converts all of them into CAP per-lane selects.
  __device__ __forceinline__ float at_uniform(int idx) const {
    const int chunk = idx >> 4;
    const int lane = idx & 15;
    float r = 0.0f;
    static_for<kChunks>([&](auto C) GH_ALWAYS_INLINE {
      constexpr int ci = decltype(C)::value;
      if (chunk == ci) {
        asm volatile("");
        static_for<16>([&](auto J) GH_ALWAYS_INLINE {
          constexpr int j = decltype(J)::value;
          if constexpr (ci * 16 + j < CAP) {
            if (lane == j) {
              asm volatile("");
              r = x[ci * 16 + j];
            }
          }
        });
      }
    });
    return r;
  }
};

template <int CAP, int I>
__device__ __forceinline__ void compare_exchange(Samples<CAP>& v) {
  constexpr int i = kNetwork<CAP>.a[I];
  constexpr int j = kNetwork<CAP>.b[I];
  const float x = v.template get<i>();
  const float y = v.template get<j>();
  // One asm block per comparator keeps min and max adjacent: if the scheduler
  // separates them, both inputs stay live in between and a 200-sample network
  // needs twice the registers.  Keys are never NaN (invalid samples are +inf).
  float lo, hi;
  asm volatile("v_min_f32 %0, %2, %3\n\tv_max_f32 %1, %2, %3" : "=&v"(lo), "=v"(hi) : "v"(x), "v"(y));
  v.template set<i>(lo);
  v.template set<j>(hi);
}

template <int CAP, int FIRST, int... I>
__device__ __forceinline__ void run_comparators(Samples<CAP>& v, std::integer_sequence<int, I...>) {
  (compare_exchange<CAP, FIRST + I>(v), ...);
}

// The network pass by pass; `fast` (wave-uniform: every lane's valid count in
// (CAP - kFastWindow, CAP]) skips each pass's comparators that feed no exact
// slot, behind one scalar branch per pass.
template <int CAP>
__device__ __forceinline__ void sort_registers(Samples<CAP>& v, bool fast) {
  static_for<Network<CAP>::kPasses>([&](auto P) GH_ALWAYS_INLINE {
    constexpr int pi = decltype(P)::value;
    constexpr int b = kNetwork<CAP>.begin[pi], m = kNetwork<CAP>.mid[pi], e = kNetwork<CAP>.begin[pi + 1];
    run_comparators<CAP, b>(v, std::make_integer_sequence<int, m - b>{});
    if constexpr (e > m) {
      if (!fast) run_comparators<CAP, m>(v, std::make_integer_sequence<int, e - m>{});
    }
  });
}

// ------------------------------------------------------ scans of the ends
// The keys are sorted, so the samples beyond a per-lane bound sit at one end
// of the array.  These walk the slots from the bottom (or the top) in groups
// of kScanGroup and stop after the first group in which no lane of the wave
// meets `hit`: the cost follows the wave's largest count of such samples,
// not the capacity.  The group test is a wave-uniform branch; the empty
// volatile asm keeps it one (a side effect cannot be speculated, so the chain
// is not if-converted into selects over every slot).
constexpr int kScanGroup = 4;

template <int CAP, bool FROM_TOP, typename Hit, typename Body>
__device__ __forceinline__ void scan_ends(const Samples<CAP>& v, bool fast, Hit&& hit, Body&& body) {
  constexpr int kGroups = (CAP + kScanGroup - 1) / kScanGroup;
  // After the pruned network (`fast`) only the end's exact slots are in
  // order: a group without hits there still ends the scan, but once the scan
  // has walked past them (some lane has more samples beyond its bound than
  // the exact end holds) it visits every remaining group.
  constexpr int kExactGroups = (FROM_TOP ? kExactTop : kExactEnd) / kScanGroup;
  bool go = true;
  static_for<kGroups>([&](auto G) GH_ALWAYS_INLINE {
    constexpr int g = decltype(G)::value;
    if (go) {
      asm volatile("");
      bool any = false;
      static_for<kScanGroup>([&](auto J) GH_ALWAYS_INLINE {
        constexpr int j = g * kScanGroup + decltype(J)::value;
        if constexpr (j < CAP) {
          constexpr int k = FROM_TOP ? CAP - 1 - j : j;
          const float x = v.template get<k>();
          const bool h = hit(x);
          any |= h;
          body(x, h);
        }
      });
      if (__ballot(any) == 0ull && !(fast && g >= kExactGroups)) go = false;
    }
  });
}

template <int CAP, typename Hit, typename Body>
__device__ __forceinline__ void scan_from_bottom(const Samples<CAP>& v, bool fast, Hit&& hit, Body&& body) {
  scan_ends<CAP, false>(v, fast, hit, body);
}

template <int CAP, typename Hit, typename Body>
__device__ __forceinline__ void scan_from_top(const Samples<CAP>& v, bool fast, Hit&& hit, Body&& body) {
  scan_ends<CAP, true>(v, fast, hit, body);
}

// ------------------------------------------------------- quartile extraction
struct Quartiles {
  float a25, b25, a50, b50, a75, b75;
  float t25, t50, t75;
  float mx;  // largest valid sample
};

// Linear-percentile ranks for `count` samples: position (count-1)*fraction is
// exact in f32 for count < 2^22, lower = floor, upper = min(lower+1, count-1)
// (rejection.py:306-310) -- the same ranks as the all-valid form
// (rejection.py:358-366) whenever the interpolation weight t is non-zero.
template <int CAP>
__device__ __forceinline__ void pick_ranks(int c, const Samples<CAP>& v, Quartiles& q) {
  const int last = c - 1;
  const int p25 = last, p50 = last * 2, p75 = last * 3;  // positions * 4
  const int l25 = p25 >> 2, l50 = p50 >> 2, l75 = p75 >> 2;
  q.a25 = v.at_uniform(l25);
  q.b25 = v.at_uniform(min(l25 + 1, last));
  q.t25 = static_cast<float>(p25 & 3) * 0.25f;
  q.a50 = v.at_uniform(l50);
  q.b50 = v.at_uniform(min(l50 + 1, last));
  q.t50 = static_cast<float>(p50 & 3) * 0.25f;
  q.a75 = v.at_uniform(l75);
  q.b75 = v.at_uniform(min(l75 + 1, last));
  q.t75 = static_cast<float>(p75 & 3) * 0.25f;
  q.mx = v.at_uniform(max(last, 0));
}

__device__ __forceinline__ float lerp_percentile(float a, float b, float t, bool all_valid) {
  if (t == 0.0f) return a;
  if (all_valid) return a + (b - a) * t;               // rejection.py:366
  return a * (1.0f - t) + b * t;                       // rejection.py:317-319
}

// np.nan_to_num(x, nan=0.0) for a float32 scalar.
__device__ __forceinline__ float nan_to_num(float x) {
  if (isnan(x)) return 0.0f;
  if (isinf(x)) return x > 0.0f ? 3.402823466e38f : -3.402823466e38f;
  return x;
}

__device__ __forceinline__ void store_count(void* base, long long p, int count_u16, float value) {
  typedef __attribute__((address_space(1))) unsigned short global_u16;
  if (count_u16) {
    reinterpret_cast<global_u16*>(reinterpret_cast<__UINTPTR_TYPE__>(base))[p] = static_cast<unsigned short>(value);
  } else {
    st_global(reinterpret_cast<float*>(base), p, value);
  }
}

// Resident DQ flags (resident_cuda.py:3110-3297, bits contracts.py:77-88).
__device__ __forceinline__ unsigned dq_flags(const IntegrateArgs& a, long long p, float master, float weight,
                                             float coverage, float low, float high) {
  constexpr unsigned kNoData = 1u, kWarpEdge = 1u << 6, kLow = 1u << 8, kHigh = 1u << 9;
  unsigned dq = 0u;
  bool invalid = !isfinite(master) || !isfinite(weight) || weight <= 0.0f;
  if (!(coverage > 0.5f)) {
    dq |= kWarpEdge;
    invalid = true;
  }
  if (a.geometric) {
    const float g = ld_global(a.geometric, p);
    if (!(isfinite(g) && g > 0.5f)) {
      dq |= kWarpEdge;
      invalid = true;
    } else if (a.dq_active_frames > 0 && g < static_cast<float>(a.dq_active_frames) - 0.5f) {
      dq |= kWarpEdge;
    }
  }
  if (invalid) dq |= kNoData;
  if (low > 0.0f) dq |= kLow;
  if (high > 0.0f) dq |= kHigh;
  return dq;
}

__device__ __forceinline__ void write_outputs(const IntegrateArgs& a, long long p, float master, float weight,
                                              float coverage, float low, float high) {
  if (a.master) st_global(a.master, p, master);
  if (a.weight) st_global(a.weight, p, weight);
  if (a.coverage) store_count(a.coverage, p, a.count_u16, coverage);
  if (a.low) store_count(a.low, p, a.count_u16, low);
  if (a.high) store_count(a.high, p, a.count_u16, high);
  typedef __attribute__((address_space(1))) unsigned global_dq;
  if (a.dq) ((global_dq*)a.dq)[p] = dq_flags(a, p, master, weight, coverage, low, high);
}

// Population standard deviation of the valid samples of pixel p, FP64 sums in
// frame order.  all_valid: unrounded mean (np.std, rejection.py:231);
// otherwise the mean is rounded to f32 first (rejection.py:323-349).
__device__ __attribute__((noinline)) float frame_axis_std(const IntegrateArgs& a, long long p, double n,
                                                          bool all_valid) {
  const unsigned off = static_cast<unsigned>(p) * 4u;
  double s = 0.0;
  for (int k = 0; k < a.n_active; ++k) {
    const float x = load_global_off(table_at(a.frame_ptr, k), off);
    s += static_cast<double>(isfinite(x) ? x : 0.0f);
  }
  double mean = s / n;
  if (!all_valid) mean = static_cast<double>(static_cast<float>(mean));
  double ss = 0.0;
  for (int k = 0; k < a.n_active; ++k) {
    const float x = load_global_off(table_at(a.frame_ptr, k), off);
    const double d = isfinite(x) ? static_cast<double>(x) - mean : 0.0;
    ss += d * d;
  }
  return static_cast<float>(sqrt(ss / n));
}

// ------------------------------------------------------------------ kernels
// Winsorized clip + weighted mean of pixel p (p >= npix: no output).
template <int CAP, bool UNIT>
__device__ __forceinline__ void winsorized_pixel(const IntegrateArgs& a, long long p) {
  const long long pp0 = p < a.npix ? p : 0;

  // Per-tile choice of the statistics arithmetic (CPUStackEngine tiles) or
  // one choice for the whole frame (weighted_integrate_stack).  Decided
  // before the samples are loaded: the division needs registers that the
  // sorted keys occupy later.
  bool all_valid = (a.variant_ptr ? *(const global_u32*)a.variant_ptr : static_cast<unsigned>(a.variant_all_valid)) != 0u;
  if (a.tile_flags) {
    const unsigned w = static_cast<unsigned>(a.width);
    const unsigned y = static_cast<unsigned>(pp0) / w;  // npix < 2^30
    const unsigned x = static_cast<unsigned>(pp0) - y * w;
    const unsigned t = static_cast<unsigned>(a.tile_size);
    all_valid = a.tile_flags[(y / t) * static_cast<unsigned>(a.tiles_x) + (x / t)] != 0;
  }
  asm volatile("" ::"v"(all_valid));

  Samples<CAP> v;
  int cnt = 0;
  double swx = 0.0, sw = 0.0;
  double su = 0.0;  // unweighted sum of the valid samples (weighted stacks)
  // Load pass: every capacity slot issues its load before any use, so all
  // loads of the pixel are in flight together.  The frame table is padded to
  // the capacity with the last active frame (re-reads hit L2), which keeps the
  // table index a compile-time constant (merged scalar loads).  Each group of
  // 16 table entries is addressed through a fresh opaque zero so the scalar
  // loads of the pointers stay next to their use: hoisted to the kernel entry,
  // CAP 64-bit pointers overflow the SGPRs and spill.
  const unsigned byte_offset = static_cast<unsigned>(pp0) * 4u;  // host guarantees npix < 2^30
  int zero = 0;
  static_for<CAP>([&](auto K) GH_ALWAYS_INLINE {
    constexpr int k = decltype(K)::value;
    if constexpr (k % 16 == 0) asm volatile("" : "+s"(zero));
    v.template set<k>(load_global_off(table_at(a.frame_ptr, k + zero), byte_offset));
  });
  // Invalid samples add +0.0, as the oracle's np.where(valid, x, 0.0) * 0.0 does.
  static_for<CAP>([&](auto K) GH_ALWAYS_INLINE {
    constexpr int k = decltype(K)::value;
    const int src = k;
    const float x = v.template get<k>();
    const bool ok = (k < a.n_active) && isfinite(x);
    cnt += ok ? 1 : 0;
    if constexpr (UNIT) {
      swx += static_cast<double>(ok ? x : 0.0f);
    } else {
      const double w = static_cast<double>(((const_f32*)a.frame_w)[src]);
      const double xd = static_cast<double>(ok ? x : 0.0f);
      swx += ok ? xd * w : 0.0;
      sw += ok ? w : 0.0;
      su += xd;
    }
    v.template set<k>(ok ? x : INFINITY);
  });
  // Force the load-pass reductions to complete here.  Otherwise the compiler
  // defers them (they are only needed at the end) and keeps every raw sample
  // alive next to its sort key: twice the registers.
  if constexpr (UNIT) {
    asm volatile("" ::"v"(swx), "v"(cnt));
  } else {
    asm volatile("" ::"v"(swx), "v"(sw), "v"(su), "v"(cnt));
  }
  // The load-pass sums, the valid count, the pixel index and the statistics
  // choice wait out the sort in LDS.  Held in VGPRs they are the registers
  // that the capacity-200 network lacks, and the allocator spills them to
  // scratch (16 B written per pixel, as HBM traffic).  The slot is
  // addressed by the lane's position in the block recomputed from mbcnt and a
  // scalar wave index, so threadIdx.x is not held across the sort either; the
  // memory clobber makes the compiler reload the values instead of keeping
  // the stored registers alive.
  __shared__ double park_sums[UNIT ? 1 : 3][kBlock];
  __shared__ unsigned park_pixel[kBlock];  // p < 2^31 (npix < 2^30), bit 31: all_valid
  __shared__ int park_count[kBlock];
  const int wave_base = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x)) & ~63;
  {
    const int slot = static_cast<int>(threadIdx.x);
    park_sums[0][slot] = swx;
    if constexpr (!UNIT) {
      park_sums[1][slot] = sw;
      park_sums[2][slot] = su;
    }
    park_pixel[slot] = static_cast<unsigned>(p) | (all_valid ? 0x80000000u : 0u);
    park_count[slot] = cnt;
  }
  // Pruned network when every lane's valid count is within kFastWindow of
  // the capacity (wave-uniform; counts below take the full network).
  const bool fast = __ballot(cnt <= CAP - kFastWindow) == 0ull;
  asm volatile("" ::: "memory");

  // No branch between the load pass and the network: a branch here lets the
  // compiler sink the key selects past it and keep raw values and masks alive
  // together.  Pixels without samples (cnt == 0) run the network on +inf keys
  // and are written as zeros at the end.
  sort_registers<CAP>(v, fast);
  asm volatile("" ::: "memory");
  const int slot = wave_base + static_cast<int>(__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)));
  cnt = park_count[slot];

  // Waterfall over the distinct valid counts of the wave; each pass fetches
  // the six rank registers with a wave-uniform register index.
  Quartiles q{};
  bool pending = true;
  while (true) {
    const unsigned long long m = __ballot(pending);
    if (m == 0ull) break;
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    const int c = __builtin_amdgcn_readfirstlane(__builtin_amdgcn_readlane(cnt, leader));
    if (pending && cnt == c) {
      // Re-read the count through readfirstlane inside the branch: otherwise
      // the optimiser substitutes the per-lane `cnt` for `c` (they are equal
      // here) and the rank index stops being provably wave-uniform.
      pick_ranks<CAP>(__builtin_amdgcn_readfirstlane(c), v, q);
      pending = false;
    }
  }

  swx = park_sums[0][slot];
  if constexpr (!UNIT) {
    sw = park_sums[1][slot];
    su = park_sums[2][slot];
  }
  const unsigned parked = park_pixel[slot];
  p = static_cast<long long>(parked & 0x7fffffffu);
  all_valid = (parked >> 31) != 0u;
  const bool in_range = p < a.npix;
  const long long pp = in_range ? p : 0;

  const float center0 = lerp_percentile(q.a50, q.b50, q.t50, all_valid);
  const float q25 = lerp_percentile(q.a25, q.b25, q.t25, all_valid);
  const float q75 = lerp_percentile(q.a75, q.b75, q.t75, all_valid);
  float scale0 = (q75 - q25) / 1.349f;
  const double n = static_cast<double>(cnt);
  const bool scale0_ok = all_valid ? (scale0 > 0.0f) : (scale0 > 0.0f && isfinite(scale0));
  if (!scale0_ok && cnt > 0) {
    // Population standard deviation of the valid samples
    // (all-valid: np.std FP64, rejection.py:231; valid: rejection.py:323-349).
    // Rare (degenerate IQR): re-read the samples in frame order from memory
    // rather than keep FP64 copies of the register array alive.
    scale0 = frame_axis_std(a, pp, n, all_valid);
  }
  const float lo_clip = center0 - a.low_sigma * scale0;
  const float hi_clip = center0 + a.high_sigma * scale0;

  // Winsorized mean and population variance in FP64.  np.clip(x, lo, hi)
  // changes only the samples below lo_clip or above hi_clip, which the sorted
  // keys hold at their ends: the winsorized sum is the sum of the valid
  // samples (load pass, frame order) with those few replaced by the bound.
  // The FP64 sums of f32 samples are exact, so equal to the oracle's frame
  // order, whenever the samples' exponents span less than 53 - 24 - log2(n)
  // bits (about 21 at 193 frames); beyond that they can differ in the last
  // FP64 bits, which moves a float32 result only at a rounding boundary.
  double sum_lo = 0.0, sum_hi = 0.0;
  int c_lo = 0, c_hi = 0;
  scan_from_bottom<CAP>(v, fast, [&](float x) { return x < lo_clip; }, [&](float x, bool h) {
    c_lo += h ? 1 : 0;
    sum_lo += static_cast<double>(h ? x : 0.0f);
  });
  // from the top: the +inf keys (slots past the valid count) come first
  scan_from_top<CAP>(v, fast, [&](float x) { return x > hi_clip; }, [&](float x, bool h) {
    const bool real = h && x != INFINITY;
    c_hi += real ? 1 : 0;
    sum_hi += static_cast<double>(real ? x : 0.0f);
  });
  const double pad = static_cast<double>(CAP - cnt);
  if constexpr (UNIT) {
    sw = static_cast<double>(cnt);
    su = swx;
  }
  double ws = ((su - sum_lo) - sum_hi) + static_cast<double>(c_lo) * static_cast<double>(lo_clip) +
              static_cast<double>(c_hi) * static_cast<double>(hi_clip);
  double wmean = ws / n;
  float center = static_cast<float>(wmean);
  if (!all_valid) wmean = static_cast<double>(center);
  // Opaque copies of the clip bounds: the clipped values are recomputed here
  // (one med3 each) instead of being kept live from the previous pass, which
  // would hold CAP doubles in registers and spill.
  float lo_clip2 = lo_clip, hi_clip2 = hi_clip;
  asm volatile("" : "+v"(lo_clip2), "+v"(hi_clip2));
  double wss = 0.0;
  static_for<CAP>([&](auto K) GH_ALWAYS_INLINE {
    constexpr int k = decltype(K)::value;
    const double d = static_cast<double>(__builtin_amdgcn_fmed3f(v.template get<k>(), lo_clip2, hi_clip2)) - wmean;
    wss = __builtin_fma(d, d, wss);
  });
  const double d_pad = static_cast<double>(hi_clip2) - wmean;
  wss -= pad * (d_pad * d_pad);
  wss = wss > 0.0 ? wss : 0.0;
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
  double rej_sum = 0.0;  // rejected samples (unit weights: survivors = all - rejected)
  const bool usable = !a.require_positive_scale || (isfinite(scale) && scale > 0.0f);
  // hi_t capped at FLT_MAX: the +inf keys past the valid count stay above it
  // (the scan walks through them) without being counted.
  const float hi_c = fminf(hi_t, 3.402823466e38f);
  scan_from_bottom<CAP>(v, fast, [&](float x) { return usable && x < lo_t; }, [&](float x, bool h) {
    n_low += h ? 1 : 0;
    if constexpr (UNIT) rej_sum += static_cast<double>(h ? x : 0.0f);
  });
  scan_from_top<CAP>(v, fast, [&](float x) { return usable && x > hi_c; }, [&](float x, bool h) {
    const bool real = h && x != INFINITY;
    n_high += real ? 1 : 0;
    if constexpr (UNIT) rej_sum += static_cast<double>(real ? x : 0.0f);
  });
  {
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
    // Survivors: every valid sample minus the rejected ones (exact whenever
    // the sums are, see the winsorized sum above).
    const int kept = cnt - n_low - n_high;
    master = static_cast<float>((swx - rej_sum) / static_cast<double>(kept));
    weight = static_cast<float>(kept);
    coverage = static_cast<float>(kept);
  } else {
    // Survivors re-summed in frame order with their weights.
    double s = 0.0, wsum = 0.0;
    for (int k = 0; k < a.n_active; ++k) {
      const float x = load_global(table_at(a.frame_ptr, k), pp);
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
  if (!in_range) return;
  if (cnt == 0) {
    write_outputs(a, p, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  write_outputs(a, p, master, weight, coverage, static_cast<float>(n_low), static_cast<float>(n_high));
}


// The register kernel for capacity CAP.  Without a job table (t.block_job
// null) block b covers pixels [256 b, 256 b + 256) of the job `a`; with one,
// block b belongs to job t.block_job[b], whose arguments are read from
// t.jobs and whose first block is t.first_block[job] (one launch over the
// pixels of many jobs).
template <int CAP, bool UNIT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(CAP > 208 ? 1 : 2, 2))) void
integrate_winsorized_kernel(IntegrateArgs a, IntegrateJobTable t) {
  long long block = blockIdx.x;
  if (t.block_job) {
    typedef __attribute__((address_space(4))) const unsigned short const_u16;
    typedef __attribute__((address_space(4))) const unsigned const_u32;
    const int j = ((const const_u16*)t.block_job)[blockIdx.x];
    a = load_uniform(t.jobs + j);
    block -= ((const const_u32*)t.first_block)[j];
  }
  winsorized_pixel<CAP, UNIT>(a, block * kBlock + threadIdx.x);
}

template <int CAP>
void launch_cap(const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid, hipStream_t stream) {
  if (a.unit_weights) {
    integrate_winsorized_kernel<CAP, true><<<grid, kBlock, 0, stream>>>(a, t);
  } else {
    integrate_winsorized_kernel<CAP, false><<<grid, kBlock, 0, stream>>>(a, t);
  }
}

}  // namespace
}  // namespace gh
