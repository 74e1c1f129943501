// Triangle-asterism descriptors and similarity estimation on gfx950.
//
// Semantics follow cpp/cuda/registration_kernels.cu:746-1821 (float32 star
// geometry, FP64 fits, '<=' nearest searches keeping the last index, mutual
// match scoring, most-inliers / lowest-RMS / lowest-index best, FP64 refit
// with 256-wide tree sums).  The layout is this library's own:
//
// * descriptors: one wave per anchor.  The nb nearest stars come from nb
//   rounds of a wave-wide minimum over 64-bit (distance bits, index) keys
//   strictly above the previous round's key (ties resolve to the lower
//   index, as the kernel's stable insertion does); the lanes then emit the
//   anchor's C(nb, 3) slots.
// * candidates: a gate-and-fit pass appends only the candidates that pass
//   the descriptor / pair gates (wave-aggregated atomics), so the scorer never
//   idles on rejected seeds.
// * scoring: one wave per candidate, one lane per moving star; catalogues and
//   the candidate's transformed moving stars sit in LDS and the forward and
//   reverse nearest searches read them as wave-wide broadcasts.  The inliers'
//   squared distances are added in moving-star order (float32), lane by lane.
// * best + refit: one 256-thread block reduces the per-block bests under the
//   total order (inliers desc, RMS asc, index asc) and runs the refit with the
//   reference's tree order over blocks of 256 moving stars.

#include <cfloat>
#include <cmath>

#include "common.hpp"
#include "triangles.hpp"

namespace gh {
namespace {

constexpr int kWave = 64;
constexpr int kWaves = 4;
constexpr int kThreads = kWave * kWaves;
constexpr int kMaxNeighbors = 16;
constexpr int kMaxScoreBlocks = 2048;
constexpr size_t kMaxScoreLds = 64 * 1024;

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x) & (kWave - 1); }

// Cross-lane LDS hand-off inside one wave (no block barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ unsigned lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
}

// Appends the lanes with keep set to the candidate list: one atomic per wave.
// Every lane of the wave must reach this call.
__device__ __forceinline__ void append(bool keep, int index, float4 params, const CandidateList& out) {
  const unsigned long long mask = __ballot(keep);
  if (mask == 0ull) return;
  const int leader = __builtin_ctzll(mask);
  unsigned base = 0;
  if (lane_id() == leader) base = atomicAdd(out.count, static_cast<unsigned>(__popcll(mask)));
  base = __shfl(base, leader);
  if (keep) {
    const unsigned pos = base + lanes_below(mask);
    if (pos < out.capacity) {
      out.index[pos] = index;
      out.params[pos] = params;
    }
  }
}

__device__ __forceinline__ bool finite2(float a, float b) { return isfinite(a) && isfinite(b); }

// --------------------------------------------------------------------------
// descriptors

__device__ __forceinline__ int pick(const int (&nbr)[kMaxNeighbors], int i) {
  int v = nbr[0];
#pragma unroll
  for (int k = 1; k < kMaxNeighbors; ++k) v = i == k ? nbr[k] : v;
  return v;
}

// Lexicographic rank -> (i, j, k) of the 3-subsets of [0, nb).
__device__ __forceinline__ void unrank3(int c, int nb, int& i, int& j, int& k) {
  i = 0;
  for (;; ++i) {
    const int n = (nb - 1 - i) * (nb - 2 - i) / 2;
    if (c < n) break;
    c -= n;
  }
  j = i + 1;
  for (;; ++j) {
    const int n = nb - 1 - j;
    if (c < n) break;
    c -= n;
  }
  k = j + 1 + c;
}

__device__ __forceinline__ float side(const float* x, const float* y, int l, int r) {
  const float dx = x[l] - x[r];
  const float dy = y[l] - y[r];
  return sqrtf(dx * dx + dy * dy);
}

// registration_kernels.cu:853-926: gates, vertex order and descriptor of one
// triangle of stars (i0, i1, i2).
__device__ RawTriangle make_slot(const float* x, const float* y, int i0, int i1, int i2) {
  RawTriangle t;
  t.d0 = t.d1 = t.area = NAN;
  t.pad = 0.0f;
  t.first = t.second = t.third = -1;
  t.valid = 0;
  if (i0 < 0 || i1 < 0 || i2 < 0 || i0 == i1 || i1 == i2 || i0 == i2) return t;
  const float x0 = x[i0], y0 = y[i0], x1 = x[i1], y1 = y[i1], x2 = x[i2], y2 = y[i2];
  const float area = fabsf(((x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0)) * 0.5f);
  if (area < 2.0f) return t;
  const float d01 = side(x, y, i0, i1), d12 = side(x, y, i1, i2), d20 = side(x, y, i2, i0);
  float s = d01, m = d12, l = d20, tmp;
  if (s > m) tmp = s, s = m, m = tmp;
  if (m > l) tmp = m, m = l, l = tmp;
  if (s > m) tmp = s, s = m, m = tmp;
  if (s <= 1.0e-6f || m <= 1.0e-6f) return t;
  int first = i0, second = i1, third = i2;
  float tf = d20, ts = d12;
  if (d12 >= d01 && d12 >= d20) {
    first = i1, second = i2, third = i0, tf = d01, ts = d20;
  } else if (d20 >= d01 && d20 >= d12) {
    first = i2, second = i0, third = i1, tf = d12, ts = d01;
  }
  if (tf < ts) {
    const int sw = first;
    first = second;
    second = sw;
  }
  t.d0 = l / m;
  t.d1 = m / s;
  t.area = area;
  t.first = first;
  t.second = second;
  t.third = third;
  t.valid = 1;
  return t;
}

__global__ __launch_bounds__(kThreads) void triangle_descriptor_kernel(const TriCatalog* catalogs, RawTriangle* raw) {
  const TriCatalog cat = load_uniform(catalogs + blockIdx.y);
  const int wave = static_cast<int>(threadIdx.x) / kWave, lane = lane_id();
  for (int a = static_cast<int>(blockIdx.x) * kWaves + wave; a < cat.count;
       a += static_cast<int>(gridDim.x) * kWaves) {
    const float ax = cat.x[a], ay = cat.y[a];
    const bool anchor_ok = finite2(ax, ay);
    int nbr[kMaxNeighbors];
    unsigned long long prev = 0ull;  // keys are (distance bits, index) + 1
#pragma unroll
    for (int r = 0; r < kMaxNeighbors; ++r) {
      nbr[r] = -1;
      if (r >= cat.neighbors || !anchor_ok) continue;
      unsigned long long best = ~0ull;
      for (int c = lane; c < cat.count; c += kWave) {
        const float cx = cat.x[c], cy = cat.y[c];
        if (!finite2(cx, cy)) continue;
        const float dx = cx - ax, dy = cy - ay;
        const float d2 = dx * dx + dy * dy;
        if (!(d2 < FLT_MAX)) continue;
        const unsigned long long key =
            ((static_cast<unsigned long long>(__float_as_uint(d2)) << 32) | static_cast<unsigned>(c)) + 1ull;
        if (key > prev && key < best) best = key;
      }
      best = wave_min_u64(best);
      if (best != ~0ull) nbr[r] = static_cast<int>(static_cast<unsigned>(best - 1ull));
      prev = best;
    }
    const long long base = static_cast<long long>(cat.raw_offset) + static_cast<long long>(a) * cat.combos;
    for (int c = lane; c < cat.combos; c += kWave) {
      int i, j, k;
      unrank3(c, cat.neighbors, i, j, k);
      raw[base + c] = anchor_ok ? make_slot(cat.x, cat.y, pick(nbr, i), pick(nbr, j), pick(nbr, k))
                                : make_slot(cat.x, cat.y, -1, -1, -1);
    }
  }
}

// --------------------------------------------------------------------------
// candidates

// FP64 least-squares similarity of three moving stars onto three reference
// stars (registration_kernels.cu:1320-1355), float32 results.
__device__ bool three_point_fit(const float (&m)[6], const float (&r)[6], float4& p) {
  const double n = 3.0;
  const double smx = static_cast<double>(m[0]) + m[2] + m[4];
  const double smy = static_cast<double>(m[1]) + m[3] + m[5];
  const double srx = static_cast<double>(r[0]) + r[2] + r[4];
  const double sry = static_cast<double>(r[1]) + r[3] + r[5];
  const double mmx = smx / n, mmy = smy / n, mrx = srx / n, mry = sry / n;
  double power = static_cast<double>(m[0]) * m[0];
  double same = static_cast<double>(m[0]) * r[0];
  double cross = static_cast<double>(m[0]) * r[1];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const double mx = m[2 * v], my = m[2 * v + 1], rx = r[2 * v], ry = r[2 * v + 1];
    if (v) {
      power = power + mx * mx;
      same = same + mx * rx;
      cross = cross + mx * ry;
    }
    power = power + my * my;
    same = same + my * ry;
    cross = cross - my * rx;
  }
  const double den = power - n * (mmx * mmx + mmy * mmy);
  if (den <= 1.0e-20) return false;
  const double na = same - n * (mmx * mrx + mmy * mry);
  const double nb = cross - n * (mmx * mry - mmy * mrx);
  const float a = static_cast<float>(na / den);
  const float b = static_cast<float>(nb / den);
  const float tx = static_cast<float>(mrx - (static_cast<double>(a) * mmx - static_cast<double>(b) * mmy));
  const float ty = static_cast<float>(mry - (static_cast<double>(b) * mmx + static_cast<double>(a) * mmy));
  if (!isfinite(a) || !isfinite(b) || !isfinite(tx) || !isfinite(ty)) return false;
  p = make_float4(a, b, tx, ty);
  return true;
}

__device__ __forceinline__ bool load_triangle(const float* x, const float* y, int n, int i0, int i1, int i2,
                                              float (&out)[6]) {
  if (i0 < 0 || i0 >= n || i1 < 0 || i1 >= n || i2 < 0 || i2 >= n) return false;
  out[0] = x[i0], out[1] = y[i0], out[2] = x[i1], out[3] = y[i1], out[4] = x[i2], out[5] = y[i2];
  return finite2(out[0], out[1]) && finite2(out[2], out[3]) && finite2(out[4], out[5]);
}

// Thread per (reference descriptor, moving descriptor) pair; both vertex
// orders of the reference triangle (registration_kernels.cu:1258-1362).
__global__ __launch_bounds__(kThreads) void triangle_candidate_kernel(ScoreCatalogs cat, const float2* rdesc,
                                                                      const int* ridx, int rn, const float2* mdesc,
                                                                      const int* midx, int mn, float radius2,
                                                                      CandidateList out) {
  const long long g = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  const long long pairs = static_cast<long long>(rn) * mn;
  bool near = false;
  int rd = 0, md = 0;
  if (g < pairs) {
    rd = static_cast<int>(g / mn);
    md = static_cast<int>(g - static_cast<long long>(rd) * mn);
    const float2 a = rdesc[rd], b = mdesc[md];
    if (finite2(a.x, a.y) && finite2(b.x, b.y)) {
      const float dx = a.x - b.x, dy = a.y - b.y;
      near = !((dx * dx + dy * dy) > radius2);
    }
  }
  for (int o = 0; o < 2; ++o) {
    bool keep = false;
    float4 p = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
    if (near) {
      int r0 = ridx[rd * 3 + 0], r1 = ridx[rd * 3 + 1];
      const int r2 = ridx[rd * 3 + 2];
      if (o) {
        const int sw = r0;
        r0 = r1;
        r1 = sw;
      }
      float rt[6], mt[6];
      if (load_triangle(cat.ref_x, cat.ref_y, cat.ref_count, r0, r1, r2, rt) &&
          load_triangle(cat.mov_x, cat.mov_y, cat.mov_count, midx[md * 3 + 0], midx[md * 3 + 1], midx[md * 3 + 2],
                        mt))
        keep = three_point_fit(mt, rt, p);
    }
    append(keep, static_cast<int>(g * 2 + o), p, out);
  }
}

__device__ __forceinline__ void ordered_pair(int p, int n, int& first, int& second) {
  first = p / (n - 1);
  const int off = p - first * (n - 1);
  second = off >= first ? off + 1 : off;
}

// Thread per ordered (reference pair, moving pair) seed with the scale,
// rotation and prior gates (registration_kernels.cu:1060-1143).
__global__ __launch_bounds__(kThreads) void pair_candidate_kernel(ScoreCatalogs cat, CatalogGates gates, int total,
                                                                  CandidateList out) {
  const long long g = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  bool keep = false;
  float4 p = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
  if (g < total) {
    const int R = cat.ref_count, M = cat.mov_count;
    const int npm = M * (M - 1);
    const int gi = static_cast<int>(g);
    const int rp = gi / npm, mp = gi - rp * npm;
    int ra, rb, ma, mb;
    ordered_pair(rp, R, ra, rb);
    ordered_pair(mp, M, ma, mb);
    const float rax = cat.ref_x[ra], ray = cat.ref_y[ra], rbx = cat.ref_x[rb], rby = cat.ref_y[rb];
    const float max_ = cat.mov_x[ma], may = cat.mov_y[ma], mbx = cat.mov_x[mb], mby = cat.mov_y[mb];
    if (finite2(rax, ray) && finite2(rbx, rby) && finite2(max_, may) && finite2(mbx, mby)) {
      const float mvx = mbx - max_, mvy = mby - may, rvx = rbx - rax, rvy = rby - ray;
      const float md2 = mvx * mvx + mvy * mvy, rd2 = rvx * rvx + rvy * rvy;
      const float mpd2 = gates.min_pair_distance * gates.min_pair_distance;
      if (!(md2 <= mpd2 || rd2 <= mpd2)) {
        const float a = (mvx * rvx + mvy * rvy) / md2;
        const float b = (mvx * rvy - mvy * rvx) / md2;
        const float tx = rax - (a * max_ - b * may);
        const float ty = ray - (b * max_ + a * may);
        const float scale = sqrtf(a * a + b * b);
        const float rotation = atan2f(b, a);
        keep = !(scale < gates.min_scale || scale > gates.max_scale);
        if (gates.max_abs_rotation_rad >= 0.0f && fabsf(rotation) > gates.max_abs_rotation_rad) keep = false;
        if (gates.prior_radius_px >= 0.0f) {
          const float ex = tx - gates.prior_dx, ey = ty - gates.prior_dy;
          if ((ex * ex + ey * ey) > gates.prior_radius_px * gates.prior_radius_px) keep = false;
        }
        p = make_float4(a, b, tx, ty);
      }
    }
  }
  append(keep, static_cast<int>(g), p, out);
}

__global__ void reset_count_kernel(unsigned* count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *count = 0u;
}

// --------------------------------------------------------------------------
// scoring

// b if it is valid and beats a (or a is empty), else a.  Evaluated without
// short-circuit branches and merged field by field with selects: the
// branchy form `if (valid && (empty || better)) a = b;` in lane-divergent code
// came out of this compiler with the pos field not updated on one path.
__device__ __forceinline__ BlockBest pick_best(const BlockBest& a, const BlockBest& b) {
  const bool beats = (b.score > a.score) |
                     ((b.score == a.score) & ((b.rms < a.rms) | ((b.rms == a.rms) & (b.index < a.index))));
  const bool take = (b.index >= 0) & ((a.index < 0) | beats);
  BlockBest r;
  r.score = take ? b.score : a.score;
  r.rms = take ? b.rms : a.rms;
  r.index = take ? b.index : a.index;
  r.pos = take ? b.pos : a.pos;
  return r;
}

__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ float2 transform(float a, float b, float tx, float ty, float2 q) {
  return make_float2(a * q.x - b * q.y + tx, b * q.x + a * q.y + ty);
}

template <bool kLds>
__global__ __launch_bounds__(kThreads) void score_kernel(ScoreCatalogs cat, float tol2, CandidateList list,
                                                         BlockBest* block_best) {
  extern __shared__ float2 lds[];
  __shared__ BlockBest wave_best[kWaves];
  const unsigned n = min(*list.count, list.capacity);
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave), lane = lane_id();
  if (blockIdx.x * kWaves >= n) {
    if (threadIdx.x == 0) block_best[blockIdx.x] = BlockBest{-1, NAN, -1, -1};
    return;
  }
  const int Rf = cat.ref_fin_count, Mf = cat.mov_fin_count;
  const float2* ref = cat.ref_fin;
  const float2* mov = cat.mov_fin;
  float2* tr = nullptr;
  if (kLds) {
    float2* lref = lds;
    float2* lmov = lds + Rf;
    for (int i = threadIdx.x; i < Rf; i += kThreads) lref[i] = cat.ref_fin[i];
    for (int i = threadIdx.x; i < Mf; i += kThreads) lmov[i] = cat.mov_fin[i];
    __syncthreads();
    ref = lref;
    mov = lmov;
    tr = lds + Rf + Mf + wave * Mf;
  }
  BlockBest mine{-1, NAN, -1, -1};
  for (unsigned c = blockIdx.x * kWaves + wave; c < n; c += gridDim.x * kWaves) {
    const int index = __builtin_amdgcn_readfirstlane(list.index[c]);
    const float4 p = list.params[c];
    const float a = uniform(p.x), b = uniform(p.y), tx = uniform(p.z), ty = uniform(p.w);
    if (kLds) {
      wave_lds_sync();  // the previous candidate's reads of tr are done
      for (int m = lane; m < Mf; m += kWave) tr[m] = transform(a, b, tx, ty, mov[m]);
      wave_lds_sync();
    }
    int inliers = 0;
    float sum = 0.0f;
    for (int base = 0; base < Mf; base += kWave) {
      const int m = base + lane;
      const bool active = m < Mf;
      const float2 me = active ? (kLds ? tr[m] : transform(a, b, tx, ty, mov[m])) : make_float2(0.0f, 0.0f);
      float bd = tol2;
      int br = -1;
#pragma unroll 4
      for (int r = 0; r < Rf; ++r) {
        const float2 q = ref[r];
        const float dx = me.x - q.x, dy = me.y - q.y;
        const float d = dx * dx + dy * dy;
        if (d <= bd) {
          bd = d;
          br = r;
        }
      }
      const float2 target = ref[br < 0 ? 0 : br];
      float rd = tol2;
      int rm = -1;
#pragma unroll 4
      for (int o = 0; o < Mf; ++o) {
        const float2 q = kLds ? tr[o] : transform(a, b, tx, ty, mov[o]);
        const float dx = q.x - target.x, dy = q.y - target.y;
        const float d = dx * dx + dy * dy;
        if (d <= rd) {
          rd = d;
          rm = o;
        }
      }
      const bool inlier = active && br >= 0 && rm == m;
      unsigned long long mask = __ballot(inlier);
      inliers += __popcll(mask);
      while (mask) {  // moving-star order: ascending lanes
        const int j = __builtin_ctzll(mask);
        sum = sum + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bd), j));
        mask &= mask - 1ull;
      }
    }
    const float rms = inliers > 0 ? sqrtf(sum / static_cast<float>(inliers)) : NAN;
    if (lane == 0) {
      list.score[c] = inliers;
      list.rms[c] = rms;
    }
    mine = pick_best(mine, BlockBest{inliers, rms, inliers > 0 ? index : -1, static_cast<int>(c)});
  }
  if (lane == 0) wave_best[wave] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    BlockBest best = wave_best[0];
    for (int w = 1; w < kWaves; ++w) best = pick_best(best, wave_best[w]);
    block_best[blockIdx.x] = best;
  }
}

// --------------------------------------------------------------------------
// best candidate and refit

// Nearest finite reference star within tolerance ('<=', last index).
__device__ __forceinline__ int nearest_reference(const ScoreCatalogs& cat, float2 q, float tol2) {
  float bd = tol2;
  int br = -1;
  for (int r = 0; r < cat.ref_fin_count; ++r) {
    const float2 s = cat.ref_fin[r];
    const float dx = q.x - s.x, dy = q.y - s.y;
    const float d = dx * dx + dy * dy;
    if (d <= bd) {
      bd = d;
      br = r;
    }
  }
  return br;
}

constexpr int kRefitValues = 8;

// Sums over the caller's moving stars in blocks of 256, each block's values
// added in the stride-halving tree order, the block totals added in block
// order from 0.0 (registration_kernels.cu:1641-1658, :1691-1702).  values(m,
// v) fills v[kRefitValues] for moving star m (zeros when it contributes
// nothing).  Result valid in thread 0.
template <typename F>
__device__ void tree_totals(double (*sums)[kThreads], int moving_count, F values, double (&total)[kRefitValues]) {
  const int t = threadIdx.x;
  for (int k = 0; k < kRefitValues; ++k) total[k] = 0.0;
  const int blocks = moving_count > 0 ? (moving_count + kThreads - 1) / kThreads : 1;
  for (int blk = 0; blk < blocks; ++blk) {
    double v[kRefitValues];
    for (int k = 0; k < kRefitValues; ++k) v[k] = 0.0;
    const int m = blk * kThreads + t;
    if (m < moving_count) values(m, v);
    for (int k = 0; k < kRefitValues; ++k) sums[k][t] = v[k];
    __syncthreads();
    for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
      if (t < stride)
        for (int k = 0; k < kRefitValues; ++k) sums[k][t] = sums[k][t] + sums[k][t + stride];
      __syncthreads();
    }
    if (t == 0)
      for (int k = 0; k < kRefitValues; ++k) total[k] = total[k] + sums[k][0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void finalize_kernel(ScoreCatalogs cat, float tol2, CandidateList list,
                                                            const BlockBest* block_best, int blocks, int refit,
                                                            SimilarityOut* out) {
  __shared__ BlockBest red[kThreads];
  __shared__ double sums[kRefitValues][kThreads];
  __shared__ float refit_matrix[6];
  __shared__ float refit_scale_rotation[2];
  __shared__ int refit_state;
  const int t = threadIdx.x;
  BlockBest mine{-1, NAN, -1, -1};
  for (int i = t; i < blocks; i += kThreads) mine = pick_best(mine, block_best[i]);
  red[t] = mine;
  __syncthreads();
  for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
    if (t < stride) red[t] = pick_best(red[t], red[t + stride]);
    __syncthreads();
  }
  const BlockBest best = red[0];
  float4 seed = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
  if (best.index >= 0) seed = list.params[best.pos];
  if (t == 0) {
    SimilarityOut r{};
    for (int i = 0; i < 9; ++i) r.matrix[i] = (i % 4 == 0) ? 1.0f : 0.0f;
    r.scale = 1.0f;
    r.rotation = 0.0f;
    r.rms = NAN;
    r.inliers = best.index >= 0 ? best.score : 0;
    r.best_index = best.index;
    r.best_pos = best.pos;
    r.refit_inliers = 0;
    r.refit_status = -1;
    r.refit_rms = NAN;
    r.scored = min(*list.count, list.capacity);
    if (best.index >= 0) {
      const float a = seed.x, b = seed.y;
      r.matrix[0] = a, r.matrix[1] = -b, r.matrix[2] = seed.z;
      r.matrix[3] = b, r.matrix[4] = a, r.matrix[5] = seed.w;
      r.scale = sqrtf(a * a + b * b);
      r.rotation = atan2f(b, a);
      r.rms = best.rms;
    }
    *out = r;
  }
  if (!refit) return;

  // refit statistics over the seed's one-sided matches
  const bool seed_ok = best.index >= 0 && isfinite(seed.x) && isfinite(seed.y) && isfinite(seed.z) &&
                       isfinite(seed.w);
  double total[kRefitValues];
  tree_totals(sums, cat.mov_count,
              [&](int m, double (&v)[kRefitValues]) {
                if (!seed_ok) return;
                const float mx = cat.mov_x[m], my = cat.mov_y[m];
                if (!finite2(mx, my)) return;
                const int br = nearest_reference(cat, transform(seed.x, seed.y, seed.z, seed.w, make_float2(mx, my)),
                                                 tol2);
                if (br < 0) return;
                const double x = mx, y = my, rx = cat.ref_fin[br].x, ry = cat.ref_fin[br].y;
                v[0] = x, v[1] = y, v[2] = rx, v[3] = ry;
                v[4] = x * x + y * y;
                v[5] = x * rx + y * ry;
                v[6] = x * ry - y * rx;
                v[7] = 1.0;
              },
              total);
  if (t == 0) {
    const int count = static_cast<int>(total[7]);
    int status = 0;
    float mat[6] = {1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f};
    float scale = 1.0f, rotation = 0.0f;
    if (count < 2) {
      status = 1;
    } else {
      const double n = static_cast<double>(count);
      const double mmx = total[0] / n, mmy = total[1] / n, mrx = total[2] / n, mry = total[3] / n;
      const double den = total[4] - n * (mmx * mmx + mmy * mmy);
      if (den <= 1.0e-20) {
        status = 2;
      } else {
        const double a = (total[5] - n * (mmx * mrx + mmy * mry)) / den;
        const double b = (total[6] - n * (mmx * mry - mmy * mrx)) / den;
        const double tx = mrx - (a * mmx - b * mmy);
        const double ty = mry - (b * mmx + a * mmy);
        mat[0] = static_cast<float>(a), mat[1] = static_cast<float>(-b), mat[2] = static_cast<float>(tx);
        mat[3] = static_cast<float>(b), mat[4] = static_cast<float>(a), mat[5] = static_cast<float>(ty);
        scale = static_cast<float>(sqrt(a * a + b * b));
        rotation = static_cast<float>(atan2(b, a));
      }
    }
    for (int i = 0; i < 6; ++i) refit_matrix[i] = mat[i];
    refit_scale_rotation[0] = scale;
    refit_scale_rotation[1] = rotation;
    refit_state = status;
    out->refit_inliers = count;
    out->refit_status = status;
  }
  __syncthreads();
  const int status = refit_state;
  double res[kRefitValues];
  const double m00 = refit_matrix[0], m01 = refit_matrix[1], m02 = refit_matrix[2];
  const double m10 = refit_matrix[3], m11 = refit_matrix[4], m12 = refit_matrix[5];
  tree_totals(sums, cat.mov_count,
              [&](int m, double (&v)[kRefitValues]) {
                if (status != 0 || !seed_ok) return;
                const float mx = cat.mov_x[m], my = cat.mov_y[m];
                if (!finite2(mx, my)) return;
                const int br = nearest_reference(cat, transform(seed.x, seed.y, seed.z, seed.w, make_float2(mx, my)),
                                                 tol2);
                if (br < 0) return;
                const double x = mx, y = my;
                const double dx = (m00 * x + m01 * y + m02) - static_cast<double>(cat.ref_fin[br].x);
                const double dy = (m10 * x + m11 * y + m12) - static_cast<double>(cat.ref_fin[br].y);
                v[0] = dx * dx + dy * dy;
              },
              res);
  if (t == 0) {
    SimilarityOut r = *out;
    float refit_rms = NAN;
    if (status == 0 && r.refit_inliers > 0) refit_rms = static_cast<float>(sqrt(res[0] / r.refit_inliers));
    r.refit_rms = refit_rms;
    if (status == 0) {
      if (r.refit_inliers < r.inliers || !isfinite(refit_rms) || !isfinite(r.rms) || refit_rms > r.rms) {
        r.refit_status = 3;  // rejected: the seed's matrix stays
      } else {
        for (int i = 0; i < 6; ++i) r.matrix[i] = refit_matrix[i];
        r.scale = refit_scale_rotation[0];
        r.rotation = refit_scale_rotation[1];
        r.rms = refit_rms;
      }
    }
    *out = r;
  }
}

}  // namespace

void launch_reset_count(unsigned* count, hipStream_t stream) {
  hipLaunchKernelGGL(reset_count_kernel, dim3(1), dim3(64), 0, stream, count);
}

void launch_triangle_descriptors(const TriCatalog* catalogs, int n_catalogs, int max_count, RawTriangle* raw,
                                 hipStream_t stream) {
  if (n_catalogs <= 0 || max_count <= 0) return;
  const unsigned gx = std::min<unsigned>(grid_for(max_count, kWaves), 1024u);
  hipLaunchKernelGGL(triangle_descriptor_kernel, dim3(gx, static_cast<unsigned>(n_catalogs)), dim3(kThreads), 0,
                     stream, catalogs, raw);
}

void launch_triangle_candidates(const ScoreCatalogs& cat, const float2* ref_desc, const int* ref_idx, int ref_n,
                                const float2* mov_desc, const int* mov_idx, int mov_n, float radius,
                                const CandidateList& out, hipStream_t stream) {
  const long long pairs = static_cast<long long>(ref_n) * mov_n;
  if (pairs <= 0) return;
  hipLaunchKernelGGL(triangle_candidate_kernel, dim3(grid_for(pairs, kThreads)), dim3(kThreads), 0, stream, cat,
                     ref_desc, ref_idx, ref_n, mov_desc, mov_idx, mov_n, radius * radius, out);
}

void launch_pair_candidates(const ScoreCatalogs& cat, const CatalogGates& gates, long long candidate_count,
                            const CandidateList& out, hipStream_t stream) {
  if (candidate_count <= 0) return;
  hipLaunchKernelGGL(pair_candidate_kernel, dim3(grid_for(candidate_count, kThreads)), dim3(kThreads), 0, stream, cat,
                     gates, static_cast<int>(candidate_count), out);
}

int score_max_blocks() { return kMaxScoreBlocks; }

int launch_score_candidates(const ScoreCatalogs& cat, float tolerance, const CandidateList& list,
                            BlockBest* block_best, int max_blocks, hipStream_t stream) {
  if (cat.ref_fin_count <= 0 || cat.mov_fin_count <= 0) return 0;  // nothing can score
  const long long upper = std::max<long long>(1, std::min<long long>(list.capacity, 1ll << 31));
  const int blocks = static_cast<int>(std::min<long long>(std::min(max_blocks, kMaxScoreBlocks),
                                                          (upper + kWaves - 1) / kWaves));
  const size_t lds = (static_cast<size_t>(cat.ref_fin_count) + static_cast<size_t>(cat.mov_fin_count) * (1 + kWaves)) *
                     sizeof(float2);
  const float tol2 = tolerance * tolerance;
  if (lds <= kMaxScoreLds)
    hipLaunchKernelGGL(score_kernel<true>, dim3(blocks), dim3(kThreads), lds, stream, cat, tol2, list, block_best);
  else
    hipLaunchKernelGGL(score_kernel<false>, dim3(blocks), dim3(kThreads), 0, stream, cat, tol2, list, block_best);
  return blocks;
}

void launch_finalize(const ScoreCatalogs& cat, float tolerance, const CandidateList& list, const BlockBest* block_best,
                     int blocks, int refit, SimilarityOut* out, hipStream_t stream) {
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, stream, cat, tolerance * tolerance, list,
                     block_best, blocks, refit, out);
}

}  // namespace gh
