// C ABI of the triangle-asterism and star-pair similarity registration:
// triangle_asterism_descriptors_f32[_batch], estimate_similarity_from_
// triangle_descriptors_f32[_batch] and estimate_similarity_from_catalogs_f32
// (src/glass_cuda.py:1726-2140; cpp/src/native_bindings.cpp:17044-18678).
//
// Inputs are host arrays, as in the reference bindings.  Each call packs them
// into one upload, runs the kernels of triangles.hip on a per-device stream
// and synchronises once; the descriptor de-duplication / area order and the
// top-k candidate list are ordered on the host from the kernels' outputs.

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "stack.hpp"
#include "triangles.hpp"

using ghi::guarded;

namespace {

constexpr int kMaxDevices = 64;

struct Workspace {
  hipStream_t stream = nullptr;
  DeviceBuffer<unsigned char> inputs;
  DeviceBuffer<gh::RawTriangle> raw;
  DeviceBuffer<int> cand_index, cand_score;
  DeviceBuffer<float4> cand_params;
  DeviceBuffer<float> cand_rms;
  DeviceBuffer<unsigned> counts;
  DeviceBuffer<gh::BlockBest> block_best;
  DeviceBuffer<gh::SimilarityOut> results;
};

std::mutex g_mutex;
Workspace g_workspaces[kMaxDevices];

// The calling device's workspace, held for one call (calls serialise).
struct Session {
  std::unique_lock<std::mutex> lock;
  Workspace* ws = nullptr;
  Session() : lock(g_mutex) {
    int device = 0;
    check(hipGetDevice(&device), "hipGetDevice");
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("HIP device id beyond the workspace table");
    ws = &g_workspaces[device];
    if (!ws->stream)
      check(hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking), "hipStreamCreate(registration)");
  }
};

// Host staging of one upload: 16-byte aligned pieces of one buffer.
struct Packer {
  std::vector<unsigned char> host;
  size_t add(const void* src, size_t bytes) {
    const size_t off = (host.size() + 15) & ~static_cast<size_t>(15);
    host.resize(off + bytes);
    if (bytes) std::memcpy(host.data() + off, src, bytes);
    return off;
  }
  size_t add_zero(size_t bytes) {
    const size_t off = (host.size() + 15) & ~static_cast<size_t>(15);
    host.resize(off + bytes, 0);
    return off;
  }
  // Device base of the upload (valid until the next call's upload).
  unsigned char* device_base(Workspace& ws) const {
    ws.inputs.ensure(std::max<size_t>(host.size(), 16), "hipMalloc(registration inputs)");
    return ws.inputs.ptr;
  }
  // Copies the staging buffer into the workspace; returns the device base.
  unsigned char* upload(Workspace& ws) {
    device_base(ws);
    if (!host.empty())
      check(hipMemcpyAsync(ws.inputs.ptr, host.data(), host.size(), hipMemcpyHostToDevice, ws.stream),
            "hipMemcpy(registration inputs)");
    return ws.inputs.ptr;
  }
};

void require(bool ok, const char* message) {
  if (!ok) throw gh::ValueError(message);
}

// --------------------------------------------------------------------------
// descriptors

struct CatalogSetup {
  int count = 0, neighbors = 3, combos = 1;
  int64_t raw = 0;
};

// native_bindings.cpp:17062-17068.
CatalogSetup setup_catalog(int64_t n, int max_stars, int neighbors) {
  CatalogSetup s;
  const int64_t count = std::min<int64_t>(n, max_stars);
  s.count = static_cast<int>(count);
  int nb = static_cast<int>(std::min<int64_t>(count, neighbors));
  nb = std::max(3, nb);
  nb = std::min(16, nb);
  s.neighbors = nb;
  s.combos = nb * (nb - 1) * (nb - 2) / 6;
  s.raw = count >= 3 ? count * s.combos : 0;
  return s;
}

struct Triangle {
  std::array<int, 3> key;
  const gh::RawTriangle* slot;
};

// Host ordering of one catalogue's raw slots (native_bindings.cpp:17146-
// 17210): valid slots, one per star triple (the lowest slot of a triple is
// kept), then by area descending with the sorted triple as the tie key.
int64_t order_descriptors(const gh::RawTriangle* raw, int64_t n_raw, int max_descriptors, float* desc, int32_t* idx,
                          float* areas) {
  std::vector<Triangle> tri;
  tri.reserve(static_cast<size_t>(n_raw));
  for (int64_t s = 0; s < n_raw; ++s) {
    const gh::RawTriangle& r = raw[s];
    if (!r.valid || r.first < 0 || r.second < 0 || r.third < 0) continue;
    std::array<int, 3> key{r.first, r.second, r.third};
    std::sort(key.begin(), key.end());
    tri.push_back(Triangle{key, &r});
  }
  std::stable_sort(tri.begin(), tri.end(), [](const Triangle& a, const Triangle& b) { return a.key < b.key; });
  tri.erase(std::unique(tri.begin(), tri.end(), [](const Triangle& a, const Triangle& b) { return a.key == b.key; }),
            tri.end());
  std::sort(tri.begin(), tri.end(), [](const Triangle& a, const Triangle& b) {
    if (a.slot->area != b.slot->area) return a.slot->area > b.slot->area;
    return a.key < b.key;
  });
  const int64_t n = std::min<int64_t>(static_cast<int64_t>(tri.size()), max_descriptors);
  for (int64_t i = 0; i < n; ++i) {
    const gh::RawTriangle& r = *tri[static_cast<size_t>(i)].slot;
    desc[2 * i] = r.d0;
    desc[2 * i + 1] = r.d1;
    idx[3 * i] = r.first;
    idx[3 * i + 1] = r.second;
    idx[3 * i + 2] = r.third;
    areas[i] = r.area;
  }
  return n;
}

void descriptors_batch(const float* const* xs, const float* const* ys, const int64_t* ns, int64_t batch,
                       int32_t max_stars, int32_t neighbors, int32_t max_descriptors, float* desc, int32_t* idx,
                       float* areas, gh_triangle_info* info) {
  require(batch >= 0, "batch count must be non-negative");
  require(max_stars >= 0, "max_stars must be non-negative");
  require(max_descriptors >= 0, "max_descriptors must be non-negative");
  if (batch == 0) return;
  require(ns && info, "catalog sizes and info outputs are required");
  require(batch <= 65535, "at most 65535 catalogues per batch");
  std::vector<CatalogSetup> setup(static_cast<size_t>(batch));
  int64_t total_raw = 0;
  int max_count = 0;
  for (int64_t b = 0; b < batch; ++b) {
    require(ns[b] >= 0, "catalog sizes must be non-negative");
    setup[b] = setup_catalog(ns[b], max_stars, neighbors);
    info[b] = gh_triangle_info{0, setup[b].raw, setup[b].neighbors, 0};
    if (setup[b].raw > 0) {
      require(xs && ys && xs[b] && ys[b], "catalog coordinate arrays are required");
      total_raw += setup[b].raw;
      max_count = std::max(max_count, setup[b].count);
    }
  }
  if (total_raw == 0 || max_descriptors == 0) return;
  require(total_raw <= INT32_MAX, "raw descriptor slots exceed 2^31");
  require(desc && idx && areas, "descriptor outputs are required");

  Session session;
  Workspace& ws = *session.ws;
  Packer pack;
  std::vector<gh::TriCatalog> table;
  std::vector<std::pair<size_t, size_t>> xy_offsets;
  std::vector<int64_t> raw_offset(static_cast<size_t>(batch), 0);
  int64_t offset = 0;
  for (int64_t b = 0; b < batch; ++b) {
    if (setup[b].raw == 0) continue;
    const size_t ox = pack.add(xs[b], sizeof(float) * setup[b].count);
    const size_t oy = pack.add(ys[b], sizeof(float) * setup[b].count);
    xy_offsets.emplace_back(ox, oy);
    raw_offset[b] = offset;
    table.push_back(gh::TriCatalog{nullptr, nullptr, setup[b].count, setup[b].neighbors, setup[b].combos,
                                   static_cast<int>(offset)});
    offset += setup[b].raw;
  }
  const size_t table_offset = pack.add_zero(sizeof(gh::TriCatalog) * table.size());
  unsigned char* base = pack.device_base(ws);
  for (size_t k = 0; k < table.size(); ++k) {
    table[k].x = reinterpret_cast<const float*>(base + xy_offsets[k].first);
    table[k].y = reinterpret_cast<const float*>(base + xy_offsets[k].second);
  }
  std::memcpy(pack.host.data() + table_offset, table.data(), sizeof(gh::TriCatalog) * table.size());
  pack.upload(ws);
  ws.raw.ensure(static_cast<size_t>(total_raw), "hipMalloc(raw descriptor slots)");
  gh::launch_triangle_descriptors(reinterpret_cast<const gh::TriCatalog*>(base + table_offset),
                                  static_cast<int>(table.size()), max_count, ws.raw.ptr, ws.stream);
  check(hipGetLastError(), "triangle descriptor launch");
  std::vector<gh::RawTriangle> host(static_cast<size_t>(total_raw));
  check(hipMemcpyAsync(host.data(), ws.raw.ptr, host.size() * sizeof(gh::RawTriangle), hipMemcpyDeviceToHost,
                       ws.stream),
        "hipMemcpy(raw descriptor slots)");
  check(hipStreamSynchronize(ws.stream), "triangle descriptors");
  for (int64_t b = 0; b < batch; ++b) {
    if (setup[b].raw == 0) continue;
    const int64_t o = static_cast<int64_t>(b) * max_descriptors;
    info[b].count = order_descriptors(host.data() + raw_offset[b], setup[b].raw, max_descriptors, desc + 2 * o,
                                      idx + 3 * o, areas + o);
  }
}

// --------------------------------------------------------------------------
// similarity

// Offsets of one catalogue inside an upload: the caller's arrays and their
// finite entries (order kept).
struct CatalogUpload {
  size_t x = 0, y = 0, fin = 0;
  int count = 0, fin_count = 0;
};

CatalogUpload pack_catalog(Packer& pack, const float* x, const float* y, int64_t n) {
  CatalogUpload c;
  c.count = static_cast<int>(n);
  std::vector<float2> fin;
  fin.reserve(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i)
    if (std::isfinite(x[i]) && std::isfinite(y[i])) fin.push_back(make_float2(x[i], y[i]));
  c.fin_count = static_cast<int>(fin.size());
  c.x = pack.add(x, sizeof(float) * static_cast<size_t>(n));
  c.y = pack.add(y, sizeof(float) * static_cast<size_t>(n));
  c.fin = pack.add(fin.data(), sizeof(float2) * fin.size());
  return c;
}

gh::ScoreCatalogs score_catalogs(unsigned char* base, const CatalogUpload& ref, const CatalogUpload& mov) {
  gh::ScoreCatalogs s{};
  s.ref_fin = reinterpret_cast<const float2*>(base + ref.fin);
  s.ref_fin_count = ref.fin_count;
  s.mov_fin = reinterpret_cast<const float2*>(base + mov.fin);
  s.mov_fin_count = mov.fin_count;
  s.ref_x = reinterpret_cast<const float*>(base + ref.x);
  s.ref_y = reinterpret_cast<const float*>(base + ref.y);
  s.ref_count = ref.count;
  s.mov_x = reinterpret_cast<const float*>(base + mov.x);
  s.mov_y = reinterpret_cast<const float*>(base + mov.y);
  s.mov_count = mov.count;
  return s;
}

gh::CandidateList candidate_list(Workspace& ws, size_t capacity, unsigned* count) {
  ws.cand_index.ensure(capacity, "hipMalloc(similarity candidates)");
  ws.cand_params.ensure(capacity, "hipMalloc(similarity candidate parameters)");
  ws.cand_score.ensure(capacity, "hipMalloc(similarity candidate scores)");
  ws.cand_rms.ensure(capacity, "hipMalloc(similarity candidate rms)");
  gh::CandidateList l{};
  l.index = ws.cand_index.ptr;
  l.params = ws.cand_params.ptr;
  l.score = ws.cand_score.ptr;
  l.rms = ws.cand_rms.ptr;
  l.count = count;
  l.capacity = static_cast<unsigned>(capacity);
  return l;
}

void fill_fit(const gh::SimilarityOut& r, int64_t candidate_count, gh_similarity_fit* out) {
  std::memcpy(out->matrix, r.matrix, sizeof(out->matrix));
  out->scale = r.scale;
  out->rotation_rad = r.rotation;
  out->rms_px = r.rms;
  out->inliers = r.inliers;
  out->best_candidate_index = r.best_index;
  out->candidate_count = candidate_count;
  out->scored_candidates = r.scored;
  out->refined_inliers = r.refit_inliers;
  out->refit_status = r.refit_status;
  out->refit_rms_px = r.refit_rms;
}

// One estimate on the workspace stream: gates / fits, scoring, best (+ refit).
template <typename Gate>
void run_estimate(Workspace& ws, const gh::ScoreCatalogs& cat, float tolerance, const gh::CandidateList& list,
                  int refit, gh::SimilarityOut* result, Gate&& gate) {
  gh::launch_reset_count(list.count, ws.stream);
  gate();
  const int blocks = gh::launch_score_candidates(cat, tolerance, list, ws.block_best.ptr, gh::score_max_blocks(),
                                                 ws.stream);
  gh::launch_finalize(cat, tolerance, list, ws.block_best.ptr, blocks, refit, result, ws.stream);
  check(hipGetLastError(), "similarity kernels");
}

void check_descriptor_table(const float* desc, const int32_t* idx, int64_t n, const char* what) {
  require(n >= 0, "descriptor counts must be non-negative");
  if (n > 0) require(desc && idx, what);
}

struct TopEntry {
  int index, score;
  float rms;
  float4 params;
};

// The top-k list of native_bindings.cpp:18424-18530 over the scored
// candidates in candidate-index order.
std::vector<TopEntry> top_candidates(std::vector<TopEntry> all, int top_k, int reference_count, int moving_count) {
  std::sort(all.begin(), all.end(), [](const TopEntry& a, const TopEntry& b) { return a.index < b.index; });
  const int per_limit = std::max(1, (top_k + 3) / 4);
  const int buckets_n = std::max(reference_count, moving_count) + 1;
  std::vector<std::vector<int>> buckets(static_cast<size_t>(buckets_n));
  std::vector<int> overall;
  auto insert = [&](std::vector<int>& sel, int pos, int limit) {
    if (static_cast<int>(sel.size()) < limit) {
      sel.push_back(pos);
      return;
    }
    size_t worst = 0;
    for (size_t q = 1; q < sel.size(); ++q) {
      const TopEntry &c = all[sel[q]], &w = all[sel[worst]];
      if (c.score < w.score || (c.score == w.score && c.rms > w.rms)) worst = q;
    }
    const TopEntry &e = all[pos], &w = all[sel[worst]];
    if (e.score > w.score || (e.score == w.score && e.rms < w.rms)) sel[worst] = pos;
  };
  for (int pos = 0; pos < static_cast<int>(all.size()); ++pos) {
    const TopEntry& e = all[pos];
    if (e.score < 0 || !std::isfinite(e.rms)) continue;
    insert(overall, pos, top_k);
    if (e.score < buckets_n) insert(buckets[e.score], pos, per_limit);
  }
  std::vector<int> chosen;
  auto take = [&](int pos) {
    if (static_cast<int>(chosen.size()) < top_k && std::find(chosen.begin(), chosen.end(), pos) == chosen.end())
      chosen.push_back(pos);
  };
  for (int score = buckets_n - 1; score >= 0 && static_cast<int>(chosen.size()) < top_k; --score) {
    auto& bucket = buckets[score];
    std::sort(bucket.begin(), bucket.end(), [&](int a, int b) {
      if (all[a].rms != all[b].rms) return all[a].rms < all[b].rms;
      return all[a].index < all[b].index;
    });
    for (int pos : bucket) take(pos);
  }
  for (int pos : overall) take(pos);
  std::sort(chosen.begin(), chosen.end(), [&](int a, int b) {
    if (all[a].score != all[b].score) return all[a].score > all[b].score;
    if (all[a].rms != all[b].rms) return all[a].rms < all[b].rms;
    return all[a].index < all[b].index;
  });
  std::vector<TopEntry> out;
  for (int pos : chosen) out.push_back(all[pos]);
  return out;
}

}  // namespace

extern "C" {

int gh_triangle_descriptors(const float* x, const float* y, int64_t n, int32_t max_stars, int32_t neighbors,
                            int32_t max_descriptors, float* descriptors, int32_t* indices, float* areas,
                            gh_triangle_info* info) {
  return guarded([&] {
    require(n >= 0, "catalog size must be non-negative");
    const float* xs[1] = {x};
    const float* ys[1] = {y};
    const int64_t ns[1] = {n};
    descriptors_batch(xs, ys, ns, 1, max_stars, neighbors, max_descriptors, descriptors, indices, areas, info);
  });
}

int gh_triangle_descriptors_batch(const float* const* x, const float* const* y, const int64_t* n, int64_t batch,
                                  int32_t max_stars, int32_t neighbors, int32_t max_descriptors, float* descriptors,
                                  int32_t* indices, float* areas, gh_triangle_info* info) {
  return guarded([&] {
    descriptors_batch(x, y, n, batch, max_stars, neighbors, max_descriptors, descriptors, indices, areas, info);
  });
}

int gh_similarity_from_triangles_batch(const float* rx, const float* ry, int64_t nr, const float* rdesc,
                                       const int32_t* ridx, int64_t nrd, int64_t batch, const float* const* mx,
                                       const float* const* my, const int64_t* nm, const float* const* mdesc,
                                       const int32_t* const* midx, const int64_t* nmd, float tolerance_px,
                                       float descriptor_radius, gh_similarity_fit* out) {
  return guarded([&] {
    require(batch >= 0, "batch count must be non-negative");
    if (batch == 0) return;
    require(tolerance_px >= 0.0f, "tolerance_px must be non-negative");
    require(descriptor_radius >= 0.0f, "descriptor_radius must be non-negative");
    require(nr >= 3 && nrd > 0, "triangle descriptor similarity requires at least one descriptor and three stars");
    require(rx && ry, "reference catalog coordinate arrays are required");
    check_descriptor_table(rdesc, ridx, nrd, "reference descriptors and indices are required");
    require(mx && my && nm && mdesc && midx && nmd && out, "moving batch arrays are required");
    int64_t max_candidates = 0;
    for (int64_t b = 0; b < batch; ++b) {
      require(nm[b] >= 3 && nmd[b] > 0, "triangle descriptor similarity requires at least one descriptor and three stars");
      require(mx[b] && my[b], "moving catalog coordinate arrays are required");
      check_descriptor_table(mdesc[b], midx[b], nmd[b], "moving descriptors and indices are required");
      const int64_t candidates = nrd * nmd[b] * 2;
      require(nrd <= INT32_MAX / 2 && nmd[b] <= INT32_MAX / 2 && candidates <= INT32_MAX,
              "descriptor pair candidates exceed 2^31");
      max_candidates = std::max(max_candidates, candidates);
    }
    Session session;
    Workspace& ws = *session.ws;
    Packer pack;
    const CatalogUpload ref = pack_catalog(pack, rx, ry, nr);
    const size_t o_rdesc = pack.add(rdesc, sizeof(float) * 2 * static_cast<size_t>(nrd));
    const size_t o_ridx = pack.add(ridx, sizeof(int32_t) * 3 * static_cast<size_t>(nrd));
    std::vector<CatalogUpload> movs;
    std::vector<std::pair<size_t, size_t>> mtab;
    for (int64_t b = 0; b < batch; ++b) {
      movs.push_back(pack_catalog(pack, mx[b], my[b], nm[b]));
      const size_t od = pack.add(mdesc[b], sizeof(float) * 2 * static_cast<size_t>(nmd[b]));
      const size_t oi = pack.add(midx[b], sizeof(int32_t) * 3 * static_cast<size_t>(nmd[b]));
      mtab.emplace_back(od, oi);
    }
    unsigned char* base = pack.upload(ws);
    ws.counts.ensure(static_cast<size_t>(batch), "hipMalloc(candidate counts)");
    ws.block_best.ensure(static_cast<size_t>(gh::score_max_blocks()), "hipMalloc(block bests)");
    ws.results.ensure(static_cast<size_t>(batch), "hipMalloc(similarity results)");
    for (int64_t b = 0; b < batch; ++b) {
      const gh::ScoreCatalogs cat = score_catalogs(base, ref, movs[b]);
      const gh::CandidateList list =
          candidate_list(ws, static_cast<size_t>(std::max<int64_t>(max_candidates, 1)), ws.counts.ptr + b);
      run_estimate(ws, cat, tolerance_px, list, 0, ws.results.ptr + b, [&] {
        gh::launch_triangle_candidates(cat, reinterpret_cast<const float2*>(base + o_rdesc),
                                       reinterpret_cast<const int*>(base + o_ridx), static_cast<int>(nrd),
                                       reinterpret_cast<const float2*>(base + mtab[b].first),
                                       reinterpret_cast<const int*>(base + mtab[b].second),
                                       static_cast<int>(nmd[b]), descriptor_radius, list, ws.stream);
      });
    }
    std::vector<gh::SimilarityOut> results(static_cast<size_t>(batch));
    check(hipMemcpyAsync(results.data(), ws.results.ptr, results.size() * sizeof(gh::SimilarityOut),
                         hipMemcpyDeviceToHost, ws.stream),
          "hipMemcpy(similarity results)");
    check(hipStreamSynchronize(ws.stream), "triangle similarity");
    for (int64_t b = 0; b < batch; ++b) fill_fit(results[b], nrd * nmd[b] * 2, out + b);
  });
}

int gh_similarity_from_triangles(const float* rx, const float* ry, int64_t nr, const float* mx, const float* my,
                                 int64_t nm, const float* rdesc, const int32_t* ridx, int64_t nrd, const float* mdesc,
                                 const int32_t* midx, int64_t nmd, float tolerance_px, float descriptor_radius,
                                 gh_similarity_fit* out) {
  const float* mxs[1] = {mx};
  const float* mys[1] = {my};
  const int64_t nms[1] = {nm};
  const float* mds[1] = {mdesc};
  const int32_t* mis[1] = {midx};
  const int64_t nmds[1] = {nmd};
  return gh_similarity_from_triangles_batch(rx, ry, nr, rdesc, ridx, nrd, 1, mxs, mys, nms, mds, mis, nmds,
                                            tolerance_px, descriptor_radius, out);
}

int gh_similarity_from_catalogs(const float* rx, const float* ry, int64_t nr, const float* mx, const float* my,
                                int64_t nm, const gh_catalog_similarity_params* params, gh_similarity_fit* out,
                                gh_similarity_candidate* top, int32_t* top_count) {
  return guarded([&] {
    require(params && out, "parameters and output are required");
    require(nr >= 2 && nm >= 2, "similarity catalog estimation requires at least two stars per catalog");
    require(rx && ry && mx && my, "catalog coordinate arrays are required");
    gh_catalog_similarity_params p = *params;
    require(p.tolerance_px >= 0.0f, "tolerance_px must be non-negative");
    require(p.min_pair_distance >= 0.0f, "min_pair_distance must be non-negative");
    if (p.prior_radius_px < 0.0f) p.prior_dx = p.prior_dy = 0.0f;
    require(p.min_scale >= 0.0f && p.max_scale >= p.min_scale,
            "scale constraints must satisfy 0 <= min_scale <= max_scale");
    if (p.max_abs_rotation_rad < 0.0f) p.max_abs_rotation_rad = -1.0f;
    require(p.top_k >= 0, "top_k must be non-negative");
    if (p.top_k > 0) require(top && top_count, "top-k outputs are required");
    const long double cc = static_cast<long double>(nr) * (nr - 1) * nm * (nm - 1);
    require(cc <= INT32_MAX, "star pair candidates exceed 2^31");
    const int64_t candidate_count = nr * (nr - 1) * nm * (nm - 1);

    Session session;
    Workspace& ws = *session.ws;
    Packer pack;
    const CatalogUpload ref = pack_catalog(pack, rx, ry, nr);
    const CatalogUpload mov = pack_catalog(pack, mx, my, nm);
    unsigned char* base = pack.upload(ws);
    ws.counts.ensure(1, "hipMalloc(candidate counts)");
    ws.block_best.ensure(static_cast<size_t>(gh::score_max_blocks()), "hipMalloc(block bests)");
    ws.results.ensure(1, "hipMalloc(similarity results)");
    const gh::ScoreCatalogs cat = score_catalogs(base, ref, mov);
    const gh::CandidateList list = candidate_list(ws, static_cast<size_t>(candidate_count), ws.counts.ptr);
    const gh::CatalogGates gates{p.min_pair_distance, p.prior_dx,  p.prior_dy,           p.prior_radius_px,
                                 p.min_scale,         p.max_scale, p.max_abs_rotation_rad};
    run_estimate(ws, cat, p.tolerance_px, list, 1, ws.results.ptr,
                 [&] { gh::launch_pair_candidates(cat, gates, candidate_count, list, ws.stream); });
    gh::SimilarityOut result{};
    check(hipMemcpyAsync(&result, ws.results.ptr, sizeof(result), hipMemcpyDeviceToHost, ws.stream),
          "hipMemcpy(similarity result)");
    check(hipStreamSynchronize(ws.stream), "catalog similarity");
    fill_fit(result, candidate_count, out);
    if (p.top_k > 0) {
      const size_t n = result.scored;
      std::vector<int> index(n), score(n);
      std::vector<float> rms(n);
      std::vector<float4> prm(n);
      if (n) {
        check(hipMemcpyAsync(index.data(), list.index, n * sizeof(int), hipMemcpyDeviceToHost, ws.stream), "index");
        check(hipMemcpyAsync(score.data(), list.score, n * sizeof(int), hipMemcpyDeviceToHost, ws.stream), "score");
        check(hipMemcpyAsync(rms.data(), list.rms, n * sizeof(float), hipMemcpyDeviceToHost, ws.stream), "rms");
        check(hipMemcpyAsync(prm.data(), list.params, n * sizeof(float4), hipMemcpyDeviceToHost, ws.stream),
              "params");
        check(hipStreamSynchronize(ws.stream), "top candidates");
      }
      std::vector<TopEntry> all(n);
      for (size_t i = 0; i < n; ++i) all[i] = TopEntry{index[i], score[i], rms[i], prm[i]};
      const std::vector<TopEntry> best =
          top_candidates(std::move(all), p.top_k, static_cast<int>(nr), static_cast<int>(nm));
      for (size_t i = 0; i < best.size(); ++i) {
        const TopEntry& e = best[i];
        top[i] = gh_similarity_candidate{e.index, e.score, e.rms, e.params.x, e.params.y, e.params.z, e.params.w};
      }
      *top_count = static_cast<int32_t>(best.size());
    } else if (top_count) {
      *top_count = 0;
    }
  });
}

}  // extern "C"
