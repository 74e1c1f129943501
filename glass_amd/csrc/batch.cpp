// Batched multi-job stacking (gh_batch_*): many resident jobs as one
// persistent warp launch plus one persistent integration launch per capacity
// class, optionally captured into a hipGraph, with a small autotuner that
// times candidate schedules (persistent grid sizes, work-item order) on the
// first run and caches the choice per batch shape for the process.

#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <sstream>

#include "stack.hpp"

using ghi::guarded;

namespace {

struct Schedule {
  unsigned warp_grid = 0;
  int warp_stride = 1;
  std::vector<unsigned> int_grid;  // per integration class
};

std::mutex g_tune_mutex;
std::map<std::string, Schedule> g_tune_cache;

// A stride coprime to n (for interleaving work items over jobs).
int coprime_stride(long long n) {
  if (n <= 2) return 1;
  for (long long s : {97LL, 89LL, 83LL, 79LL, 73LL, 71LL, 67LL, 61LL}) {
    if (s < n && std::gcd(s, n) == 1) return static_cast<int>(s);
  }
  return 1;
}

}  // namespace

struct gh_batch {
  struct Job {
    gh_stack* stack = nullptr;
    float* warped = nullptr;          // n_warp frames, owned
    float* coverage = nullptr;        // owned, when geometric coverage is needed
    int n_unwarped = 0;               // frames adding full coverage
    bool mean_only = false;           // rejection none: per-job mean kernel
    gh::IntegrateArgs args{};
    DeviceBuffer<const float*> table;
    DeviceBuffer<float> weights;
    DeviceBuffer<const unsigned*> flag_ptrs;
  };
  struct Class {
    int capacity = 0;
    bool unit = false;
    std::vector<int> jobs;
    long long total_blocks = 0;
    DeviceBuffer<gh::IntegrateArgs> args;
    DeviceBuffer<long long> first_block;
  };
  std::vector<Job> jobs;
  std::vector<Class> classes;
  DeviceBuffer<gh::BatchWarpJob> warp_jobs;
  DeviceBuffer<gh::BatchWarpFrame> warp_frames;
  DeviceBuffer<unsigned> warped_flags;
  DeviceBuffer<float> lut;
  int n_warp_jobs = 0;
  long long total_items = 0;
  size_t total_warped = 0;
  uint32_t flags = 0;
  std::string signature;
  Schedule sched;
  bool tuned = false;
  int trials = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~gh_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (Job& j : jobs) {
      (void)hipFree(j.warped);
      (void)hipFree(j.coverage);
      j.table.release();
      j.weights.release();
      j.flag_ptrs.release();
    }
    for (Class& c : classes) {
      c.args.release();
      c.first_block.release();
    }
    warp_jobs.release();
    warp_frames.release();
    warped_flags.release();
    lut.release();
    if (stream) (void)hipStreamDestroy(stream);
  }

  // The whole batch on `st`: clear flags / coverage, warp, integrate.
  void enqueue(hipStream_t st, const Schedule& sc) {
    if (total_warped) {
      gh::launch_fill(reinterpret_cast<float*>(warped_flags.ptr), static_cast<long long>(total_warped), 0.0f, st);
    }
    for (Job& j : jobs) {
      if (!j.coverage) continue;
      gh::launch_fill(j.coverage, j.stack->npix, static_cast<float>(j.n_unwarped), st);
    }
    if (total_items) {
      gh::launch_warp_batched(warp_jobs.ptr, n_warp_jobs, warp_frames.ptr, total_items,
                              reinterpret_cast<const float4*>(lut.ptr), sc.warp_grid, sc.warp_stride, st);
    }
    for (size_t c = 0; c < classes.size(); ++c) {
      const Class& k = classes[c];
      gh::launch_integrate_batched(k.capacity, k.unit, k.args.ptr, k.first_block.ptr, static_cast<int>(k.jobs.size()),
                                   k.total_blocks, sc.int_grid[c], st);
    }
    for (Job& j : jobs) {
      if (j.mean_only) gh::launch_integrate(j.args, st);
    }
  }

  float time_ms(const Schedule& sc) {
    hipEvent_t a, b;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    check(hipEventRecord(a, stream), "hipEventRecord");
    enqueue(stream, sc);
    check(hipEventRecord(b, stream), "hipEventRecord");
    check(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ++trials;
    return ms;
  }

  // Coordinate descent over the warp schedule, then each class's grid.
  void autotune() {
    const unsigned base_w = gh::warp_batched_max_blocks();
    Schedule best = sched;
    float best_ms = time_ms(best);  // also the warm-up
    best_ms = std::min(best_ms, time_ms(best));
    auto try_sched = [&](const Schedule& cand) {
      const float ms = std::min(time_ms(cand), time_ms(cand));
      if (ms < best_ms) {
        best_ms = ms;
        best = cand;
      }
    };
    for (unsigned mult : {1u, 2u, 4u}) {
      for (int stride : {1, coprime_stride(total_items)}) {
        Schedule c = best;
        c.warp_grid = base_w * mult;
        c.warp_stride = stride;
        if (c.warp_grid == best.warp_grid && c.warp_stride == best.warp_stride) continue;
        try_sched(c);
      }
    }
    int dev = 0, cus = 1;
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device CU count");
    for (size_t k = 0; k < classes.size(); ++k) {
      for (unsigned mult : {2u, 4u, 8u}) {
        Schedule c = best;
        c.int_grid[k] = static_cast<unsigned>(cus) * mult;
        if (c.int_grid[k] == best.int_grid[k]) continue;
        try_sched(c);
      }
    }
    sched = best;
  }
};

extern "C" {

int gh_batch_create(const gh_batch_job* jobs, int64_t n_jobs, uint32_t flags, gh_batch** out) {
  return guarded([&] {
    if (!out) throw gh::ValueError("output pointer is NULL");
    *out = nullptr;
    if (n_jobs <= 0 || !jobs) throw gh::ValueError("a batch needs at least one job");
    auto b = std::make_unique<gh_batch>();
    b->flags = flags;
    check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate(batch)");
    // Lanczos3 weight table shared by all jobs.
    const size_t lut_floats = static_cast<size_t>(gh::kLanczosLutIntervals + 1) * gh::kLanczosLutStride;
    std::vector<float> lut(lut_floats);
    gh::make_lanczos_lut(lut.data());
    b->lut.ensure(lut_floats, "hipMalloc(batch lanczos table)");
    check(hipMemcpy(b->lut.ptr, lut.data(), lut_floats * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy(lut)");

    int tile_w = 0, tile_h = 0;
    gh::warp_tile_shape(&tile_w, &tile_h);
    std::ostringstream sig;
    size_t total_warped = 0;
    for (int64_t j = 0; j < n_jobs; ++j) total_warped += static_cast<size_t>(std::max<int64_t>(jobs[j].n_warp, 0));
    b->total_warped = total_warped;
    if (total_warped) b->warped_flags.ensure(total_warped, "hipMalloc(batch warp flags)");
    std::vector<gh::BatchWarpJob> wjobs;
    std::vector<gh::BatchWarpFrame> wframes;
    size_t flag_base = 0;
    b->jobs.resize(static_cast<size_t>(n_jobs));
    std::map<std::pair<int, bool>, int> class_of;
    for (int64_t j = 0; j < n_jobs; ++j) {
      const gh_batch_job& in = jobs[j];
      gh_batch::Job& job = b->jobs[static_cast<size_t>(j)];
      gh_stack* s = in.stack;
      if (!s) throw gh::ValueError("batch job has no stack");
      if (s->loaded_count != s->frames) throw std::runtime_error("all resident frames must be loaded before a batch");
      if (s->npix >= (int64_t(1) << 30)) throw gh::ValueError("frames of 2^30 pixels or more are not supported");
      job.stack = s;
      const gh_integrate_params& prm = in.params;
      if (prm.tile_size != 0) throw gh::ValueError("batched integration needs tile_size 0");
      if (prm.rejection != GH_REJECT_NONE && prm.rejection != GH_REJECT_WINSORIZED_SIGMA)
        throw gh::ValueError("unsupported rejection mode");
      if (prm.rejection == GH_REJECT_WINSORIZED_SIGMA) {
        if (!(prm.low_sigma > 0.0f) || !(prm.high_sigma > 0.0f))
          throw gh::ValueError("low_sigma and high_sigma must be positive");
        if (prm.min_samples < 1) throw gh::ValueError("min_samples must be at least 1");
        if (!(prm.max_reject_fraction >= 0.0f && prm.max_reject_fraction <= 1.0f))
          throw gh::ValueError("max_reject_fraction must be between 0 and 1");
      }
      if (prm.count_dtype != GH_COUNT_F32 && prm.count_dtype != GH_COUNT_U16)
        throw gh::ValueError("count_map_dtype must be float32 or uint16");
      if (in.n_warp < 0 || in.n_warp > s->frames || (in.n_warp > 0 && (!in.warp_indices || !in.inverse)))
        throw gh::ValueError("invalid warp list");
      if (in.n_warp > 0 && in.interpolation != GH_INTERP_BILINEAR && in.interpolation != GH_INTERP_LANCZOS3)
        throw gh::ValueError("unsupported warp interpolation");
      // Warped copies and warp work.
      std::vector<int64_t> warped_slot(static_cast<size_t>(s->frames), -1);
      if (in.n_warp > 0) {
        check(hipMalloc(reinterpret_cast<void**>(&job.warped), static_cast<size_t>(in.n_warp) * s->frame_bytes()),
              "hipMalloc(batch warped frames)");
        gh::BatchWarpJob wj{};
        wj.height = static_cast<int>(s->height);
        wj.width = static_cast<int>(s->width);
        wj.tiles_x = static_cast<int>((s->width + tile_w - 1) / tile_w);
        wj.tiles = wj.tiles_x * static_cast<int>((s->height + tile_h - 1) / tile_h);
        wj.interpolation = in.interpolation;
        wj.skip_nonfinite = (in.warp_flags & GH_WARP_SKIP_NONFINITE_TAPS) ? 1 : 0;
        wj.fill = in.fill;
        wj.clamp = in.clamping_threshold;
        wj.first_frame = static_cast<int>(wframes.size());
        wj.n_frames = static_cast<int>(in.n_warp);
        for (int64_t k = 0; k < in.n_warp; ++k) {
          const int64_t f = in.warp_indices[k];
          s->require_loaded(f, "batched warp");
          if (warped_slot[f] >= 0) throw gh::ValueError("warp indices must be unique");
          warped_slot[f] = k;
          const double* m = in.inverse + k * 9;
          for (int e = 0; e < 9; ++e)
            if (!std::isfinite(m[e])) throw gh::ValueError("warp inverse matrix must be finite");
          if (!(m[6] == 0.0 && m[7] == 0.0 && m[8] == 1.0))
            throw gh::ValueError("batched warps need affine inverses (last row 0 0 1)");
          gh::BatchWarpFrame fr{};
          fr.src = s->slot[f];
          fr.dst = job.warped + static_cast<size_t>(k) * static_cast<size_t>(s->npix);
          std::memcpy(fr.inv, m, sizeof(fr.inv));
          fr.src_flag = s->nonfinite + f;
          fr.dst_flag = b->warped_flags.ptr + flag_base + static_cast<size_t>(k);
          wframes.push_back(fr);
        }
        wjobs.push_back(wj);
      }
      const bool need_cov = in.outputs.dq && in.outputs.dq_use_geometric;
      job.n_unwarped = static_cast<int>(s->frames - in.n_warp);
      if (need_cov) {
        check(hipMalloc(reinterpret_cast<void**>(&job.coverage), s->frame_bytes()), "hipMalloc(batch coverage)");
        if (!wjobs.empty() && in.n_warp > 0) wjobs.back().coverage = job.coverage;
      }
      // Integration table (gh_integrate's active-frame rules).
      std::vector<const float*> table;
      std::vector<float> w;
      std::vector<const unsigned*> fl;
      bool unit = true;
      for (int64_t f = 0; f < s->frames; ++f) {
        const float wf = in.weights ? in.weights[f] : 1.0f;
        if (!(std::isfinite(wf) && wf > 0.0f)) continue;
        const int64_t k = warped_slot[f];
        table.push_back(k >= 0 ? job.warped + static_cast<size_t>(k) * static_cast<size_t>(s->npix) : s->slot[f]);
        fl.push_back(k >= 0 ? b->warped_flags.ptr + flag_base + static_cast<size_t>(k) : s->nonfinite + f);
        w.push_back(wf);
        unit = unit && (wf == 1.0f);
      }
      flag_base += static_cast<size_t>(std::max<int64_t>(in.n_warp, 0));
      const int n_active = static_cast<int>(table.size());
      if (prm.rejection == GH_REJECT_WINSORIZED_SIGMA && n_active > gh::integrate_max_frames())
        throw gh::ValueError("winsorized integration supports at most 256 positive-weight frames");
      const size_t padded = static_cast<size_t>(std::max(gh::integrate_max_frames(), n_active));
      const float* pad_ptr = n_active ? table.back() : s->slot[0];
      table.resize(padded, pad_ptr);
      w.resize(padded, 0.0f);
      job.table.ensure(padded, "hipMalloc(batch frame table)");
      job.weights.ensure(padded, "hipMalloc(batch frame weights)");
      job.flag_ptrs.ensure(std::max<size_t>(fl.size(), 1), "hipMalloc(batch frame flags)");
      check(hipMemcpy(job.table.ptr, table.data(), padded * sizeof(const float*), hipMemcpyHostToDevice),
            "hipMemcpy(batch frame table)");
      check(hipMemcpy(job.weights.ptr, w.data(), padded * sizeof(float), hipMemcpyHostToDevice),
            "hipMemcpy(batch frame weights)");
      if (!fl.empty()) {
        check(hipMemcpy(job.flag_ptrs.ptr, fl.data(), fl.size() * sizeof(const unsigned*), hipMemcpyHostToDevice),
              "hipMemcpy(batch frame flags)");
      }
      gh::IntegrateArgs& a = job.args;
      a.frame_ptr = job.table.ptr;
      a.frame_w = job.weights.ptr;
      a.n_active = n_active;
      a.npix = s->npix;
      a.width = static_cast<int>(s->width);
      a.rejection = prm.rejection;
      a.low_sigma = prm.low_sigma;
      a.high_sigma = prm.high_sigma;
      a.min_samples = prm.min_samples;
      a.max_frac = prm.max_reject_fraction;
      a.unit_weights = unit ? 1 : 0;
      a.require_positive_scale = prm.require_positive_scale;
      a.all_weights_positive = (n_active == s->frames) ? 1 : 0;
      if (prm.statistics_variant >= 0) {
        a.variant_all_valid = prm.statistics_variant;
      } else {
        a.active_flags = job.flag_ptrs.ptr;
      }
      const gh_integrate_outputs& o = in.outputs;
      for (const void* ptr : {static_cast<const void*>(o.master), static_cast<const void*>(o.weight),
                              static_cast<const void*>(o.coverage), static_cast<const void*>(o.low),
                              static_cast<const void*>(o.high), static_cast<const void*>(o.dq)}) {
        if (ptr && !is_device_pointer(ptr)) throw gh::ValueError("batch outputs must be device pointers");
      }
      a.master = o.master;
      a.weight = o.weight;
      a.coverage = o.coverage;
      a.low = o.low;
      a.high = o.high;
      a.dq = o.dq;
      a.count_u16 = (prm.count_dtype == GH_COUNT_U16) ? 1 : 0;
      a.geometric = need_cov ? job.coverage : nullptr;
      a.dq_active_frames = o.dq_active_frames;
      if (prm.rejection == GH_REJECT_NONE) {
        job.mean_only = true;
        if (a.active_flags) a.variant_all_valid = 0;  // unused by the mean kernel
      } else {
        const std::pair<int, bool> key{gh::integrate_capacity(n_active), unit};
        auto it = class_of.find(key);
        if (it == class_of.end()) {
          it = class_of.emplace(key, static_cast<int>(b->classes.size())).first;
          b->classes.emplace_back();
          b->classes.back().capacity = key.first;
          b->classes.back().unit = key.second;
        }
        b->classes[it->second].jobs.push_back(static_cast<int>(j));
      }
      sig << s->height << 'x' << s->width << ':' << in.n_warp << ':' << in.interpolation << ':' << n_active << ';';
    }
    // Device job tables.
    long long items = 0;
    for (gh::BatchWarpJob& wj : wjobs) {
      wj.first_item = items;
      items += gh::warp_batched_items(wj);
    }
    b->total_items = items;
    b->n_warp_jobs = static_cast<int>(wjobs.size());
    if (!wjobs.empty()) {
      b->warp_jobs.ensure(wjobs.size(), "hipMalloc(batch warp jobs)");
      b->warp_frames.ensure(wframes.size(), "hipMalloc(batch warp frames)");
      check(hipMemcpy(b->warp_jobs.ptr, wjobs.data(), wjobs.size() * sizeof(gh::BatchWarpJob), hipMemcpyHostToDevice),
            "hipMemcpy(batch warp jobs)");
      check(hipMemcpy(b->warp_frames.ptr, wframes.data(), wframes.size() * sizeof(gh::BatchWarpFrame),
                      hipMemcpyHostToDevice),
            "hipMemcpy(batch warp frames)");
    }
    int dev = 0, cus = 1;
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device CU count");
    b->sched.warp_grid = gh::warp_batched_max_blocks();
    b->sched.warp_stride = 1;
    for (gh_batch::Class& c : b->classes) {
      std::vector<gh::IntegrateArgs> args;
      std::vector<long long> first;
      long long blocks = 0;
      for (int j : c.jobs) {
        args.push_back(b->jobs[static_cast<size_t>(j)].args);
        first.push_back(blocks);
        blocks += gh::integrate_blocks(b->jobs[static_cast<size_t>(j)].args.npix);
      }
      c.total_blocks = blocks;
      c.args.ensure(args.size(), "hipMalloc(batch integrate jobs)");
      c.first_block.ensure(first.size(), "hipMalloc(batch integrate offsets)");
      check(hipMemcpy(c.args.ptr, args.data(), args.size() * sizeof(gh::IntegrateArgs), hipMemcpyHostToDevice),
            "hipMemcpy(batch integrate jobs)");
      check(hipMemcpy(c.first_block.ptr, first.data(), first.size() * sizeof(long long), hipMemcpyHostToDevice),
            "hipMemcpy(batch integrate offsets)");
      b->sched.int_grid.push_back(static_cast<unsigned>(cus) * 4u);
    }
    b->signature = sig.str();
    *out = b.release();
  });
}

int gh_batch_run(gh_batch* b, void* stream_ptr) {
  return guarded([&] {
    if (!b) throw gh::ValueError("batch is NULL");
    hipStream_t st = stream_ptr ? static_cast<hipStream_t>(stream_ptr) : b->stream;
    if (!b->tuned) {
      // Work queued on the caller's stream before this call must be complete
      // before the batch's own stream runs tuning or capture.
      if (st != b->stream) check(hipStreamSynchronize(st), "synchronize(caller stream)");
      if (b->flags & GH_BATCH_AUTOTUNE) {
        bool cached = false;
        {
          std::lock_guard<std::mutex> lock(g_tune_mutex);
          auto it = g_tune_cache.find(b->signature);
          if (it != g_tune_cache.end() && it->second.int_grid.size() == b->sched.int_grid.size()) {
            b->sched = it->second;
            cached = true;
          }
        }
        if (!cached) {
          b->autotune();
          std::lock_guard<std::mutex> lock(g_tune_mutex);
          g_tune_cache[b->signature] = b->sched;
        }
      }
      if (b->flags & GH_BATCH_USE_GRAPH) {
        check(hipStreamSynchronize(b->stream), "synchronize(batch)");
        check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
          b->enqueue(b->stream, b->sched);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(b->stream, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        check(hipStreamEndCapture(b->stream, &b->graph), "hipStreamEndCapture");
        check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      }
      b->tuned = true;
      if (st != b->stream) check(hipStreamSynchronize(b->stream), "synchronize(batch)");
    }
    if (b->exec) {
      check(hipGraphLaunch(b->exec, st), "hipGraphLaunch");
    } else {
      b->enqueue(st, b->sched);
    }
  });
}

int gh_batch_info(gh_batch* b, int64_t* out) {
  return guarded([&] {
    if (!b || !out) throw gh::ValueError("argument is NULL");
    out[0] = static_cast<int64_t>(b->jobs.size());
    out[1] = b->total_items;
    int64_t launches = static_cast<int64_t>(b->classes.size());
    for (const auto& j : b->jobs) launches += j.mean_only ? 1 : 0;
    out[2] = launches;
    out[3] = b->exec ? 1 : 0;
    out[4] = b->sched.warp_grid;
    out[5] = b->sched.warp_stride;
    out[6] = b->sched.int_grid.empty() ? 0 : b->sched.int_grid[0];
    out[7] = b->trials;
  });
}

int gh_batch_destroy(gh_batch* b) {
  return guarded([&] { delete b; });
}

}  // extern "C"
