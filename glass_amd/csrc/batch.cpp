// Batched multi-job stacking (gh_batch_*): the warps of many resident jobs
// run as one launch over (job, frame group, output tile) work items; a small
// kernel then decides each job's statistics arithmetic from the frames'
// non-finite flags; the integrations run as one launch per sort-network
// capacity class over the pixels of all jobs of the class (a job table read on
// the device: block -> job, the job's arguments, its first block), each job
// reading its statistics decision from device memory.  The whole sequence
// needs no host round trip and is captured into a hipGraph once and replayed.
// An autotuner times candidate schedules on the first run and caches the
// choice per batch shape for the process: the warp's grid size, work-item
// order and frames per work item, and the integration's capacity classes
// (each job at its smallest capacity, powers of two, or one launch at the
// largest capacity).  None of these choices changes a result: the warp
// computes each pixel of a frame the same way whatever the item holding it,
// and a sorting network of any capacity >= the sample count sorts the same
// keys (the padding sorts after them).

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
  int frames_per_item = gh::kWarpBatch;  // 2, 4 or 8 frames of one job per warp work item
  int classes = 0;                       // integration capacity classes (kClassSets)
};

// Frames per warp work item the schedule may use (divisors of kWarpBatch,
// so an item never straddles two of the per-job path's 8-frame groups).
constexpr int kItemFrames[] = {2, 4, 8};
constexpr int kItemFrameOptions = 3;
int item_frames_index(int f) { return f == 2 ? 0 : f == 4 ? 1 : 2; }

// Capacity class sets of the integration: 0 each job at its smallest
// register capacity, 1 powers of two, 2 every job at the largest capacity
// any job needs (a single launch per weight kind: unit or per-frame).
constexpr int kClassSets = 3;

int class_capacity(int set, int n_active, int largest) {
  if (set == 0) return gh::integrate_cap_for(n_active);
  if (set == 1) {
    int c = 16;
    while (c < n_active) c <<= 1;
    return c;
  }
  return largest;
}

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
  // Warp work items for one frames-per-item choice.
  struct WarpPlan {
    DeviceBuffer<gh::BatchWarpJob> jobs;
    DeviceBuffer<unsigned short> item_job;  // work item -> warp job
    long long total_items = 0;
  };
  // One integration launch over the jobs of a capacity class.
  struct ClassLaunch {
    int cap = 0;
    unsigned grid = 0;
    gh::IntegrateArgs first{};  // kernel argument (the table overrides it)
    DeviceBuffer<gh::IntegrateArgs> args;
    DeviceBuffer<unsigned short> block_job;
    DeviceBuffer<unsigned> first_block;
  };
  struct IntegratePlan {
    std::vector<ClassLaunch> launches;
    std::vector<size_t> per_job;  // jobs launched on their own (mean, more than 256 frames)
  };
  struct Job {
    gh_stack* stack = nullptr;
    float* warped = nullptr;          // n_warp frames, owned
    float* coverage = nullptr;        // owned, when geometric coverage is needed
    int n_unwarped = 0;               // frames adding full coverage
    gh::IntegrateArgs args{};
    DeviceBuffer<const float*> table;
    DeviceBuffer<float> weights;
    DeviceBuffer<const unsigned*> flag_ptrs;
  };
  std::vector<Job> jobs;
  // [0]: the Lanczos3 jobs' items, [1]: the bilinear jobs' items (a launch
  // each), per frames-per-item choice
  WarpPlan wplans[2][kItemFrameOptions];
  IntegratePlan iplans[kClassSets];
  DeviceBuffer<gh::BatchWarpFrame> warp_frames;
  DeviceBuffer<gh::VariantJob> variant_jobs;
  DeviceBuffer<int> variants;              // per variant job, written on the device
  int n_variant_jobs = 0;
  DeviceBuffer<unsigned> warped_flags;
  DeviceBuffer<float> lut;
  int n_warp_jobs = 0;
  size_t total_warped = 0;
  uint32_t flags = 0;
  std::string signature;
  Schedule sched;
  bool tuned = false;
  bool pinned = false;  // schedule set by gh_batch_set_schedule (no autotune)
  int trials = 0;
  hipStream_t stream = nullptr;
  // Side streams and events of the fork / join in enqueue (default priority).
  static constexpr int kSides = 3;
  hipStream_t side[kSides] = {};
  hipEvent_t ev_fork = nullptr, ev_warp = nullptr, ev_decided = nullptr;
  hipEvent_t ev_join[kSides] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~gh_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (hipStream_t s : side)
      if (s) (void)hipStreamSynchronize(s);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (Job& j : jobs) {
      (void)hipFree(j.warped);
      (void)hipFree(j.coverage);
      j.table.release();
      j.weights.release();
      j.flag_ptrs.release();
    }
    for (auto& kind : wplans) {
      for (WarpPlan& w : kind) {
        w.jobs.release();
        w.item_job.release();
      }
    }
    for (IntegratePlan& ip : iplans) {
      for (ClassLaunch& l : ip.launches) {
        l.args.release();
        l.block_job.release();
        l.first_block.release();
      }
    }
    warp_frames.release();
    variant_jobs.release();
    variants.release();
    warped_flags.release();
    lut.release();
    for (hipEvent_t e : {ev_fork, ev_warp, ev_decided})
      if (e) (void)hipEventDestroy(e);
    for (int k = 0; k < kSides; ++k) {
      if (ev_join[k]) (void)hipEventDestroy(ev_join[k]);
      if (side[k]) (void)hipStreamDestroy(side[k]);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }

  // The whole batch, ordered after the work already queued on `st` and
  // complete when `st` reaches its end: clear flags / coverage, warp,
  // integrate.  The launches that do not depend on each other run side by
  // side (a fork / join over `st` and the side streams, recorded as parallel
  // branches when captured into the graph): the Lanczos3 and the bilinear
  // warp launches (different jobs), then the integration class launches.
  void enqueue(hipStream_t st, const Schedule& sc) {
    if (total_warped) {
      gh::launch_fill(reinterpret_cast<float*>(warped_flags.ptr), static_cast<long long>(total_warped), 0.0f, st);
    }
    for (Job& j : jobs) {
      if (!j.coverage) continue;
      gh::launch_fill(j.coverage, j.stack->npix, static_cast<float>(j.n_unwarped), st);
    }
    const int o = item_frames_index(sc.frames_per_item);
    const bool both = wplans[0][o].total_items && wplans[1][o].total_items;
    if (both) {
      check(hipEventRecord(ev_fork, st), "hipEventRecord(batch fork)");
      check(hipStreamWaitEvent(side[0], ev_fork, 0), "hipStreamWaitEvent(batch fork)");
      warp_kind(0, side[0], sc);
      check(hipEventRecord(ev_warp, side[0]), "hipEventRecord(batch warp)");
      warp_kind(1, st, sc);
      check(hipStreamWaitEvent(st, ev_warp, 0), "hipStreamWaitEvent(batch warp)");
    } else {
      warp(st, sc);
    }
    gh::launch_decide_variants(variant_jobs.ptr, n_variant_jobs, variants.ptr, st);
    integrate(st, sc, true);
  }

  // The Lanczos3 launch takes the schedule's grid and item order; the
  // bilinear launch runs one workgroup per item in order.
  void warp_kind(int kind, hipStream_t st, const Schedule& sc) {
    const int o = item_frames_index(sc.frames_per_item);
    const WarpPlan& w = wplans[kind][o];
    if (!w.total_items) return;
    if (kind == 0) {
      gh::launch_warp_batched(true, w.jobs.ptr, w.item_job.ptr, warp_frames.ptr, w.total_items,
                              reinterpret_cast<const float4*>(lut.ptr), sc.warp_grid, sc.warp_stride, st);
    } else {
      gh::launch_warp_batched(false, w.jobs.ptr, w.item_job.ptr, warp_frames.ptr, w.total_items, nullptr,
                              static_cast<unsigned>(std::min<long long>(w.total_items, 0x7fffffffLL)), 1, st);
    }
  }
  void warp(hipStream_t st, const Schedule& sc) {
    warp_kind(0, st, sc);
    warp_kind(1, st, sc);
  }

  // The integration launches; with `fork`, class launch i runs on stream
  // i mod (1 + kSides) (st or a side stream, each after the launches queued
  // on st so far) and st waits for all of them at the end.
  void integrate(hipStream_t st, const Schedule& sc, bool fork = false) {
    const IntegratePlan& ip = iplans[sc.classes];
    bool used[kSides] = {};
    if (fork && ip.launches.size() > 1) check(hipEventRecord(ev_decided, st), "hipEventRecord(batch decided)");
    for (size_t i = 0; i < ip.launches.size(); ++i) {
      const ClassLaunch& l = ip.launches[i];
      hipStream_t s = st;
      const int k = static_cast<int>(i % (1 + kSides)) - 1;
      if (fork && ip.launches.size() > 1 && k >= 0) {
        s = side[k];
        if (!used[k]) check(hipStreamWaitEvent(s, ev_decided, 0), "hipStreamWaitEvent(batch decided)");
        used[k] = true;
      }
      gh::launch_integrate_cap(l.cap, l.first, gh::IntegrateJobTable{l.args.ptr, l.block_job.ptr, l.first_block.ptr},
                               l.grid, s);
    }
    for (size_t j : ip.per_job) gh::launch_integrate(jobs[j].args, st);
    for (int k = 0; k < kSides; ++k) {
      if (!used[k]) continue;
      check(hipEventRecord(ev_join[k], side[k]), "hipEventRecord(batch join)");
      check(hipStreamWaitEvent(st, ev_join[k], 0), "hipStreamWaitEvent(batch join)");
    }
  }

  long long lanczos_items(int frames_per_item) const {
    return wplans[0][item_frames_index(frames_per_item)].total_items;
  }
  long long total_items() const {
    const int o = item_frames_index(sched.frames_per_item);
    return wplans[0][o].total_items + wplans[1][o].total_items;
  }
  int integrate_launches() const {
    const IntegratePlan& ip = iplans[sched.classes];
    return static_cast<int>(ip.launches.size() + ip.per_job.size());
  }

  // Device time of `launch` on the batch's stream (best of two runs).
  template <typename F>
  float time_launch(F&& launch) {
    hipEvent_t a, b;
    check(hipEventCreate(&a), "hipEventCreate");
    check(hipEventCreate(&b), "hipEventCreate");
    float best = 0.0f;
    for (int rep = 0; rep < 2; ++rep) {
      check(hipEventRecord(a, stream), "hipEventRecord");
      launch();
      check(hipEventRecord(b, stream), "hipEventRecord");
      check(hipEventSynchronize(b), "hipEventSynchronize");
      float ms = 0.0f;
      check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
      best = rep == 0 ? ms : std::min(best, ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ++trials;
    return best;
  }

  // The warp launch is tuned on its own timing: first the frames per work
  // item at the default grid, then the grid (persistent multiples of the
  // resident-block count, or one workgroup per item) and the item order.  The
  // integration is tuned over its capacity class sets.  Tuning runs write the
  // batch's own buffers; every real run starts by clearing them.
  void autotune() {
    enqueue(stream, sched);  // warm-up
    check(hipStreamSynchronize(stream), "synchronize(batch)");
    if (lanczos_items(gh::kWarpBatch) + wplans[1][item_frames_index(gh::kWarpBatch)].total_items) {
      const unsigned base_w = gh::warp_batched_max_blocks(true);
      auto warp_time = [&](const Schedule& sc) { return time_launch([&] { warp(stream, sc); }); };
      float best = warp_time(sched);
      for (int fpi : kItemFrames) {
        if (fpi == sched.frames_per_item) continue;
        Schedule sc = sched;
        sc.frames_per_item = fpi;
        const float ms = warp_time(sc);
        if (ms < best) {
          best = ms;
          sched = sc;
        }
      }
      const long long items = lanczos_items(sched.frames_per_item);
      const unsigned all_items = static_cast<unsigned>(std::min<long long>(std::max(items, 1LL), 0x7fffffffLL));
      const Schedule base = sched;
      for (unsigned grid : {base_w, base_w * 4u, all_items}) {
        for (int stride : {1, coprime_stride(items)}) {
          Schedule sc = base;
          sc.warp_grid = grid;
          sc.warp_stride = stride;
          if (sc.warp_grid == base.warp_grid && sc.warp_stride == base.warp_stride) continue;
          const float ms = warp_time(sc);
          if (ms < best) {
            best = ms;
            sched = sc;
          }
        }
      }
    }
    warp(stream, sched);  // the integration candidates read these warped frames
    gh::launch_decide_variants(variant_jobs.ptr, n_variant_jobs, variants.ptr, stream);
    float best_i = time_launch([&] { integrate(stream, sched, true); });
    const Schedule base = sched;
    for (int set = 0; set < kClassSets; ++set) {
      if (set == base.classes) continue;
      Schedule sc = base;
      sc.classes = set;
      const float ms = time_launch([&] { integrate(stream, sc, true); });
      if (ms < best_i) {
        best_i = ms;
        sched = sc;
      }
    }
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
    for (int k = 0; k < gh_batch::kSides; ++k) {
      check(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking), "hipStreamCreate(batch side)");
      check(hipEventCreateWithFlags(&b->ev_join[k], hipEventDisableTiming), "hipEventCreate(batch)");
    }
    for (hipEvent_t* e : {&b->ev_fork, &b->ev_warp, &b->ev_decided})
      check(hipEventCreateWithFlags(e, hipEventDisableTiming), "hipEventCreate(batch)");
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
    std::vector<gh::VariantJob> vjobs;
    std::vector<size_t> vjob_of;  // job index of each variant job
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
        // Lanczos3 kernel choice per group of kWarpBatch frames, as the
        // per-job path launches them.
        for (int64_t g0 = 0; g0 < in.n_warp; g0 += gh::kWarpBatch) {
          const int n = static_cast<int>(std::min<int64_t>(gh::kWarpBatch, in.n_warp - g0));
          double inv[gh::kWarpBatch][9];
          for (int j = 0; j < n; ++j) std::memcpy(inv[j], in.inverse + (g0 + j) * 9, sizeof(inv[j]));
          const int streamed = in.interpolation == GH_INTERP_LANCZOS3 &&
                                       gh::lanczos_stream_eligible(inv, n, static_cast<int>(s->width),
                                                                   static_cast<int>(s->height))
                                   ? 1
                                   : 0;
          for (int j = 0; j < n; ++j) wframes[wj.first_frame + g0 + j].stream = streamed;
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
        throw gh::ValueError("winsorized integration supports at most 16384 positive-weight frames");
      const size_t padded = static_cast<size_t>(std::max(gh::integrate_register_capacity(), n_active));
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
      if (prm.statistics_variant >= 0 || prm.rejection == GH_REJECT_NONE) {
        a.variant_all_valid = std::max(prm.statistics_variant, 0);  // unused by the mean kernel
      } else {
        // decided on the device after the warp (gh_integrate's host rule)
        vjobs.push_back(gh::VariantJob{job.flag_ptrs.ptr, n_active, n_active == s->frames ? 1 : 0});
        vjob_of.push_back(static_cast<size_t>(j));
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
      sig << s->height << 'x' << s->width << ':' << in.n_warp << ':' << in.interpolation << ':' << n_active << ';';
    }
    // Warp work items, one table per frames-per-item choice.
    b->n_warp_jobs = static_cast<int>(wjobs.size());
    if (wjobs.size() > 65535) throw gh::ValueError("a batch holds at most 65535 warped jobs");
    if (!wjobs.empty()) {
      b->warp_frames.ensure(wframes.size(), "hipMalloc(batch warp frames)");
      check(hipMemcpy(b->warp_frames.ptr, wframes.data(), wframes.size() * sizeof(gh::BatchWarpFrame),
                      hipMemcpyHostToDevice),
            "hipMemcpy(batch warp frames)");
      for (int kind = 0; kind < 2; ++kind) {
        std::vector<gh::BatchWarpJob> kj;
        for (const gh::BatchWarpJob& wj : wjobs)
          if ((wj.interpolation == GH_INTERP_LANCZOS3) == (kind == 0)) kj.push_back(wj);
        if (kj.empty()) continue;
        for (int o = 0; o < kItemFrameOptions; ++o) {
          std::vector<gh::BatchWarpJob> pj = kj;
          long long items = 0;
          for (gh::BatchWarpJob& wj : pj) {
            wj.frames_per_item = kItemFrames[o];
            wj.first_item = items;
            items += gh::warp_batched_items(wj);
          }
          // Work item -> warp job, so a workgroup finds its job with one load.
          std::vector<unsigned short> item_job(static_cast<size_t>(items));
          for (size_t k = 0; k < pj.size(); ++k) {
            const long long end = k + 1 < pj.size() ? pj[k + 1].first_item : items;
            std::fill(item_job.begin() + pj[k].first_item, item_job.begin() + end, static_cast<unsigned short>(k));
          }
          gh_batch::WarpPlan& w = b->wplans[kind][o];
          w.total_items = items;
          w.item_job.ensure(item_job.size(), "hipMalloc(batch item table)");
          check(hipMemcpy(w.item_job.ptr, item_job.data(), item_job.size() * sizeof(unsigned short),
                          hipMemcpyHostToDevice),
                "hipMemcpy(batch item table)");
          w.jobs.ensure(pj.size(), "hipMalloc(batch warp jobs)");
          check(hipMemcpy(w.jobs.ptr, pj.data(), pj.size() * sizeof(gh::BatchWarpJob), hipMemcpyHostToDevice),
                "hipMemcpy(batch warp jobs)");
        }
      }
    }
    b->n_variant_jobs = static_cast<int>(vjobs.size());
    if (!vjobs.empty()) {
      b->variant_jobs.ensure(vjobs.size(), "hipMalloc(batch variant jobs)");
      b->variants.ensure(vjobs.size(), "hipMalloc(batch variants)");
      check(hipMemcpy(b->variant_jobs.ptr, vjobs.data(), vjobs.size() * sizeof(gh::VariantJob),
                      hipMemcpyHostToDevice),
            "hipMemcpy(batch variant jobs)");
      for (size_t k = 0; k < vjob_of.size(); ++k) b->jobs[vjob_of[k]].args.variant_ptr = b->variants.ptr + k;
    }
    // Integration launches per capacity class set: winsorized jobs of at most
    // 256 positive-weight frames share register-kernel launches through job
    // tables; rejection-free and larger jobs launch on their own.
    int largest = 0;
    for (const gh_batch::Job& job : b->jobs) {
      const gh::IntegrateArgs& a = job.args;
      if (a.rejection != 0 && a.n_active <= gh::integrate_register_capacity())
        largest = std::max(largest, gh::integrate_cap_for(a.n_active));
    }
    for (int set = 0; set < kClassSets; ++set) {
      gh_batch::IntegratePlan& ip = b->iplans[set];
      // (capacity, unit weights): the kernel variant of a launch
      std::map<std::pair<int, int>, std::vector<size_t>> classes;
      for (size_t j = 0; j < b->jobs.size(); ++j) {
        const gh::IntegrateArgs& a = b->jobs[j].args;
        if (a.rejection != 0 && a.n_active <= gh::integrate_register_capacity())
          classes[{class_capacity(set, a.n_active, largest), a.unit_weights}].push_back(j);
        else
          ip.per_job.push_back(j);
      }
      for (const auto& [key, members] : classes) {
        gh_batch::ClassLaunch l;
        l.cap = key.first;
        std::vector<gh::IntegrateArgs> args;
        std::vector<unsigned> first;
        std::vector<unsigned short> block_job;
        for (size_t k = 0; k < members.size(); ++k) {
          const gh::IntegrateArgs& a = b->jobs[members[k]].args;
          args.push_back(a);
          first.push_back(static_cast<unsigned>(block_job.size()));
          const long long blocks = (a.npix + gh::kBlock - 1) / gh::kBlock;
          if (static_cast<long long>(block_job.size()) + blocks > 0x7fffffffLL)
            throw gh::ValueError("a batch integration class covers at most 2^31 - 1 blocks of pixels");
          block_job.insert(block_job.end(), static_cast<size_t>(blocks), static_cast<unsigned short>(k));
        }
        l.first = args.front();
        l.grid = static_cast<unsigned>(block_job.size());
        l.args.ensure(args.size(), "hipMalloc(batch integration jobs)");
        l.first_block.ensure(first.size(), "hipMalloc(batch integration job blocks)");
        l.block_job.ensure(block_job.size(), "hipMalloc(batch integration block table)");
        check(hipMemcpy(l.args.ptr, args.data(), args.size() * sizeof(gh::IntegrateArgs), hipMemcpyHostToDevice),
              "hipMemcpy(batch integration jobs)");
        check(hipMemcpy(l.first_block.ptr, first.data(), first.size() * sizeof(unsigned), hipMemcpyHostToDevice),
              "hipMemcpy(batch integration job blocks)");
        check(hipMemcpy(l.block_job.ptr, block_job.data(), block_job.size() * sizeof(unsigned short),
                        hipMemcpyHostToDevice),
              "hipMemcpy(batch integration block table)");
        ip.launches.push_back(l);
      }
    }
    b->sched.warp_grid = gh::warp_batched_max_blocks(true);
    b->sched.warp_stride = 1;
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
      if ((b->flags & GH_BATCH_AUTOTUNE) && !b->pinned) {
        bool cached = false;
        {
          std::lock_guard<std::mutex> lock(g_tune_mutex);
          auto it = g_tune_cache.find(b->signature);
          if (it != g_tune_cache.end()) {
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
    if (!stream_ptr) check(hipStreamSynchronize(st), "synchronize(batch)");
  });
}

int gh_batch_info(gh_batch* b, int64_t* out) {
  return guarded([&] {
    if (!b || !out) throw gh::ValueError("argument is NULL");
    out[0] = static_cast<int64_t>(b->jobs.size());
    out[1] = b->total_items();
    out[2] = b->integrate_launches();
    out[3] = b->exec ? 1 : 0;
    out[4] = b->sched.warp_grid;
    out[5] = b->sched.warp_stride;
    out[6] = b->n_variant_jobs;
    out[7] = b->trials;
    out[8] = b->sched.frames_per_item;
    out[9] = b->sched.classes;
  });
}

int gh_batch_set_schedule(gh_batch* b, int64_t warp_grid, int32_t warp_item_stride, int32_t frames_per_item,
                          int32_t integrate_classes) {
  return guarded([&] {
    if (!b) throw gh::ValueError("batch is NULL");
    if (frames_per_item != 2 && frames_per_item != 4 && frames_per_item != 8)
      throw gh::ValueError("frames_per_item must be 2, 4 or 8");
    if (integrate_classes < 0 || integrate_classes >= kClassSets)
      throw gh::ValueError("integrate_classes must be 0, 1 or 2");
    if (warp_grid < 1 || warp_grid > 0x7fffffffLL) throw gh::ValueError("warp_grid must be in [1, 2^31 - 1]");
    if (warp_item_stride < 1) throw gh::ValueError("warp_item_stride must be positive");
    const long long items = b->lanczos_items(frames_per_item);
    if (warp_item_stride > 1 && items > 0 && std::gcd(static_cast<long long>(warp_item_stride), items) != 1)
      throw gh::ValueError("warp_item_stride must be coprime to the number of warp work items");
    check(hipStreamSynchronize(b->stream), "synchronize(batch)");
    if (b->exec) check(hipGraphExecDestroy(b->exec), "hipGraphExecDestroy");
    if (b->graph) check(hipGraphDestroy(b->graph), "hipGraphDestroy");
    b->exec = nullptr;
    b->graph = nullptr;
    b->sched.warp_grid = static_cast<unsigned>(warp_grid);
    b->sched.warp_stride = warp_item_stride;
    b->sched.frames_per_item = frames_per_item;
    b->sched.classes = integrate_classes;
    b->pinned = true;
    b->tuned = false;  // the next run captures the graph with this schedule
  });
}

int gh_batch_destroy(gh_batch* b) {
  return guarded([&] { delete b; });
}

}  // extern "C"
