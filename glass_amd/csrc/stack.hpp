// Internal state of a gh_stack handle and the C-ABI error plumbing, shared by
// the translation units that implement include/glass_hip.h.
//
// gh_stack mirrors the device state of the reference ResidentCalibratedStack
// (cpp/src/native_bindings.cpp:2915-2996): an F x H x W float32 frame stack in
// HBM, optional masters, a geometric warp-coverage accumulator and a stream.
// Frames are addressed through a slot table so a warp can write a frame into a
// spare slot and swap, instead of warping into scratch and copying back
// (native_bindings.cpp:15442-15524 + cpp/cuda/warp_kernels.cu:806-914).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/glass_hip.h"
#include "common.hpp"
#include "kernels.hpp"

using gh::check;

namespace ghi {

inline thread_local std::string g_last_error;

inline int fail(int code, const std::string& message) {
  g_last_error = message;
  return code;
}

template <typename F>
int guarded(F&& body) {
  try {
    body();
    return GH_OK;
  } catch (const gh::ValueError& e) {
    return fail(GH_ERR_VALUE, e.what());
  } catch (const gh::IndexError& e) {
    return fail(GH_ERR_INDEX, e.what());
  } catch (const std::bad_alloc& e) {
    return fail(GH_ERR_RUNTIME, std::string("host allocation failed: ") + e.what());
  } catch (const std::exception& e) {
    return fail(GH_ERR_RUNTIME, e.what());
  }
}

inline bool is_device_pointer(const void* ptr) {
  if (ptr == nullptr) return false;
  hipPointerAttribute_t attr;
  const hipError_t err = hipPointerGetAttributes(&attr, ptr);
  if (err != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

template <typename T>
struct DeviceBuffer {
  T* ptr = nullptr;
  size_t count = 0;
  void ensure(size_t n, const char* what) {
    if (n <= count && ptr != nullptr) return;
    release();
    check(hipMalloc(reinterpret_cast<void**>(&ptr), std::max<size_t>(n, 1) * sizeof(T)), what);
    count = n;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
};

// frame_global_stats from FP64 moments (native_bindings.cpp:10626-10640):
// mean = sum / n, std = sqrt(max(sum_sq / n - mean^2, 0)); zeros for n = 0.
inline void fill_frame_statistics(const gh::FrameMoments& m, int64_t npix, gh_frame_statistics* out) {
  const int64_t n = static_cast<int64_t>(m.finite);
  out->sum = m.sum;
  out->sum_sq = m.sum_sq;
  out->valid_pixels = n;
  out->total_pixels = npix;
  out->nonfinite_pixels = npix - n;
  out->mean = n ? m.sum / static_cast<double>(n) : 0.0;
  double var = n ? m.sum_sq / static_cast<double>(n) - out->mean * out->mean : 0.0;
  if (var < 0.0) var = 0.0;
  out->std = std::sqrt(var);
  out->min = n ? m.min : 0.0f;
  out->max = n ? m.max : 0.0f;
  out->median = 0.0;
  out->mad = 0.0;
}

}  // namespace ghi

using ghi::DeviceBuffer;
using ghi::is_device_pointer;

struct gh_stack {
  int64_t frames = 0, height = 0, width = 0, npix = 0;
  float* pool = nullptr;                 // frames x npix allocation
  std::vector<float*> slot;              // frame -> current slot
  std::vector<float*> spare;             // free slots for out-of-place warps
  std::vector<float*> spare_allocations; // separately allocated spare slots
  std::vector<char> loaded;
  int64_t loaded_count = 0;
  unsigned* nonfinite = nullptr;         // [frames] flags
  float* bias = nullptr;
  float* dark = nullptr;
  float* flat = nullptr;
  float* geometric = nullptr;            // warp coverage accumulator
  int64_t warp_coverage_frames = 0;
  hipStream_t stream = nullptr;

  DeviceBuffer<gh::StatPartial> partials;
  DeviceBuffer<gh::StatPartial> calib_stats;  // [frames]
  std::vector<char> calib_stats_valid;
  DeviceBuffer<unsigned char> raw_staging;  // kRawRing host-raw upload slots
  DeviceBuffer<const float*> frame_table;
  DeviceBuffer<float> frame_weights;
  DeviceBuffer<unsigned char> tile_flags;
  DeviceBuffer<float> out_f32[5];
  DeviceBuffer<unsigned> out_dq;
  DeviceBuffer<gh::FrameMoments> moments;
  DeviceBuffer<unsigned> radix;
  DeviceBuffer<double> adj_ws;              // refinement partials + sums
  DeviceBuffer<unsigned> star_cand;         // [1 + capacity] count, candidate pixels
  DeviceBuffer<float> star_windows;         // gathered candidate windows
  DeviceBuffer<unsigned> warp_flags;        // [kWarpBatch] source-flag snapshot
  DeviceBuffer<float> lanczos_lut;          // normalised Lanczos3 weight cubics
  DeviceBuffer<double> ln_sums;             // local normalisation: per-tile FP64 sums
  DeviceBuffer<unsigned long long> ln_counts;  //   and paired-pixel counts
  DeviceBuffer<float> ln_coef;              //   host coefficients staged for the application
  DeviceBuffer<unsigned char> clip_tables;  // clip integration: frame table, weights, inverses
  DeviceBuffer<float> clip_geometric;       //   host geometric-coverage output staging
  DeviceBuffer<float> clip_tile_map;        //   tile-local per-pixel multiplier
  // star candidate catalogues: rank keys, sorted keys, per-cell top-k, cell
  // scratch, sort temporaries, decoded list, centroid status, counters
  DeviceBuffer<unsigned long long> cand_keys, cand_sorted, cand_selected;
  DeviceBuffer<unsigned> cand_u32, cand_count;
  DeviceBuffer<unsigned char> cand_keep, cand_temp, cand_status;
  DeviceBuffer<float> cand_xyf;
  DeviceBuffer<float> align_inv;            // alignment-metric candidates: float32 inverses
  DeviceBuffer<double> align_partials;      //   (candidate, row chunk) FP64 partial sums
  DeviceBuffer<const float*> align_frames;  //   moving frame of each candidate
  DeviceBuffer<double> ncc_ws;              // translation search partials + scores

  hipStream_t pick(void* s) const { return s ? static_cast<hipStream_t>(s) : stream; }

  size_t frame_bytes() const { return static_cast<size_t>(npix) * sizeof(float); }

  void require_index(int64_t i) const {
    if (i < 0 || i >= frames) throw gh::IndexError("resident frame index is out of range");
  }
  void require_loaded(int64_t i, const char* op) const {
    require_index(i);
    if (!loaded[i]) throw std::runtime_error(std::string("resident frame must be loaded before ") + op);
  }
  void mark_loaded(int64_t i) {
    if (!loaded[i]) {
      loaded[i] = 1;
      ++loaded_count;
    }
  }

  int64_t bytes_allocated() const {
    int64_t total = static_cast<int64_t>(frames) * static_cast<int64_t>(frame_bytes());
    total += static_cast<int64_t>(spare_allocations.size()) * static_cast<int64_t>(frame_bytes());
    if (bias) total += frame_bytes();
    if (dark) total += frame_bytes();
    if (flat) total += frame_bytes();
    if (geometric) total += frame_bytes();
    return total;
  }

  // The coverage entry points (reset / full-frame accumulate) take no stream
  // and run on the stack's own stream; a warp or integrate still queued on a
  // caller stream may touch the coverage map, so they wait on this event,
  // recorded on that caller stream after such work.
  hipEvent_t caller_done = nullptr;
  bool caller_recorded = false;

  void mark_caller(hipStream_t st) {
    if (st == stream) return;
    if (!caller_done) check(hipEventCreateWithFlags(&caller_done, hipEventDisableTiming), "hipEventCreate");
    check(hipEventRecord(caller_done, st), "hipEventRecord(caller stream)");
    caller_recorded = true;
  }

  void wait_caller() {
    if (caller_recorded) check(hipStreamWaitEvent(stream, caller_done, 0), "hipStreamWaitEvent(caller stream)");
  }

  // Device tables that the host rewrites before each integration launch
  // (frame pointers, weights, inverses, tile flags): a new upload first waits
  // for the launch that last read them, whichever stream it ran on.
  hipEvent_t tables_done = nullptr;

  bool tables_recorded = false;

  void wait_tables() {
    if (tables_recorded) check(hipEventSynchronize(tables_done), "integration tables in use");
  }

  void mark_tables(hipStream_t st) {
    if (!tables_done) check(hipEventCreateWithFlags(&tables_done, hipEventDisableTiming), "hipEventCreate");
    check(hipEventRecord(tables_done, st), "hipEventRecord(integration tables)");
    tables_recorded = true;
  }

  // Host raw frames go through a ring of device slots filled on a copy
  // stream: the upload of frame k + 1 overlaps the calibration of frame k.
  // Slot i is free again once the calibration that read it has run
  // (ring_free[i], recorded on the calibrating stream).
  static constexpr int kRawRing = 3;
  hipStream_t copy_stream = nullptr;
  hipEvent_t ring_copied[kRawRing] = {};
  hipEvent_t ring_free[kRawRing] = {};
  int ring_next = 0;
  size_t ring_slot_bytes = 0;

  unsigned char* ring_slot(size_t frame_raw_bytes, int& slot) {
    const size_t need = (frame_raw_bytes + 255) / 256 * 256;
    if (!copy_stream) {
      check(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking), "hipStreamCreate(raw upload)");
      for (int i = 0; i < kRawRing; ++i) {
        check(hipEventCreateWithFlags(&ring_copied[i], hipEventDisableTiming), "hipEventCreate(raw ring)");
        check(hipEventCreateWithFlags(&ring_free[i], hipEventDisableTiming), "hipEventCreate(raw ring)");
        check(hipEventRecord(ring_free[i], stream), "hipEventRecord(raw ring)");
      }
    }
    if (need > ring_slot_bytes) {
      // every slot must be idle before the ring is reallocated
      for (int i = 0; i < kRawRing; ++i) check(hipEventSynchronize(ring_free[i]), "raw ring drain");
      check(hipStreamSynchronize(copy_stream), "raw ring drain");
      raw_staging.release();
      raw_staging.ensure(need * kRawRing, "hipMalloc(raw upload ring)");
      ring_slot_bytes = need;
    }
    slot = ring_next;
    ring_next = (ring_next + 1) % kRawRing;
    return raw_staging.ptr + static_cast<size_t>(slot) * ring_slot_bytes;
  }

  void ensure_geometric() {
    if (geometric) return;
    check(hipMalloc(reinterpret_cast<void**>(&geometric), frame_bytes()), "hipMalloc(warp coverage)");
    gh::launch_fill(geometric, npix, 0.0f, stream);
    warp_coverage_frames = 0;
  }

  void ensure_spares(size_t n) {
    while (spare.size() < n) {
      float* p = nullptr;
      check(hipMalloc(reinterpret_cast<void**>(&p), frame_bytes()), "hipMalloc(warp spare slot)");
      spare.push_back(p);
      spare_allocations.push_back(p);
    }
  }

  void upload_master(const float* src, float** dst, const char* what) {
    if (src == nullptr) {
      if (*dst) (void)hipFree(*dst);
      *dst = nullptr;
      return;
    }
    if (*dst == nullptr) check(hipMalloc(reinterpret_cast<void**>(dst), frame_bytes()), what);
    check(hipMemcpyAsync(*dst, src, frame_bytes(), is_device_pointer(src) ? hipMemcpyDeviceToDevice
                                                                            : hipMemcpyHostToDevice,
                         stream),
          what);
  }

  ~gh_stack() {
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipFree(pool);
    for (float* p : spare_allocations) (void)hipFree(p);
    (void)hipFree(nonfinite);
    (void)hipFree(bias);
    (void)hipFree(dark);
    (void)hipFree(flat);
    (void)hipFree(geometric);
    partials.release();
    calib_stats.release();
    raw_staging.release();
    frame_table.release();
    frame_weights.release();
    tile_flags.release();
    for (auto& b : out_f32) b.release();
    out_dq.release();
    moments.release();
    radix.release();
    warp_flags.release();
    lanczos_lut.release();
    star_cand.release();
    star_windows.release();
    adj_ws.release();
    ln_sums.release();
    ln_counts.release();
    ln_coef.release();
    clip_tables.release();
    clip_geometric.release();
    clip_tile_map.release();
    cand_keys.release();
    cand_sorted.release();
    cand_selected.release();
    cand_u32.release();
    cand_count.release();
    cand_keep.release();
    cand_temp.release();
    cand_status.release();
    cand_xyf.release();
    align_inv.release();
    align_partials.release();
    align_frames.release();
    ncc_ws.release();
    if (copy_stream) (void)hipStreamSynchronize(copy_stream);
    for (int i = 0; i < kRawRing; ++i) {
      if (ring_copied[i]) (void)hipEventDestroy(ring_copied[i]);
      if (ring_free[i]) (void)hipEventDestroy(ring_free[i]);
    }
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (caller_done) (void)hipEventDestroy(caller_done);
    if (tables_done) (void)hipEventDestroy(tables_done);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace ghi {

// Device memory an integration launch reads besides the frames and outputs:
// the frame-pointer table, the frame weights and the per-tile statistics
// flags.  integrate_stack takes them from here when `scratch` is given
// (and then accepts device outputs only), else from the stack's own buffers,
// grown on demand.
struct IntegrationScratch {
  const float** frame_table;
  float* weights;
  size_t table_entries;
  unsigned char* tile_flags;
  size_t tile_entries;
};

void integrate_stack(gh_stack* s, const float* weights, const gh_integrate_params* params,
                     const gh_integrate_outputs* outputs, void* stream_ptr, const IntegrationScratch* scratch);

}  // namespace ghi

