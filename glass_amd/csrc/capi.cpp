// C ABI of libglass_hip.so (declared in include/glass_hip.h): stack
// lifecycle, calibration, warp, integration, statistics, star detection and
// registration refinement.  The stack state lives in stack.hpp.

#include "stack.hpp"
#include "stars.hpp"
#include "adjoint.hpp"

using ghi::guarded;
using ghi::g_last_error;

extern "C" {

int gh_last_error(char* buffer, size_t size) {
  if (buffer && size) {
    const size_t n = std::min(size - 1, g_last_error.size());
    std::memcpy(buffer, g_last_error.data(), n);
    buffer[n] = '\0';
  }
  return static_cast<int>(g_last_error.size());
}

int gh_version(void) { return 1; }

int gh_device_count(int* count) {
  return guarded([&] {
    int n = 0;
    check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    *count = n;
  });
}

int gh_device_info(int device, gh_device_properties* out) {
  return guarded([&] {
    if (!out) throw gh::ValueError("device properties pointer is NULL");
    int n = 0;
    check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw gh::IndexError("HIP device id is out of range");
    hipDeviceProp_t props{};
    check(hipGetDeviceProperties(&props, device), "hipGetDeviceProperties");
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, props.name, sizeof(out->name) - 1);
    std::strncpy(out->arch, props.gcnArchName, sizeof(out->arch) - 1);
    out->memory_total_bytes = static_cast<int64_t>(props.totalGlobalMem);
    out->multiprocessor_count = props.multiProcessorCount;
    out->major = props.major;
    out->minor = props.minor;
    out->wavefront_size = props.warpSize;
  });
}

int gh_host_alloc(size_t bytes, void** out) {
  return guarded([&] {
    if (!out) throw gh::ValueError("output pointer is NULL");
    *out = nullptr;
    if (bytes == 0) throw gh::ValueError("pinned host allocation must be non-empty");
    check(hipHostMalloc(out, bytes, hipHostMallocPortable), "hipHostMalloc");
  });
}

int gh_host_free(void* ptr) {
  return guarded([&] {
    if (ptr) check(hipHostFree(ptr), "hipHostFree");
  });
}

int gh_stack_create(int64_t frame_count, int64_t height, int64_t width, gh_stack** out) {
  return guarded([&] {
    if (!out) throw gh::ValueError("output handle pointer is NULL");
    *out = nullptr;
    if (frame_count <= 0 || height <= 0 || width <= 0)
      throw gh::ValueError("resident stack dimensions must be non-empty");
    if (height > (1 << 30) || width > (1 << 30)) throw gh::ValueError("resident stack dimensions are too large");
    // Kernels address a frame with 32-bit byte offsets.
    if (height * width >= (int64_t(1) << 30)) throw gh::ValueError("frames of 2^30 pixels or more are not supported");
    auto* s = new gh_stack();
    try {
      s->frames = frame_count;
      s->height = height;
      s->width = width;
      s->npix = height * width;
      check(hipStreamCreate(&s->stream), "hipStreamCreate(resident stack)");
      // Created here, not on first use, so recording them never creates a
      // runtime object inside a call (the no-allocation entry point records
      // caller_done).
      check(hipEventCreateWithFlags(&s->caller_done, hipEventDisableTiming), "hipEventCreate(caller stream)");
      check(hipEventCreateWithFlags(&s->tables_done, hipEventDisableTiming), "hipEventCreate(integration tables)");
      check(hipMalloc(reinterpret_cast<void**>(&s->pool), static_cast<size_t>(frame_count) * s->frame_bytes()),
            "hipMalloc(resident calibrated stack)");
      check(hipMalloc(reinterpret_cast<void**>(&s->nonfinite), static_cast<size_t>(frame_count) * sizeof(unsigned)),
            "hipMalloc(nonfinite flags)");
      check(hipMemsetAsync(s->nonfinite, 0, static_cast<size_t>(frame_count) * sizeof(unsigned), s->stream),
            "hipMemset(nonfinite flags)");
      s->slot.resize(frame_count);
      for (int64_t i = 0; i < frame_count; ++i) s->slot[i] = s->pool + i * s->npix;
      s->loaded.assign(frame_count, 0);
      s->calib_stats_valid.assign(frame_count, 0);
      check(hipStreamSynchronize(s->stream), "resident stack init");
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

int gh_stack_destroy(gh_stack* stack) {
  return guarded([&] { delete stack; });
}

int gh_stack_info(const gh_stack* s, int64_t out[8]) {
  return guarded([&] {
    if (!s) throw gh::ValueError("stack handle is NULL");
    out[0] = s->frames;
    out[1] = s->height;
    out[2] = s->width;
    out[3] = s->npix;
    out[4] = s->loaded_count;
    out[5] = s->bytes_allocated();
    out[6] = s->warp_coverage_frames;
    out[7] = static_cast<int64_t>(s->spare_allocations.size()) * static_cast<int64_t>(s->frame_bytes());
  });
}

int gh_stack_frame_ptr(gh_stack* s, int64_t index, float** device_ptr) {
  return guarded([&] {
    s->require_index(index);
    *device_ptr = s->slot[index];
  });
}

int gh_stack_frame_nonfinite(gh_stack* s, int64_t index, int64_t* count) {
  return guarded([&] {
    s->require_index(index);
    unsigned flag = 0;
    check(hipStreamSynchronize(s->stream), "synchronize(nonfinite flag)");
    check(hipMemcpy(&flag, s->nonfinite + index, sizeof(unsigned), hipMemcpyDeviceToHost), "hipMemcpy(nonfinite)");
    *count = flag;
  });
}

int gh_set_masters(gh_stack* s, const float* bias, const float* dark, const float* flat) {
  return guarded([&] {
    s->upload_master(bias, &s->bias, "upload bias master");
    s->upload_master(dark, &s->dark, "upload dark master");
    s->upload_master(flat, &s->flat, "upload flat master");
    check(hipStreamSynchronize(s->stream), "set_calibration_masters");
  });
}

int gh_upload_frame(gh_stack* s, int64_t index, const float* frame) {
  return guarded([&] {
    s->require_index(index);
    if (!frame) throw gh::ValueError("frame pointer is NULL");
    check(hipMemcpyAsync(s->slot[index], frame, s->frame_bytes(),
                         is_device_pointer(frame) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s->stream),
          "hipMemcpy(resident calibrated frame)");
    check(hipMemsetAsync(s->nonfinite + index, 0, sizeof(unsigned), s->stream), "reset nonfinite flag");
    gh::launch_flag_nonfinite(s->slot[index], s->npix, s->nonfinite + index, s->stream);
    s->calib_stats_valid[index] = 0;
    s->mark_loaded(index);
    check(hipStreamSynchronize(s->stream), "upload_calibrated_frame");
  });
}

int gh_download_tile(gh_stack* s, const int64_t* indices, int64_t count, int64_t x0, int64_t y0, int64_t x1,
                     int64_t y1, float* out) {
  return guarded([&] {
    if (count <= 0) throw gh::ValueError("indices must contain at least one frame index");
    if (x0 < 0 || y0 < 0 || x0 >= x1 || y0 >= y1 || x1 > s->width || y1 > s->height)
      throw gh::IndexError("resident frame tile bounds are invalid");
    const size_t tw = static_cast<size_t>(x1 - x0), th = static_cast<size_t>(y1 - y0);
    check(hipStreamSynchronize(s->stream), "synchronize(download)");
    for (int64_t k = 0; k < count; ++k) {
      s->require_loaded(indices[k], "resident frame tile download");
      const float* src = s->slot[indices[k]] + y0 * s->width + x0;
      check(hipMemcpy2D(out + k * tw * th, tw * sizeof(float), src, static_cast<size_t>(s->width) * sizeof(float),
                        tw * sizeof(float), th, hipMemcpyDeviceToHost),
            "hipMemcpy2D(resident frame tile)");
    }
  });
}

int gh_calibrate_frames(gh_stack* s, const int64_t* indices, int64_t count, const void* raw,
                        int64_t raw_frame_stride_bytes, int32_t format, const float* dark_scales,
                        const gh_calib_policy* policy, void* stream_ptr) {
  return guarded([&] {
    if (count <= 0) return;
    if (!raw) throw gh::ValueError("raw frame pointer is NULL");
    if (format != GH_FMT_F32 && format != GH_FMT_U16 && format != GH_FMT_FITS_U16BE_BZERO)
      throw gh::ValueError("unsupported raw sample format");
    const int64_t sample_bytes = (format == GH_FMT_F32) ? 4 : 2;
    const int64_t frame_raw_bytes = s->npix * sample_bytes;
    if (raw_frame_stride_bytes < frame_raw_bytes) throw gh::ValueError("raw frame stride is smaller than a frame");
    for (int64_t k = 0; k < count; ++k) s->require_index(indices[k]);
    gh_calib_policy pol{1, 1.0e-6f, 0.0f, 0};
    if (policy) pol = *policy;
    hipStream_t st = s->pick(stream_ptr);
    const bool device_raw = is_device_pointer(raw);
    gh::CalibParams p{};
    p.bias = s->bias;
    p.dark = s->dark;
    p.flat = s->flat;
    p.has_bias = s->bias != nullptr;
    p.has_dark = s->dark != nullptr;
    p.has_flat = s->flat != nullptr;
    p.dark_includes_bias = pol.master_dark_includes_bias;
    p.flat_floor = pol.flat_floor;
    p.pedestal = pol.pedestal;
    p.format = format;
    p.accumulate_stats = pol.accumulate_stats;
    const bool aligned = (s->npix % 4 == 0) && (raw_frame_stride_bytes % 16 == 0) &&
                         (reinterpret_cast<uintptr_t>(raw) % 16 == 0);
    if (!aligned && s->npix % 4 == 0 && device_raw)
      throw gh::ValueError("device raw frames must be 16-byte aligned with a 16-byte multiple stride");
    const unsigned blocks = gh::calibrate_blocks_per_frame(s->npix);
    if (pol.accumulate_stats) {
      s->partials.ensure(static_cast<size_t>(gh::kCalibBatch) * blocks, "hipMalloc(calibration partials)");
      s->calib_stats.ensure(static_cast<size_t>(s->frames), "hipMalloc(calibration stats)");
    }
    // Host raw: one frame per launch from the upload ring (the copy of the
    // next frame runs on the copy stream meanwhile).  Device raw: batches.
    const int64_t batch_cap = device_raw ? gh::kCalibBatch : 1;
    int slot = -1;
    if (st == s->stream) s->wait_caller();
    for (int64_t start = 0; start < count; start += batch_cap) {
      const int n = static_cast<int>(std::min<int64_t>(batch_cap, count - start));
      gh::CalibBatch b{};
      for (int j = 0; j < n; ++j) {
        const int64_t k = start + j;
        const int64_t f = indices[k];
        const unsigned char* src = static_cast<const unsigned char*>(raw) + k * raw_frame_stride_bytes;
        if (!device_raw) {
          unsigned char* dev = s->ring_slot(static_cast<size_t>(frame_raw_bytes), slot);
          check(hipStreamWaitEvent(s->copy_stream, s->ring_free[slot], 0), "hipStreamWaitEvent(raw ring)");
          check(hipMemcpyAsync(dev, src, frame_raw_bytes, hipMemcpyHostToDevice, s->copy_stream),
                "hipMemcpy(raw light)");
          check(hipEventRecord(s->ring_copied[slot], s->copy_stream), "hipEventRecord(raw ring)");
          check(hipStreamWaitEvent(st, s->ring_copied[slot], 0), "hipStreamWaitEvent(raw ring)");
          src = dev;
        }
        b.src[j] = src;
        b.dst[j] = s->slot[f];
        b.dark_scale[j] = dark_scales ? dark_scales[k] : 1.0f;
        b.nonfinite[j] = s->nonfinite + f;
        b.frame[j] = static_cast<int>(f);
      }
      gh::launch_clear_flags(s->nonfinite, b, n, st);
      gh::launch_calibrate(b, n, p, s->npix, s->partials.ptr, st);
      if (pol.accumulate_stats) gh::launch_stats_reduce(s->partials.ptr, b, n, blocks, s->calib_stats.ptr, st);
      for (int j = 0; j < n; ++j) {
        const int64_t f = indices[start + j];
        s->calib_stats_valid[f] = pol.accumulate_stats ? 1 : 0;
        s->mark_loaded(f);
      }
      if (!device_raw) check(hipEventRecord(s->ring_free[slot], st), "hipEventRecord(raw ring)");
    }
    s->mark_caller(st);
    if (!stream_ptr) check(hipStreamSynchronize(st), "calibrate");
  });
}

int gh_normalize_frames(gh_stack* s, const int64_t* indices, int64_t count, const float* offsets,
                        const float* divisors, void* stream_ptr) {
  return guarded([&] {
    if (count <= 0) return;
    if (!offsets || !divisors) throw gh::ValueError("offsets and divisors must not be NULL");
    hipStream_t st = s->pick(stream_ptr);
    for (int64_t k = 0; k < count; ++k) {
      const int64_t f = indices[k];
      s->require_loaded(f, "frame normalisation");
      // The frame keeps its non-finite flag; normalisation can only add to it.
      gh::launch_normalize(s->slot[f], s->npix, offsets[k], divisors[k], s->nonfinite + f, st);
      s->calib_stats_valid[f] = 0;
    }
    if (!stream_ptr) check(hipStreamSynchronize(st), "normalize_frames");
  });
}

int gh_warp_frames(gh_stack* s, const int64_t* indices, int64_t count, const double* inverse, int32_t interpolation,
                   float fill, float clamping_threshold, uint32_t flags, void* stream_ptr) {
  return guarded([&] {
    if (count <= 0) return;
    if (interpolation != GH_INTERP_BILINEAR && interpolation != GH_INTERP_LANCZOS3)
      throw gh::ValueError("unsupported warp interpolation");
    if (!inverse) throw gh::ValueError("inverse matrices pointer is NULL");
    for (int64_t k = 0; k < count; ++k) {
      s->require_loaded(indices[k], "resident matrix warp");
      for (int64_t j = 0; j < k; ++j)
        if (indices[j] == indices[k]) throw gh::ValueError("warp indices must be unique");
      for (int e = 0; e < 9; ++e)
        if (!std::isfinite(inverse[k * 9 + e])) throw gh::ValueError("warp inverse matrix must be finite");
    }
    hipStream_t st = s->pick(stream_ptr);
    gh::WarpParams p{};
    p.height = static_cast<int>(s->height);
    p.width = static_cast<int>(s->width);
    p.interpolation = interpolation;
    p.fill = fill;
    p.clamp = clamping_threshold;
    p.skip_nonfinite = (flags & GH_WARP_SKIP_NONFINITE_TAPS) ? 1 : 0;
    p.affine = (flags & GH_WARP_AFFINE) ? 1 : 0;
    if (p.affine) {
      for (int64_t k = 0; k < count; ++k) {
        const double* m = inverse + k * 9;
        if (!(m[6] == 0.0 && m[7] == 0.0 && m[8] == 1.0))
          throw gh::ValueError("GH_WARP_AFFINE requires an inverse whose last row is (0, 0, 1)");
      }
    }
    if (flags & GH_WARP_TRACK_COVERAGE) {
      s->ensure_geometric();
      p.coverage = s->geometric;
    }
    if (interpolation == GH_INTERP_LANCZOS3 && p.affine) {
      constexpr size_t kLutFloats = (gh::kLanczosLutIntervals + 1) * gh::kLanczosLutStride;
      if (s->lanczos_lut.ptr == nullptr) {
        std::vector<float> lut(kLutFloats);
        gh::make_lanczos_lut(lut.data());
        s->lanczos_lut.ensure(kLutFloats, "hipMalloc(lanczos table)");
        check(hipMemcpy(s->lanczos_lut.ptr, lut.data(), kLutFloats * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy(lanczos table)");
      }
      p.lanczos_lut = reinterpret_cast<const float4*>(s->lanczos_lut.ptr);
    }
    s->warp_flags.ensure(gh::kWarpBatch, "hipMalloc(warp flags)");
    s->ensure_spares(static_cast<size_t>(std::min<int64_t>(count, gh::kWarpBatch)));
    for (int64_t start = 0; start < count; start += gh::kWarpBatch) {
      const int n = static_cast<int>(std::min<int64_t>(gh::kWarpBatch, count - start));
      gh::WarpBatch b{};
      for (int j = 0; j < n; ++j) {
        const int64_t f = indices[start + j];
        b.src[j] = s->slot[f];
        b.dst[j] = s->spare[j];
        std::memcpy(b.inv[j], inverse + (start + j) * 9, 9 * sizeof(double));
        b.nonfinite[j] = s->nonfinite + f;
      }
      b.src_flags = s->warp_flags.ptr;
      gh::launch_snapshot_warp_flags(b, n, s->warp_flags.ptr, st);
      gh::launch_warp(b, n, p, st);
      for (int j = 0; j < n; ++j) {
        const int64_t f = indices[start + j];
        std::swap(s->slot[f], s->spare[j]);
        s->calib_stats_valid[f] = 0;
      }
      if (flags & GH_WARP_TRACK_COVERAGE) s->warp_coverage_frames += n;
    }
    s->mark_caller(st);
  });
}

int gh_translate_frame(gh_stack* s, int64_t index, int32_t mode, double dx, double dy, float fill,
                       void* stream_ptr) {
  return guarded([&] {
    s->require_loaded(index, "translation warp");
    if (mode != GH_TRANSLATE_INTEGER && mode != GH_INTERP_BILINEAR)
      throw gh::ValueError("translation mode must be integer or bilinear");
    if (!std::isfinite(dx) || !std::isfinite(dy)) throw gh::ValueError("translation must be finite");
    int idx = 0, idy = 0;
    if (mode == GH_TRANSLATE_INTEGER) {
      if (dx != std::floor(dx) || dy != std::floor(dy) || std::fabs(dx) > 2.0e9 || std::fabs(dy) > 2.0e9)
        throw gh::ValueError("integer translation needs whole-pixel offsets");
      idx = static_cast<int>(dx);
      idy = static_cast<int>(dy);
    }
    hipStream_t st = s->pick(stream_ptr);
    if (st == s->stream) s->wait_caller();
    s->ensure_geometric();
    s->ensure_spares(1);
    unsigned* flag = s->nonfinite + index;
    gh::launch_fill(reinterpret_cast<float*>(flag), 1, 0.0f, st);  // all-zero bits: flag cleared
    gh::launch_translate(s->slot[index], s->spare[0], static_cast<int>(s->width), static_cast<int>(s->height),
                         mode == GH_INTERP_BILINEAR, idx, idy, static_cast<float>(dx), static_cast<float>(dy), fill,
                         s->geometric, flag, st);
    std::swap(s->slot[index], s->spare[0]);
    s->calib_stats_valid[index] = 0;
    s->warp_coverage_frames += 1;
    s->mark_caller(st);
    if (!stream_ptr) check(hipStreamSynchronize(st), "translation warp");
  });
}

int gh_reset_warp_coverage(gh_stack* s) {
  return guarded([&] {
    s->ensure_geometric();
    s->wait_caller();
    gh::launch_fill(s->geometric, s->npix, 0.0f, s->stream);
    s->warp_coverage_frames = 0;
    check(hipStreamSynchronize(s->stream), "reset_warp_coverage");
  });
}

int gh_accumulate_full_warp_coverage(gh_stack* s) {
  return guarded([&] {
    s->ensure_geometric();
    s->wait_caller();
    gh::launch_add_scalar(s->geometric, s->npix, 1.0f, s->stream);
    ++s->warp_coverage_frames;
    check(hipStreamSynchronize(s->stream), "accumulate_full_warp_coverage_frame");
  });
}

int gh_warp_coverage_map(gh_stack* s, float* out) {
  return guarded([&] {
    if (!out) throw gh::ValueError("output pointer is NULL");
    s->wait_caller();
    check(hipStreamSynchronize(s->stream), "synchronize(warp coverage)");
    if (!s->geometric) {
      if (is_device_pointer(out)) {
        gh::launch_fill(out, s->npix, 0.0f, s->stream);
        check(hipStreamSynchronize(s->stream), "warp_coverage_map");
      } else {
        std::fill(out, out + s->npix, 0.0f);
      }
      return;
    }
    check(hipMemcpy(out, s->geometric, s->frame_bytes(), is_device_pointer(out) ? hipMemcpyDeviceToDevice
                                                                                 : hipMemcpyDeviceToHost),
          "hipMemcpy(warp coverage map)");
  });
}

}  // extern "C"

namespace ghi {

void integrate_stack(gh_stack* s, const float* weights, const gh_integrate_params* params,
                     const gh_integrate_outputs* outputs, void* stream_ptr, const IntegrationScratch* scratch) {
    if (!s) throw gh::ValueError("stack handle is NULL");
    if (!params || !outputs) throw gh::ValueError("params and outputs must not be NULL");
    if (s->loaded_count != s->frames)
      throw std::runtime_error("all resident frames must be loaded before integration");
    const gh_integrate_params& prm = *params;
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
    hipStream_t st = s->pick(stream_ptr);

    // Active frames: finite, positive weights (cpu/integration.py:67), frame order.
    std::vector<const float*> table;
    std::vector<float> w;
    bool unit = true;
    for (int64_t f = 0; f < s->frames; ++f) {
      const float wf = weights ? weights[f] : 1.0f;
      if (!(std::isfinite(wf) && wf > 0.0f)) continue;
      table.push_back(s->slot[f]);
      w.push_back(wf);
      unit = unit && (wf == 1.0f);
    }
    const int n_active = static_cast<int>(table.size());
    if (prm.rejection == GH_REJECT_WINSORIZED_SIGMA && n_active > gh::integrate_max_frames())
      throw gh::ValueError("winsorized integration supports at most 16384 positive-weight frames");
    if (s->npix >= (int64_t(1) << 30)) throw gh::ValueError("frames of 2^30 pixels or more are not supported");
    // The winsorized kernel reads table entries [0, capacity) unconditionally
    // (compile-time indices); pad to the largest capacity with a valid frame
    // (the last active one, or frame 0 when none is active) and weight 0.
    const size_t padded = static_cast<size_t>(std::max(gh::integrate_register_capacity(), n_active));
    const float* pad_ptr = n_active ? table.back() : s->slot[0];
    table.resize(padded, pad_ptr);
    w.resize(padded, 0.0f);
    const float** d_table;
    float* d_weights;
    if (scratch) {
      if (padded > scratch->table_entries)
        throw gh::ValueError("frame table exceeds the no-allocation scratch");
      // Every call on this device shares the scratch: the last launch that
      // read it on this stream must have finished before it is rewritten.
      check(hipStreamSynchronize(st), "synchronize(integration scratch)");
      d_table = scratch->frame_table;
      d_weights = scratch->weights;
    } else {
      s->wait_tables();
      s->frame_table.ensure(padded, "hipMalloc(frame table)");
      s->frame_weights.ensure(padded, "hipMalloc(frame weights)");
      d_table = s->frame_table.ptr;
      d_weights = s->frame_weights.ptr;
    }
    check(hipMemcpy(d_table, table.data(), padded * sizeof(const float*), hipMemcpyHostToDevice),
          "hipMemcpy(frame table)");
    check(hipMemcpy(d_weights, w.data(), padded * sizeof(float), hipMemcpyHostToDevice),
          "hipMemcpy(frame weights)");

    // Statistics arithmetic (rejection.py:39-44): all-valid only when every
    // weight is positive and every sample of every frame is finite.
    bool any_nonfinite = false;
    {
      std::vector<unsigned> flags(static_cast<size_t>(s->frames));
      check(hipStreamSynchronize(st), "synchronize(nonfinite flags)");
      check(hipMemcpy(flags.data(), s->nonfinite, flags.size() * sizeof(unsigned), hipMemcpyDeviceToHost),
            "hipMemcpy(nonfinite flags)");
      for (int64_t f = 0; f < s->frames; ++f) {
        const float wf = weights ? weights[f] : 1.0f;
        if (std::isfinite(wf) && wf > 0.0f && flags[f]) any_nonfinite = true;
      }
    }
    gh::IntegrateArgs a{};
    a.frame_ptr = d_table;
    a.frame_w = d_weights;
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
    const bool all_weights_positive = (n_active == s->frames);
    if (prm.statistics_variant >= 0) {
      a.variant_all_valid = prm.statistics_variant;
    } else if (prm.tile_size > 0 && all_weights_positive && any_nonfinite) {
      const int tiles_x = static_cast<int>((s->width + prm.tile_size - 1) / prm.tile_size);
      const int tiles_y = static_cast<int>((s->height + prm.tile_size - 1) / prm.tile_size);
      const size_t tiles = static_cast<size_t>(tiles_x) * tiles_y;
      unsigned char* d_tiles;
      if (scratch) {
        if (tiles > scratch->tile_entries) throw gh::ValueError("tile grid exceeds the no-allocation scratch");
        d_tiles = scratch->tile_flags;
      } else {
        s->tile_flags.ensure(tiles, "hipMalloc(tile flags)");
        d_tiles = s->tile_flags.ptr;
      }
      gh::launch_tile_all_finite(d_table, n_active, static_cast<int>(s->height),
                                 static_cast<int>(s->width), prm.tile_size, d_tiles, st);
      a.tile_flags = d_tiles;
      a.tile_size = prm.tile_size;
      a.tiles_x = tiles_x;
    } else {
      a.variant_all_valid = (all_weights_positive && !any_nonfinite) ? 1 : 0;
    }

    // Outputs: device pointers are written in place; host pointers go through
    // device temporaries and are copied back.
    const gh_integrate_outputs& o = *outputs;
    const size_t count_bytes = (prm.count_dtype == GH_COUNT_U16) ? 2 : 4;
    void* host_dst[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t host_bytes[6] = {0, 0, 0, 0, 0, 0};
    auto route = [&](void* user, int slot, size_t elem_bytes) -> void* {
      if (!user) return nullptr;
      if (is_device_pointer(user)) return user;
      if (scratch) throw gh::ValueError("the no-allocation integration writes device outputs only");
      host_dst[slot] = user;
      host_bytes[slot] = elem_bytes * static_cast<size_t>(s->npix);
      if (slot < 5) {
        s->out_f32[slot].ensure(static_cast<size_t>(s->npix), "hipMalloc(integration output)");
        return s->out_f32[slot].ptr;
      }
      s->out_dq.ensure(static_cast<size_t>(s->npix), "hipMalloc(dq output)");
      return s->out_dq.ptr;
    };
    a.master = static_cast<float*>(route(o.master, 0, 4));
    a.weight = static_cast<float*>(route(o.weight, 1, 4));
    a.coverage = route(o.coverage, 2, count_bytes);
    a.low = route(o.low, 3, count_bytes);
    a.high = route(o.high, 4, count_bytes);
    a.dq = static_cast<unsigned*>(route(o.dq, 5, 4));
    a.count_u16 = (prm.count_dtype == GH_COUNT_U16) ? 1 : 0;
    if (o.dq && o.dq_use_geometric) {
      if (!s->geometric) throw gh::ValueError("DQ with geometric coverage requested but no warp coverage exists");
      a.geometric = s->geometric;
    }
    a.dq_active_frames = o.dq_active_frames;
    gh::launch_integrate(a, st);
    if (!scratch) s->mark_tables(st);
    const void* dev_src[6] = {a.master, a.weight, a.coverage, a.low, a.high, a.dq};
    bool any_host = false;
    for (int i = 0; i < 6; ++i) {
      if (!host_dst[i]) continue;
      any_host = true;
      check(hipMemcpyAsync(host_dst[i], dev_src[i], host_bytes[i], hipMemcpyDeviceToHost, st),
            "hipMemcpy(integration output)");
    }
    s->mark_caller(st);
    if (any_host || !stream_ptr) check(hipStreamSynchronize(st), "integrate synchronize");
}

}  // namespace ghi

extern "C" {

int gh_integrate(gh_stack* s, const float* weights, const gh_integrate_params* params,
                 const gh_integrate_outputs* outputs, void* stream_ptr) {
  return guarded([&] { ghi::integrate_stack(s, weights, params, outputs, stream_ptr, nullptr); });
}

int gh_frame_stats(gh_stack* s, int64_t index, int32_t exact_median, gh_frame_statistics* out) {
  return guarded([&] {
    s->require_loaded(index, "global frame statistics");
    if (!out) throw gh::ValueError("output pointer is NULL");
    const float* frame = s->slot[index];
    const unsigned blocks = gh::frame_moments_blocks(s->npix);
    s->moments.ensure(blocks + 1, "hipMalloc(frame moments)");
    gh::launch_frame_moments(frame, s->npix, s->moments.ptr + 1, blocks, s->moments.ptr, s->stream);
    gh::FrameMoments m{};
    check(hipMemcpyAsync(&m, s->moments.ptr, sizeof(m), hipMemcpyDeviceToHost, s->stream), "hipMemcpy(moments)");
    check(hipStreamSynchronize(s->stream), "frame stats");
    const int64_t n = static_cast<int64_t>(m.finite);
    ghi::fill_frame_statistics(m, s->npix, out);
    if (!exact_median || n == 0) return;
    s->radix.ensure(256 + 4, "hipMalloc(radix select)");
    auto select = [&](int abs_mode, float center, uint64_t rank) -> float {
      gh::launch_radix_select(frame, s->npix, abs_mode, center, rank, s->radix.ptr, s->radix.ptr + 256, s->stream);
      unsigned key = 0;
      check(hipMemcpyAsync(&key, s->radix.ptr + 256, sizeof(unsigned), hipMemcpyDeviceToHost, s->stream),
            "hipMemcpy(radix key)");
      check(hipStreamSynchronize(s->stream), "radix select");
      const unsigned bits = (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
      float v;
      std::memcpy(&v, &bits, sizeof(v));
      return v;
    };
    // quality.py:170-180: partition median, even counts average the middle pair in FP64.
    const uint64_t mid = static_cast<uint64_t>(n / 2);
    double median = select(0, 0.0f, mid);
    if (n % 2 == 0) median = (static_cast<double>(select(0, 0.0f, mid - 1)) + median) / 2.0;
    const float center = static_cast<float>(median);
    double mad = select(1, center, mid);
    if (n % 2 == 0) mad = (static_cast<double>(select(1, center, mid - 1)) + mad) / 2.0;
    out->median = median;
    out->mad = mad;
  });
}

int gh_calibration_stats(gh_stack* s, int64_t index, double* sum, double* sum_sq, int64_t* finite) {
  return guarded([&] {
    s->require_index(index);
    if (!s->calib_stats_valid[index] || !s->calib_stats.ptr)
      throw gh::ValueError("no fused calibration statistics for this frame");
    gh::StatPartial r{};
    check(hipStreamSynchronize(s->stream), "synchronize(calibration stats)");
    check(hipMemcpy(&r, s->calib_stats.ptr + index, sizeof(r), hipMemcpyDeviceToHost), "hipMemcpy(calib stats)");
    *sum = r.sum;
    *sum_sq = r.sum_sq;
    *finite = static_cast<int64_t>(r.finite);
  });
}

int gh_synchronize(gh_stack* s) {
  return guarded([&] { check(hipStreamSynchronize(s->stream), "synchronize"); });
}

int gh_detect_stars(gh_stack* s, int64_t index, double threshold_sigma, int64_t max_stars, double background,
                    double noise, int32_t window_radius, double min_separation_px, int64_t origin_x,
                    int64_t origin_y, gh_star* out, int64_t capacity, int64_t* count, double* background_out) {
  return guarded([&] {
    s->require_loaded(index, "star detection");
    if (!count || (capacity > 0 && !out)) throw gh::ValueError("output pointer is NULL");
    if (window_radius > gh::kMaxStarRadius) throw gh::ValueError("window_radius must be at most 8");
    const int radius = std::max(1, static_cast<int>(window_radius));
    const float* frame = s->slot[index];
    const int h = static_cast<int>(s->height), w = static_cast<int>(s->width);
    *count = 0;
    // robust_background (star_detect.py:28-42): numpy float32 medians.
    const unsigned blocks = gh::frame_moments_blocks(s->npix);
    s->moments.ensure(blocks + 1, "hipMalloc(frame moments)");
    gh::launch_frame_moments(frame, s->npix, s->moments.ptr + 1, blocks, s->moments.ptr, s->stream);
    gh::FrameMoments m{};
    check(hipMemcpyAsync(&m, s->moments.ptr, sizeof(m), hipMemcpyDeviceToHost, s->stream), "hipMemcpy(moments)");
    check(hipStreamSynchronize(s->stream), "star background");
    const uint64_t n = m.finite;
    double median = 0.0, mad = 0.0, sigma_bg = 1.0;
    if (n > 0) {
      s->radix.ensure(256 + 4, "hipMalloc(radix select)");
      auto select = [&](int abs_mode, float center, uint64_t rank) -> float {
        gh::launch_radix_select(frame, s->npix, abs_mode, center, rank, s->radix.ptr, s->radix.ptr + 256, s->stream);
        unsigned key = 0;
        check(hipMemcpyAsync(&key, s->radix.ptr + 256, sizeof(unsigned), hipMemcpyDeviceToHost, s->stream),
              "hipMemcpy(radix key)");
        check(hipStreamSynchronize(s->stream), "radix select");
        const unsigned bits = (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
        float v;
        std::memcpy(&v, &bits, sizeof(v));
        return v;
      };
      // np.median of float32: the middle pair is averaged in float32.
      auto median32 = [&](int abs_mode, float center) {
        const float hi = select(abs_mode, center, n / 2);
        if (n % 2) return hi;
        const float lo = select(abs_mode, center, n / 2 - 1);
        return (lo + hi) / 2.0f;
      };
      const float med = median32(0, 0.0f);
      median = med;
      mad = median32(1, med);
      sigma_bg = 1.4826 * mad;
      if (sigma_bg <= 0.0) {
        // np.std of the finite float32 samples, in numpy's float32 pairwise order.
        std::vector<float> host(static_cast<size_t>(s->npix));
        check(hipMemcpy(host.data(), frame, s->frame_bytes(), hipMemcpyDeviceToHost), "hipMemcpy(frame)");
        size_t k = 0;
        for (float v : host)
          if (std::isfinite(v)) host[k++] = v;
        sigma_bg = gh::numpy_std_f32(host.data(), k);
      }
      if (sigma_bg <= 0.0) sigma_bg = 1.0;
    }
    if (background_out) {
      background_out[0] = median;
      background_out[1] = mad;
      background_out[2] = sigma_bg;
    }
    const double bg = std::isnan(background) ? median : background;
    const double sigma = std::isnan(noise) ? sigma_bg : noise;
    const double threshold = bg + threshold_sigma * std::max(sigma, 1.0e-6);
    if (h < 3 || w < 3) return;
    // Candidates (star_detect.py:148-160); the FP64 threshold meets the float32
    // image as a float32 (numpy weak-scalar comparison).
    unsigned cap = 1u << 16;
    unsigned found = 0;
    for (int pass = 0; pass < 2; ++pass) {
      s->star_cand.ensure(static_cast<size_t>(cap) + 1, "hipMalloc(star candidates)");
      gh::launch_fill(reinterpret_cast<float*>(s->star_cand.ptr), 1, 0.0f, s->stream);
      gh::launch_star_candidates(frame, h, w, static_cast<float>(threshold), s->star_cand.ptr + 1, cap,
                                 s->star_cand.ptr, s->stream);
      check(hipMemcpyAsync(&found, s->star_cand.ptr, sizeof(unsigned), hipMemcpyDeviceToHost, s->stream),
            "hipMemcpy(candidate count)");
      check(hipStreamSynchronize(s->stream), "star candidates");
      if (found <= cap) break;
      cap = found;
    }
    std::vector<unsigned> cand(found);
    if (found) {
      check(hipMemcpy(cand.data(), s->star_cand.ptr + 1, found * sizeof(unsigned), hipMemcpyDeviceToHost),
            "hipMemcpy(candidates)");
    }
    std::sort(cand.begin(), cand.end());  // np.nonzero order: row-major
    const int side = 2 * radius + 1, cells = side * side;
    std::vector<float> windows(static_cast<size_t>(found) * cells);
    if (found) {
      check(hipMemcpy(s->star_cand.ptr + 1, cand.data(), found * sizeof(unsigned), hipMemcpyHostToDevice),
            "hipMemcpy(sorted candidates)");
      s->star_windows.ensure(windows.size(), "hipMalloc(star windows)");
      gh::launch_gather_windows(frame, h, w, s->star_cand.ptr + 1, found, radius, s->star_windows.ptr, s->stream);
      check(hipMemcpyAsync(windows.data(), s->star_windows.ptr, windows.size() * sizeof(float),
                           hipMemcpyDeviceToHost, s->stream),
            "hipMemcpy(star windows)");
      check(hipStreamSynchronize(s->stream), "star windows");
    }
    std::vector<gh::Star> stars;
    stars.reserve(found);
    for (unsigned k = 0; k < found; ++k) {
      gh::CandidatePatch c;
      c.x = static_cast<int>(cand[k] % static_cast<unsigned>(w));
      c.y = static_cast<int>(cand[k] / static_cast<unsigned>(w));
      c.values = windows.data() + static_cast<size_t>(k) * cells;
      gh::Star st;
      if (gh::measure_candidate(c, h, w, radius, bg, sigma, origin_x, origin_y, st)) stars.push_back(st);
    }
    stars = gh::finish_star_list(std::move(stars), min_separation_px, max_stars);
    const int64_t n_out = std::min<int64_t>(static_cast<int64_t>(stars.size()), std::max<int64_t>(capacity, 0));
    for (int64_t k = 0; k < n_out; ++k) {
      const gh::Star& st = stars[k];
      out[k] = gh_star{st.x, st.y, st.flux, st.peak, st.fwhm_px, st.eccentricity, st.snr, st.sharpness, st.npix,
                       st.has_fwhm ? 1 : 0, st.has_eccentricity ? 1 : 0};
    }
    *count = static_cast<int64_t>(stars.size());
  });
}

namespace {

gh::AffineMap affine_map_of(const double* map6) {
  if (!map6) throw gh::ValueError("map pointer is NULL");
  gh::AffineMap m{};
  for (int k = 0; k < 6; ++k) {
    if (!std::isfinite(map6[k])) throw gh::ValueError("affine map must be finite");
    m.b[k] = map6[k];
  }
  return m;
}

void check_image_dims(int64_t h, int64_t w) {
  if (h < 1 || w < 1 || h > (1 << 24) || w > (1 << 24) || h * w >= (int64_t{1} << 31))
    throw gh::ValueError("image dimensions must be positive and below 2^31 pixels");
}

// Inverse of a 3x3 affine matrix (last row 0 0 1), FP64.
void invert_affine3(const double* m, double* inv) {
  if (!(m[6] == 0.0 && m[7] == 0.0 && m[8] == 1.0)) throw gh::ValueError("matrix must be affine (last row 0 0 1)");
  for (int k = 0; k < 9; ++k)
    if (!std::isfinite(m[k])) throw gh::ValueError("matrix must be finite");
  const double det = m[0] * m[4] - m[1] * m[3];
  if (!(std::fabs(det) > 1.0e-20)) throw gh::ValueError("matrix is singular");
  inv[0] = m[4] / det;
  inv[1] = -m[1] / det;
  inv[3] = -m[3] / det;
  inv[4] = m[0] / det;
  inv[2] = -(inv[0] * m[2] + inv[1] * m[5]);
  inv[5] = -(inv[3] * m[2] + inv[4] * m[5]);
  inv[6] = 0.0;
  inv[7] = 0.0;
  inv[8] = 1.0;
}

// Solves the symmetric positive definite 6x6 system a x = b (Cholesky, FP64,
// Jacobi-scaled).  Returns false when the matrix is not positive definite.
bool solve_spd6(const double (&a_in)[6][6], const double (&b_in)[6], double (&x)[6]) {
  double d[6], a[6][6], b[6];
  for (int i = 0; i < 6; ++i) d[i] = a_in[i][i] > 0.0 ? std::sqrt(a_in[i][i]) : 1.0;
  for (int i = 0; i < 6; ++i) {
    b[i] = b_in[i] / d[i];
    for (int j = 0; j < 6; ++j) a[i][j] = a_in[i][j] / (d[i] * d[j]);
  }
  double l[6][6] = {};
  for (int i = 0; i < 6; ++i) {
    for (int j = 0; j <= i; ++j) {
      double s = a[i][j];
      for (int k = 0; k < j; ++k) s -= l[i][k] * l[j][k];
      if (i == j) {
        if (!(s > 1.0e-14)) return false;
        l[i][i] = std::sqrt(s);
      } else {
        l[i][j] = s / l[j][j];
      }
    }
  }
  double y[6];
  for (int i = 0; i < 6; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= l[i][k] * y[k];
    y[i] = s / l[i][i];
  }
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < 6; ++k) s -= l[k][i] * x[k];
    x[i] = s / l[i][i];
  }
  for (int i = 0; i < 6; ++i) x[i] /= d[i];
  return true;
}

}  // namespace

int gh_affine_workspace_size(int64_t height, int64_t width, int64_t* doubles) {
  return guarded([&] {
    check_image_dims(height, width);
    if (!doubles) throw gh::ValueError("output pointer is NULL");
    *doubles = static_cast<int64_t>(gh::adjoint_workspace_doubles(height * width));
  });
}

int gh_affine_lanczos3_forward(const float* src, int64_t height, int64_t width, const double* map6, float fill,
                               float* out, void* stream) {
  return guarded([&] {
    check_image_dims(height, width);
    if (!src || !out) throw gh::ValueError("image pointer is NULL");
    if (!is_device_pointer(src) || !is_device_pointer(out)) throw gh::ValueError("images must be device memory");
    gh::launch_affine_forward(src, static_cast<int>(height), static_cast<int>(width), affine_map_of(map6), fill, out,
                              static_cast<hipStream_t>(stream));
  });
}

int gh_affine_lanczos3_backward(const float* src, int64_t height, int64_t width, const double* map6,
                                const float* grad_out, float* grad_src, double* grad_map6, double* workspace,
                                int64_t workspace_doubles, void* stream) {
  return guarded([&] {
    check_image_dims(height, width);
    if (!src || !grad_out) throw gh::ValueError("image pointer is NULL");
    if (!is_device_pointer(src) || !is_device_pointer(grad_out) || (grad_src && !is_device_pointer(grad_src)))
      throw gh::ValueError("images must be device memory");
    if (grad_map6) {
      if (!is_device_pointer(grad_map6) || !workspace || !is_device_pointer(workspace))
        throw gh::ValueError("grad_map6 and workspace must be device memory");
      if (workspace_doubles < static_cast<int64_t>(gh::adjoint_workspace_doubles(height * width)))
        throw gh::ValueError("workspace is smaller than gh_affine_workspace_size");
    }
    gh::launch_affine_backward(src, static_cast<int>(height), static_cast<int>(width), affine_map_of(map6), grad_out,
                               grad_src, workspace, grad_map6, static_cast<hipStream_t>(stream));
  });
}

int gh_refine_affine(gh_stack* s, int64_t reference_index, int64_t moving_index, const double* matrix_in,
                     int32_t max_iterations, double tolerance_px, double* matrix_out, double* report) {
  return guarded([&] {
    s->require_loaded(reference_index, "affine refinement");
    s->require_loaded(moving_index, "affine refinement");
    if (!matrix_in || !matrix_out) throw gh::ValueError("matrix pointer is NULL");
    if (max_iterations < 0) throw gh::ValueError("max_iterations must be >= 0");
    const int h = static_cast<int>(s->height), w = static_cast<int>(s->width);
    double inv[9];
    invert_affine3(matrix_in, inv);  // reference pixel -> moving pixel
    const double cx = 0.5 * (w - 1), cy = 0.5 * (h - 1);
    // Centred parameters: s = Bc (x - cx, y - cy, 1).
    double bc[6] = {inv[0], inv[1], inv[0] * cx + inv[1] * cy + inv[2],
                    inv[3], inv[4], inv[3] * cx + inv[4] * cy + inv[5]};
    auto to_map = [&](const double (&c)[6]) {
      gh::AffineMap m{};
      m.b[0] = c[0];
      m.b[1] = c[1];
      m.b[2] = c[2] - c[0] * cx - c[1] * cy;
      m.b[3] = c[3];
      m.b[4] = c[4];
      m.b[5] = c[5] - c[3] * cx - c[4] * cy;
      return m;
    };
    const size_t ws = gh::adjoint_workspace_doubles(s->npix);
    s->adj_ws.ensure(ws, "hipMalloc(refinement workspace)");
    double* sums_dev = s->adj_ws.ptr + (ws - gh::kNormalSums);
    double sums[gh::kNormalSums];
    auto normal_sums = [&](const gh::AffineMap& m) {
      gh::launch_normal_equations(s->slot[reference_index], s->slot[moving_index], h, w, m, cx, cy, s->adj_ws.ptr,
                                  sums_dev, s->stream);
      check(hipMemcpyAsync(sums, sums_dev, sizeof(sums), hipMemcpyDeviceToHost, s->stream), "hipMemcpy(sums)");
      check(hipStreamSynchronize(s->stream), "normal equations");
    };
    int it = 0;
    double last_step = 0.0;
    for (; it < max_iterations; ++it) {
      normal_sums(to_map(bc));
      if (sums[28] < 6.0) throw std::runtime_error("affine refinement: too few overlapping pixels");
      double a[6][6], g[6], delta[6];
      int t = 0;
      for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) a[i][j] = a[j][i] = sums[t++];
      for (int i = 0; i < 6; ++i) g[i] = -sums[21 + i];
      if (!solve_spd6(a, g, delta)) throw std::runtime_error("affine refinement: singular normal equations");
      for (int i = 0; i < 6; ++i) bc[i] += delta[i];
      // largest displacement of an image corner under the update
      last_step = 0.0;
      for (int c = 0; c < 4; ++c) {
        const double xc = ((c & 1) ? w - 1 : 0) - cx, yc = ((c & 2) ? h - 1 : 0) - cy;
        const double ux = delta[0] * xc + delta[1] * yc + delta[2];
        const double uy = delta[3] * xc + delta[4] * yc + delta[5];
        last_step = std::max(last_step, std::sqrt(ux * ux + uy * uy));
      }
      if (last_step < tolerance_px) {
        ++it;
        break;
      }
    }
    const gh::AffineMap fin = to_map(bc);
    double out_inv[9] = {fin.b[0], fin.b[1], fin.b[2], fin.b[3], fin.b[4], fin.b[5], 0.0, 0.0, 1.0};
    invert_affine3(out_inv, matrix_out);
    if (report) {
      normal_sums(fin);
      report[0] = static_cast<double>(it);
      report[1] = sums[28] > 0.0 ? std::sqrt(sums[27] / sums[28]) : NAN;
      report[2] = sums[28];
      report[3] = last_step;
    }
  });
}

int gh_estimate_star_transform(const double* ref_xy, int64_t n_ref, const double* mov_xy, int64_t n_mov,
                               int32_t model, int64_t min_inliers, double tolerance_px, int64_t max_stars,
                               double* matrix, int64_t* pairs, int64_t* n_pairs, double* rms, int32_t* ok) {
  return guarded([&] {
    if ((n_ref > 0 && !ref_xy) || (n_mov > 0 && !mov_xy) || !matrix || !n_pairs || !rms || !ok)
      throw gh::ValueError("argument pointer is NULL");
    const gh::TransformResult r =
        gh::estimate_star_transform(ref_xy, n_ref, mov_xy, n_mov, model, min_inliers, tolerance_px, max_stars);
    std::copy(r.matrix, r.matrix + 9, matrix);
    *n_pairs = static_cast<int64_t>(r.pairs.size());
    if (pairs) {
      for (size_t k = 0; k < r.pairs.size(); ++k) {
        pairs[2 * k] = r.pairs[k].first;
        pairs[2 * k + 1] = r.pairs[k].second;
      }
    }
    *rms = std::isfinite(r.rms) ? r.rms : NAN;
    *ok = r.ok ? 1 : 0;
  });
}

}  // extern "C"
