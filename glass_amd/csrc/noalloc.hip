// gh_integrate_noalloc: the resident integration (gh_integrate) for callers
// that cannot allocate device memory.  It has no counterpart in the
// reference.
//
// The integration reads, besides the frames and the outputs, a frame-pointer
// table, the frame weights and (StackEngine tiles) per-tile statistics flags.
// gh_integrate keeps those in buffers of the stack, allocated on first use;
// here they live in one file-scope device array, `scratch`, of fixed size
// (kScratchBytes, held by each GPU until the process ends).  Outputs must be
// device memory (no host staging).  The call allocates no device memory and
// creates no runtime object.
//
// In host code `scratch` names a host-side stand-in, not a device address:
// its device address is looked up once per device with hipGetSymbolAddress
// (on the first call with that device current) and only that address is used.
// Every call on one GPU shares its copy of the array and leaves its last
// contents there; the entry point waits for the previous launch on the same
// stream before rewriting it, but calls that overlap on one GPU from
// different streams or threads race.

#include <mutex>

#include "stack.hpp"

namespace {

constexpr size_t kTableEntries = 16384;  // frames: integrate_max_frames()
constexpr size_t kTileEntries = 16384;   // StackEngine statistics tiles
constexpr size_t kTableBytes = kTableEntries * sizeof(const float*);
constexpr size_t kWeightBytes = kTableEntries * sizeof(float);
constexpr size_t kScratchBytes = kTableBytes + kWeightBytes + kTileEntries;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's `scratch`, looked up on the first call.
unsigned char* scratch_on_current_device() {
  int device = 0;
  check(hipGetDevice(&device), "hipGetDevice");
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("HIP device id beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(integration scratch)");
    if (!ptr) throw std::runtime_error("hipGetSymbolAddress returned no address for the integration scratch");
    g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
  }
  return g_scratch_ptr[device];
}

}  // namespace

extern "C" {

int gh_integrate_noalloc(gh_stack* s, const float* weights, const gh_integrate_params* params,
                         const gh_integrate_outputs* outputs, void* stream) {
  return ghi::guarded([&] {
    if (static_cast<size_t>(gh::integrate_max_frames()) > kTableEntries ||
        static_cast<size_t>(gh::integrate_register_capacity()) > kTableEntries)
      throw std::runtime_error("integration frame limit exceeds the no-allocation scratch");
    unsigned char* base = scratch_on_current_device();
    ghi::IntegrationScratch sc{};
    sc.frame_table = reinterpret_cast<const float**>(base);
    sc.weights = reinterpret_cast<float*>(base + kTableBytes);
    sc.table_entries = kTableEntries;
    sc.tile_flags = base + kTableBytes + kWeightBytes;
    sc.tile_entries = kTileEntries;
    ghi::integrate_stack(s, weights, params, outputs, stream, &sc);
  });
}

}  // extern "C"
