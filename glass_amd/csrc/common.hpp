// Shared helpers for the gfx950 resident stacking library (libglass_hip.so).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace gh {

// Error classes map onto the C-ABI status codes; the Python host turns them
// into ValueError / IndexError / RuntimeError like the reference bindings
// (cpp/src/native_bindings.cpp:1011-1018, invalid_argument / out_of_range).
struct ValueError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct IndexError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(hipError_t err, const char* what) {
  if (err != hipSuccess) {
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(err));
  }
}

constexpr int kBlock = 256;

// Wave-uniform reads of device descriptor tables.  A descriptor read from
// memory (rather than passed as a kernel argument) holds generic pointers;
// loads through generic pointers are flat loads, which the compiler treats as
// lane-varying, so every address derived from them would live in VGPR pairs.
// Reading the table through the global address space and broadcasting the
// first active lane's words marks the values wave-uniform: pointers become
// scalar (SGPR) bases.  Valid only where every lane reads the same entry.
typedef __attribute__((address_space(1))) const unsigned global_u32;
typedef __attribute__((address_space(1))) const char global_char;
typedef __attribute__((address_space(1))) const unsigned long long global_u64;

typedef __attribute__((address_space(1))) float global_f32;
typedef __attribute__((address_space(1))) const float global_cf32;

// Frame memory is global: these accessors keep loads and stores off the flat
// path when the frame pointer was read from a descriptor table.
__device__ __forceinline__ float ld_global(const float* base, long long i) { return ((const global_cf32*)base)[i]; }
__device__ __forceinline__ void st_global(float* base, long long i, float v) { ((global_f32*)base)[i] = v; }
// Element i < 2^30 of a wave-uniform base: scalar base + 32-bit byte offset.
__device__ __forceinline__ void st_global_u32(float* base, unsigned i, float v) {
  typedef __attribute__((address_space(1))) char global_char_rw;
  *(global_f32*)((global_char_rw*)base + (i << 2)) = v;
}

template <typename T>
__device__ __forceinline__ T load_uniform(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "descriptor must be whole 32-bit words");
  constexpr int kWords = static_cast<int>(sizeof(T) / 4);
  unsigned w[kWords];
  const global_u32* src = (const global_u32*)p;
#pragma unroll
  for (int i = 0; i < kWords; ++i) w[i] = __builtin_amdgcn_readfirstlane(src[i]);
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

// A wave-uniform descriptor read by scalar loads (constant address space):
// unlike load_uniform it issues no vector memory operation, so it does not
// wait behind a wave's outstanding vector loads.  Valid only where every
// lane reads the same entry and no earlier work of this kernel wrote it.
template <typename T>
__device__ __forceinline__ T load_scalar(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "descriptor must be whole 32-bit words");
  constexpr int kWords = static_cast<int>(sizeof(T) / 4);
  typedef __attribute__((address_space(4))) const unsigned const_u32_word;
  const const_u32_word* src = (const const_u32_word*)p;
  unsigned w[kWords];
#pragma unroll
  for (int i = 0; i < kWords; ++i) w[i] = src[i];
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

// Entry k of a device table of pointers, as a wave-uniform pointer.  The
// table is read through the constant address space, so a wave-uniform k
// becomes a scalar load into SGPRs (a global load would fetch the same 8
// bytes into every lane and then need two v_readfirstlane per entry).
typedef __attribute__((address_space(4))) const unsigned long long const_u64;
typedef __attribute__((address_space(4))) const float const_f32;  // wave-uniform reads: scalar loads
template <typename T>
__device__ __forceinline__ T* table_at(T* const* table, long long k) {
  const unsigned long long v = ((const_u64*)table)[k];
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(v >> 32));
  return reinterpret_cast<T*>((static_cast<unsigned long long>(hi) << 32) | lo);
}

inline unsigned grid_for(long long n, int per_block) {
  long long g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

}  // namespace gh
