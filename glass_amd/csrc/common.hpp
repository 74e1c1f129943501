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

inline unsigned grid_for(long long n, int per_block) {
  long long g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

}  // namespace gh
