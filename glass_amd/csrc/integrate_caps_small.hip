// Register-sort winsorized integration kernels of capacities 16, 24, 32, 48, 64
// (up to 64 frames); one translation unit per capacity set so that the sorting networks
// compile in parallel.  The device code is integrate_pixel.hpp.

#include "integrate_pixel.hpp"

namespace gh {

bool launch_integrate_caps_small(int cap, const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid,
                                 hipStream_t stream) {
  switch (cap) {
    case 16:
      launch_cap<16>(a, t, grid, stream);
      return true;
    case 24:
      launch_cap<24>(a, t, grid, stream);
      return true;
    case 32:
      launch_cap<32>(a, t, grid, stream);
      return true;
    case 48:
      launch_cap<48>(a, t, grid, stream);
      return true;
    case 64:
      launch_cap<64>(a, t, grid, stream);
      return true;
    default:
      return false;
  }
}

}  // namespace gh
