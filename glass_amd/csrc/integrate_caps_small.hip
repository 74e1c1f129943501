// Register-sort winsorized integration kernels of capacities 16, 24, 32, 40, 48, 56, 64
// (1 to 64 frames); one translation unit per capacity set so that the
// sorting networks compile in parallel.  The device code is integrate_pixel.hpp.

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
    case 40:
      launch_cap<40>(a, t, grid, stream);
      return true;
    case 48:
      launch_cap<48>(a, t, grid, stream);
      return true;
    case 56:
      launch_cap<56>(a, t, grid, stream);
      return true;
    case 64:
      launch_cap<64>(a, t, grid, stream);
      return true;
    default:
      return false;
  }
}

}  // namespace gh
