// Register-sort winsorized integration kernels of capacities 160, 200, 256
// (129 to 256 frames); one translation unit per capacity set so that the sorting networks
// compile in parallel.  The device code is integrate_pixel.hpp.

#include "integrate_pixel.hpp"

namespace gh {

bool launch_integrate_caps_large(int cap, const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid,
                                 hipStream_t stream) {
  switch (cap) {
    case 160:
      launch_cap<160>(a, t, grid, stream);
      return true;
    case 200:
      launch_cap<200>(a, t, grid, stream);
      return true;
    case 256:
      launch_cap<256>(a, t, grid, stream);
      return true;
    default:
      return false;
  }
}

}  // namespace gh
