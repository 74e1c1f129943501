// Register-sort winsorized integration kernels of capacities 112, 120, 128
// (105 to 128 frames); one translation unit per capacity set so that the
// sorting networks compile in parallel.  The device code is integrate_pixel.hpp.

#include "integrate_pixel.hpp"

namespace gh {

bool launch_integrate_caps_upper(int cap, const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid,
                                 hipStream_t stream) {
  switch (cap) {
    case 112:
      launch_cap<112>(a, t, grid, stream);
      return true;
    case 120:
      launch_cap<120>(a, t, grid, stream);
      return true;
    case 128:
      launch_cap<128>(a, t, grid, stream);
      return true;
    default:
      return false;
  }
}

}  // namespace gh
