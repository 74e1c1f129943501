// Register-sort winsorized integration kernels of capacities 72, 80, 88, 96, 104
// (65 to 104 frames); one translation unit per capacity set so that the
// sorting networks compile in parallel.  The device code is integrate_pixel.hpp.

#include "integrate_pixel.hpp"

namespace gh {

bool launch_integrate_caps_mid(int cap, const IntegrateArgs& a, const IntegrateJobTable& t, unsigned grid,
                               hipStream_t stream) {
  switch (cap) {
    case 72:
      launch_cap<72>(a, t, grid, stream);
      return true;
    case 80:
      launch_cap<80>(a, t, grid, stream);
      return true;
    case 88:
      launch_cap<88>(a, t, grid, stream);
      return true;
    case 96:
      launch_cap<96>(a, t, grid, stream);
      return true;
    case 104:
      launch_cap<104>(a, t, grid, stream);
      return true;
    default:
      return false;
  }
}

}  // namespace gh
