// Affine Lanczos3 warp with its adjoint (photometric registration refinement).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace gh {

// s_x = b0 x + b1 y + b2, s_y = b3 x + b4 y + b5: reference pixel -> source pixel.
struct AffineMap {
  double b[6];
};

// Per-block sums of the Gauss-Newton pass: J^T J upper triangle (21),
// J^T r (6), sum r^2, valid-pixel count.
constexpr int kNormalSums = 29;

// Doubles of device workspace the backward / normal-equation passes need for
// an image of npix pixels (block partials plus the reduced sums).
size_t adjoint_workspace_doubles(long long npix);

void launch_affine_forward(const float* src, int h, int w, const AffineMap& m, float fill, float* out,
                           hipStream_t stream);
// grad_src (nullable) += adjoint(grad_out); grad_map (nullable, device, 6
// doubles) = d <grad_out, warp> / d B.  workspace: adjoint_workspace_doubles.
void launch_affine_backward(const float* src, int h, int w, const AffineMap& m, const float* grad_out,
                            float* grad_src, double* workspace, double* grad_map, hipStream_t stream);
// sums (device, kNormalSums doubles) for the centred Jacobian (x - cx, y - cy).
void launch_normal_equations(const float* ref, const float* mov, int h, int w, const AffineMap& m, double cx,
                             double cy, double* workspace, double* sums, hipStream_t stream);

}  // namespace gh
