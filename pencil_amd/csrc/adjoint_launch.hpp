// adjoint_launch.hpp - host-side entry into the adjoint kernels (adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "render.hpp"

namespace pcl {
namespace adj {

// 8x8 pixel tiles of the image (one adjoint block each).
int tiles(const RSettings& st);
// Size of the block-partials buffer: tiles x (n_mat*6 + 1) doubles.
size_t partial_doubles(const RSettings& st, int n_mat);
// Enqueues the adjoint pass and the reduction on stream s.  With adjw, the
// image adjoint is given; otherwise it is 2 (img - tgt) and the L2 loss is
// written to out[n_mat*6].  out: n_mat*6 (+1) doubles, device.
hipError_t launch(hipStream_t s, const DevScene& D, const CamRec& C, const RSettings& st, int n_mat,
                  const float* img, const float* tgt, const float* adjw, double* partial, double* out);

}  // namespace adj
}  // namespace pcl
