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
// Enqueues the adjoint pass and the reduction on stream s.  The pass renders
// the image itself (into img_out when that is not null, the forward render's
// bits).  With adjw, the image adjoint is given; otherwise it is
// 2 (image - tgt) and the L2 loss is written to out[n_mat*6].  out: n_mat*6
// (+1) doubles, device.  ctr: 3 device words that get the rays traced and the
// closest-hit and any-hit queries added.
hipError_t launch(hipStream_t s, const DevScene& D, const CamRec& C, const RSettings& st, int n_mat,
                  const float* tgt, const float* adjw, float* img_out, double* partial, double* out,
                  unsigned long long* ctr);

}  // namespace adj
}  // namespace pcl
