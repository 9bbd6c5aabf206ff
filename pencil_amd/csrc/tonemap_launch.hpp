// tonemap_launch.hpp - host side of the device tone map (tonemap.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcl {
namespace tm {

// Enqueues the 8-bit encode of n_px HDR pixels (RGB float, row-major) on s.
// scale = 2^exposure as the host computes it; d_thr: the 256-float threshold
// table of tonemap_thresholds.  d_hdr 16-byte aligned, d_out 4-byte aligned.
hipError_t launch(hipStream_t s, const float* d_hdr, int64_t n_px, float scale, const float* d_thr, uint8_t* d_out);

}  // namespace tm
}  // namespace pcl
