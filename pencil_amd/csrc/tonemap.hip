// tonemap.hip - HDR -> 8-bit on the device: exposure, ACES, clamp, gamma
// 1/2.2, (uint8)(x * 255 + 0.5), i.e. color_to_rgb8 (src/image_io.hpp:46-59)
// with tonemap_color's ACES curve (src/tonemap.hpp:27-61).
//
// Exposure scale and ACES are float + - * / in the reference's order, exact
// here.  The gamma step goes through a threshold table instead of a device
// pow: encode(v) = (uint8)(powf(v, 1/2.2f) * 255 + 0.5) is a non-decreasing
// step function of v in [0, 1], so the host finds, with its own powf (the
// reference's libm), the smallest float t[k] with encode(t[k]) >= k for k =
// 1..255, and the device returns the number of thresholds <= v.  The output
// is then the host's bit for bit.  One lane encodes 4 pixels: 48 B read as
// three dwordx4, 12 B written as three dwords.
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "tonemap_launch.hpp"

namespace pcl {
namespace {

constexpr int kTmBlock = 256;

__device__ __forceinline__ float aces_dev(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    const float y = (x * (a * x + b)) / (x * (c * x + d) + e);
    return (0.0f < y) ? y : 0.0f;   // std::max(0.0f, y): NaN -> 0
}

__device__ __forceinline__ uint32_t encode(float hdr, float scale, const float* t) {
    float m = aces_dev(hdr * scale);
    m = clampf(m, 0.0f, 1.0f);
    // count of t[1..255] <= m: binary search over the sorted thresholds
    int lo = 0;   // answer in [lo, lo + 256)
#pragma unroll
    for (int step = 128; step >= 1; step >>= 1)
        if (t[lo + step] <= m) lo += step;
    return (uint32_t)lo;
}

__global__ __launch_bounds__(kTmBlock) void k_tonemap(const float* __restrict__ hdr, int64_t n_px, float scale,
                                                      const float* __restrict__ thr, uint8_t* __restrict__ out) {
    __shared__ float t[256];
    t[threadIdx.x] = thr[threadIdx.x];   // t[0] = -inf never counts: see tonemap_thresholds
    __syncthreads();
    const int64_t q = (int64_t)blockIdx.x * kTmBlock + threadIdx.x;   // quad of pixels
    const int64_t p0 = q * 4;
    if (p0 >= n_px) return;
    if (p0 + 4 <= n_px) {
        const float4* src = reinterpret_cast<const float4*>(hdr + p0 * 3);
        const float4 a = gload_f4(src), b = gload_f4(src + 1), c = gload_f4(src + 2);
        const float v[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
        uint32_t w[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            w[k] = encode(v[4 * k], scale, t) | (encode(v[4 * k + 1], scale, t) << 8) |
                   (encode(v[4 * k + 2], scale, t) << 16) | (encode(v[4 * k + 3], scale, t) << 24);
        }
        uint32_t* dst = reinterpret_cast<uint32_t*>(out + p0 * 3);
        dst[0] = w[0]; dst[1] = w[1]; dst[2] = w[2];
    } else {
        for (int64_t i = p0 * 3; i < n_px * 3; ++i) out[i] = (uint8_t)encode(hdr[i], scale, t);
    }
}

}  // namespace

namespace tm {

hipError_t launch(hipStream_t s, const float* d_hdr, int64_t n_px, float scale, const float* d_thr, uint8_t* d_out) {
    const int64_t quads = (n_px + 3) / 4;
    const int64_t blocks = (quads + kTmBlock - 1) / kTmBlock;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tonemap, dim3((unsigned)blocks), dim3(kTmBlock), 0, s, d_hdr, n_px, scale, d_thr, d_out);
    return hipGetLastError();
}

}  // namespace tm
}  // namespace pcl
