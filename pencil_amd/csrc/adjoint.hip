// adjoint.hip - the adjoint pass: per-pixel replay with gradient
// accumulation, and the deterministic FP64 reduction of the partials.
//
// k_adjoint runs one 8x8 pixel tile per 64-lane block.  Each lane renders its
// pixel with render_pixel's program and the adjoint of adjoint.hpp, into the
// pixel's value and a private FP64 Jacobian, weights it by the pixel's image
// adjoint, then the block sums its 64 lanes with a fixed butterfly and writes
// one partial per parameter.  k_reduce_params sums the
// block partials of each parameter in block order with a fixed tree.  No
// atomics touch a gradient, so the result is the same on every run.
#include <hip/hip_runtime.h>

#include "adjoint.hpp"
#include "adjoint_launch.hpp"

namespace pcl {
namespace {

constexpr int kAdjBlock = 64;
constexpr int kReduceBlock = 256;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// One pass per gradient.  Each lane replays its pixel's samples with unit
// weight.  That yields two things at once:
//  - the pixel's value, with the forward render's bits: the replay performs
//    render_pixel's float operations in the same order;
//  - the pixel's Jacobian J.  Albedo and emission channel c reach only image
//    channel c, so J holds one number per parameter, accumulated in FP64.
// The pixel's adjoint weight W is then known: given (adj, the VJP), or the L2
// loss's 2 (I - target), whose loss is reduced into the last slot.  W scales J
// once: the gradient is the sum over pixels of W_c J / n^2.
// 128 VGPRs, 4 waves per SIMD: on the 512^2 64-spp, 8-material step the
// pass measured 37.2 ms, against 42.5 ms at 3 waves and 45.3 ms at 5; 8x8
// tiles and 64-pixel row segments per block measured the same.
template <int NM>
__global__ __launch_bounds__(kAdjBlock, 4) void k_adjoint(DevScene S, CamRec C, RSettings st, const float* tgt,
                                                       const float* adj, float* img_out, double* partial,
                                                       int n_par, unsigned long long* ctr) {
    __shared__ unsigned long long lds[kStackK * kAdjBlock];
    const int tiles_x = (st.width + 7) / 8;
    const int lane = threadIdx.x;
    const int i = (blockIdx.x % tiles_x) * 8 + (lane & 7);
    const int j = (blockIdx.x / tiles_x) * 8 + (lane >> 3);
    Grad<NM> G;
    G.clear();
    double loss = 0.0;
    LaneCount<false> c;
    if (i < st.width && j < st.height) {
        const int n = st.samples;
        alignas(16) unsigned char fbuf[sizeof(AdjFrame) * kMaxFrames];   // no zeroing constructors
        AdjFrame* F = reinterpret_cast<AdjFrame*>(fbuf);
        Stack stk;
        stk.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
        stk.set_stride(kAdjBlock);
        // render_pixel (render.hpp), replayed
        const int Wd = st.width, Hd = st.height;
        const f3 unit(1.0f, 1.0f, 1.0f);
        Rng rng((uint32_t)(i * 97861u + j * 1973u + 1u));
        f3 acc(0, 0, 0);
        for (int sy = 0; sy < n; ++sy) {
            for (int sx = 0; sx < n; ++sx) {
                float u = (i + (sx + 0.5f) / n) / (float)Wd;
                float v = 1.0f - (j + (sy + 0.5f) / n) / (float)Hd;
                u += rng.sym() * (0.5f / (Wd * n));
                v += rng.sym() * (0.5f / (Hd * n));
                Ray r = camera_ray(C, u, v, rng);
                if (st.pathtrace) acc += trace_path_adj(S, st, r, rng, F, stk, c, unit, G);
                else acc += whitted_adj(S, st, r, st.max_depth, rng, F, 0, stk, c, unit, G);
            }
        }
        const f3 I = acc / (float)(n * n);
        const size_t px = ((size_t)j * st.width + i) * 3;
        if (img_out) {
            img_out[px] = I.x; img_out[px + 1] = I.y; img_out[px + 2] = I.z;
        }
        f3 Wp;
        if (adj) {
            Wp = f3(adj[px], adj[px + 1], adj[px + 2]);
        } else {
            const f3 d(I.x - tgt[px], I.y - tgt[px + 1], I.z - tgt[px + 2]);
            loss = (double)d.x * d.x + (double)d.y * d.y + (double)d.z * d.z;
            Wp = d * 2.0f;
        }
        const double inv = 1.0 / ((double)n * (double)n);   // the image is the mean of n*n samples
        const double w[3] = {(double)Wp.x * inv, (double)Wp.y * inv, (double)Wp.z * inv};
#pragma unroll
        for (int q = 0; q < NM * 6; ++q) G.g[q] *= w[q % 3];
    }
    double* out = partial + (size_t)blockIdx.x * (n_par + 1);
    for (int p = 0; p < n_par; ++p) {
        const double v = wave_sum(G.g[p]);
        if (lane == 0) out[p] = v;
    }
    const double l = wave_sum(loss);
    if (lane == 0) out[n_par] = l;
    // work counters: rays, closest-hit and any-hit queries, one atomic each per wave
    const uint32_t k0 = wave_sum_u(c.rays), k1 = wave_sum_u(c.closest), k2 = wave_sum_u(c.anyhit);
    if (lane == 0) {
        atomicAdd(ctr, (unsigned long long)k0);
        atomicAdd(ctr + 1, (unsigned long long)k1);
        atomicAdd(ctr + 2, (unsigned long long)k2);
    }
}

// out[p] = sum over blocks b (in order of b within each thread's stride, then
// a fixed tree over threads) of partial[b][p].
__global__ __launch_bounds__(kReduceBlock) void k_reduce_params(const double* partial, int n_blocks, int stride,
                                                               double* out) {
    __shared__ double sh[kReduceBlock];
    const int p = blockIdx.x;
    double v = 0.0;
    for (int b = threadIdx.x; b < n_blocks; b += kReduceBlock) v += partial[(size_t)b * stride + p];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kReduceBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[p] = sh[0];
}

template <int NM>
void launch_nm(hipStream_t s, const DevScene& D, const CamRec& C, const RSettings& st, int tiles, const float* tgt,
               const float* adj, float* img_out, double* partial, int n_par, unsigned long long* ctr) {
    hipLaunchKernelGGL(k_adjoint<NM>, dim3(tiles), dim3(kAdjBlock), 0, s, D, C, st, tgt, adj, img_out, partial, n_par,
                       ctr);
}

}  // namespace

namespace adj {

int tiles(const RSettings& st) { return ((st.width + 7) / 8) * ((st.height + 7) / 8); }

size_t partial_doubles(const RSettings& st, int n_mat) { return (size_t)tiles(st) * (size_t)(n_mat * 6 + 1); }

hipError_t launch(hipStream_t s, const DevScene& D, const CamRec& C, const RSettings& st, int n_mat,
                  const float* tgt, const float* adjw, float* img_out, double* partial, double* out,
                  unsigned long long* ctr) {
    const int t = tiles(st);
    const int n_par = n_mat * 6;
    if (n_mat <= 4) launch_nm<4>(s, D, C, st, t, tgt, adjw, img_out, partial, n_par, ctr);
    else if (n_mat <= 8) launch_nm<8>(s, D, C, st, t, tgt, adjw, img_out, partial, n_par, ctr);
    else if (n_mat <= 16) launch_nm<16>(s, D, C, st, t, tgt, adjw, img_out, partial, n_par, ctr);
    else if (n_mat <= 32) launch_nm<32>(s, D, C, st, t, tgt, adjw, img_out, partial, n_par, ctr);
    else launch_nm<kAdjMaxMats>(s, D, C, st, t, tgt, adjw, img_out, partial, n_par, ctr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_params, dim3(n_par + 1), dim3(kReduceBlock), 0, s, partial, t, n_par + 1, out);
    return hipGetLastError();
}

}  // namespace adj
}  // namespace pcl
