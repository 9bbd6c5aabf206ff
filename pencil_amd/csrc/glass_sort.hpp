// glass_sort.hpp - the order in which a glass launch takes its tasks.
//
// Under soft lights each deferred glass vertex's subtree was walked and its
// shaded nodes counted (render.hpp: walk_wave, GlassSink::keys); under hard
// lights, with no walk, the key is the subtree's depth limit.  The glass
// launches (core.hip: k_glass, wavefront.hip: k_finish_glass) take the tasks
// by that key, largest first, so that a wave's lanes hold subtrees of about
// one size instead of waiting for the largest of 64 random ones.  The values
// do not depend on the order: each task writes its own sample.  A counting
// sort over task indices: histogram, offsets, and a scatter that reserves
// each block's share of a bin with one atomic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcl {
namespace {

// Counts of 255 and more share the last bin.
constexpr int kSortBins = 256, kSortChunk = 4096;   // tasks per block of 256 threads

// Clears the 512 words of counts and offsets (a kernel, not a memset).
__global__ __launch_bounds__(256) void k_sort_clear(unsigned long long* bins) { bins[threadIdx.x] = 0ull; }

__global__ __launch_bounds__(256) void k_glass_hist(const uint32_t* __restrict__ keys, const uint32_t* count,
                                                    uint32_t cap, uint32_t* bins) {
    __shared__ uint32_t h[kSortBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = min(*count, cap), b0 = blockIdx.x * kSortChunk, b1 = min(n, b0 + kSortChunk);
    for (uint32_t t = b0 + threadIdx.x; t < b1; t += 256) atomicAdd(&h[min(keys[t], (uint32_t)kSortBins - 1)], 1u);
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&bins[threadIdx.x], h[threadIdx.x]);
}

// One block of 256: bins[256 + k] = first position of key k, keys in
// descending order.
__global__ __launch_bounds__(256) void k_glass_scan(uint32_t* bins) {
    __shared__ uint32_t v[kSortBins];
    const int i = threadIdx.x;
    const uint32_t own = bins[kSortBins - 1 - i];
    v[i] = own;
    __syncthreads();
    for (int d = 1; d < kSortBins; d <<= 1) {
        const uint32_t x = i >= d ? v[i - d] : 0u;
        __syncthreads();
        v[i] += x;
        __syncthreads();
    }
    bins[kSortBins + (kSortBins - 1 - i)] = v[i] - own;
}

__global__ __launch_bounds__(256) void k_glass_scatter(const uint32_t* __restrict__ keys, const uint32_t* count,
                                                       uint32_t cap, uint32_t* bins, uint32_t* __restrict__ perm) {
    __shared__ uint32_t h[kSortBins], base[kSortBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = min(*count, cap), b0 = blockIdx.x * kSortChunk, b1 = min(n, b0 + kSortChunk);
    for (uint32_t t = b0 + threadIdx.x; t < b1; t += 256) atomicAdd(&h[min(keys[t], (uint32_t)kSortBins - 1)], 1u);
    __syncthreads();
    base[threadIdx.x] = h[threadIdx.x] ? atomicAdd(&bins[kSortBins + threadIdx.x], h[threadIdx.x]) : 0u;
    h[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t t = b0 + threadIdx.x; t < b1; t += 256) {
        const uint32_t k = min(keys[t], (uint32_t)kSortBins - 1);
        perm[base[k] + atomicAdd(&h[k], 1u)] = t;
    }
}

// The four launches on `st`: bins (512 words) are cleared here; perm gets
// the order of tasks 0 .. min(*count, cap) - 1.
inline void glass_sort(const uint32_t* keys, const uint32_t* count, uint32_t cap, uint32_t* bins, uint32_t* perm,
                       hipStream_t st) {
    const unsigned sb = (unsigned)((cap + kSortChunk - 1) / kSortChunk);
    hipLaunchKernelGGL(k_sort_clear, dim3(1), dim3(256), 0, st, reinterpret_cast<unsigned long long*>(bins));
    hipLaunchKernelGGL(k_glass_hist, dim3(sb), dim3(256), 0, st, keys, count, cap, bins);
    hipLaunchKernelGGL(k_glass_scan, dim3(1), dim3(256), 0, st, bins);
    hipLaunchKernelGGL(k_glass_scatter, dim3(sb), dim3(256), 0, st, keys, count, cap, bins, perm);
}

}  // namespace
}  // namespace pcl
