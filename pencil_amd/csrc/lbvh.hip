// lbvh.hip - BVH built on the device (the fast builder behind use_bvh = 2).
//
// A linear BVH (Morton-ordered radix tree, after Karras, "Maximizing
// parallelism in the construction of BVHs, octrees, and k-d trees", HPG
// 2012): primitive box centers are quantized to 10 bits per axis inside the
// centers' bounds, interleaved into a 30-bit Morton code, made unique by the
// insertion id in the low 32 bits of a 64-bit key, and radix-sorted; every
// internal node then finds its key range and split from longest common
// prefixes, independently; boxes are filled bottom-up, the second child to
// arrive at a node combining both.  Same NodeRec / PrimRec layout as the host
// build, so the traversal kernels are unchanged.  The topology is not the
// reference's SAH tree: hits are the same except where equal t ties (or a
// grazing box test) resolve by DFS order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "lbvh.hpp"

namespace pcl {
namespace lbvh {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t spread10(uint32_t v) {   // 10 bits -> every third bit
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// centers' bounds: one block, grid-stride, then a tree in LDS
__global__ __launch_bounds__(1024) void k_center_bounds(const float* boxes, int n, float* out6) {
    __shared__ float sh[6][1024];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float* b = boxes + (size_t)i * 6;
        for (int a = 0; a < 3; ++a) {
            const float c = (b[a] + b[3 + a]) * 0.5f;
            mn[a] = fminf(mn[a], c);
            mx[a] = fmaxf(mx[a], c);
        }
    }
    for (int a = 0; a < 3; ++a) { sh[a][threadIdx.x] = mn[a]; sh[3 + a][threadIdx.x] = mx[a]; }
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 3; ++a) {
                sh[a][threadIdx.x] = fminf(sh[a][threadIdx.x], sh[a][threadIdx.x + w]);
                sh[3 + a][threadIdx.x] = fmaxf(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out6[threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(kBlock) void k_keys(const float* boxes, int n, const float* cb, uint64_t* keys) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float* b = boxes + (size_t)i * 6;
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        const float c = (b[a] + b[3 + a]) * 0.5f;
        const float ext = cb[3 + a] - cb[a];
        float u = ext > 0.0f ? (c - cb[a]) / ext : 0.0f;
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        q[a] = (uint32_t)fminf(u * 1024.0f, 1023.0f);
    }
    const uint32_t m = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    keys[i] = ((uint64_t)m << 32) | (uint32_t)i;
}

__device__ __forceinline__ int delta(const uint64_t* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    return __clzll((long long)(k[i] ^ k[j]));   // keys are unique: never 64
}

// Karras 2012, Fig. 4: internal node i of n - 1.  Child refs: internal
// index, or ~slot for a leaf.
__global__ __launch_bounds__(kBlock) void k_tree(const uint64_t* k, int n, NodeRec* nodes, int32_t* parent_int,
                                                 int32_t* parent_leaf) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(k, n, i, i - d);
    int lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(k, n, i, j);
    int s = 0;
    for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) {
        if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    NodeRec& nd = nodes[i];
    if (lo == gamma) { nd.left = ~gamma; parent_leaf[gamma] = i; }
    else { nd.left = gamma; parent_int[gamma] = i; }
    if (hi == gamma + 1) { nd.right = ~(gamma + 1); parent_leaf[gamma + 1] = i; }
    else { nd.right = gamma + 1; parent_int[gamma + 1] = i; }
    nd.pad0 = nd.pad1 = 0;
}

// Leaves in key order: permuted primitive records and the slot <-> id maps.
__global__ __launch_bounds__(kBlock) void k_leaves(const uint64_t* k, int n, const PrimRec* in, PrimRec* out,
                                                   int32_t* leaf_order, int32_t* slot_of) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const int id = (int)(uint32_t)k[s];
    out[s] = in[id];
    leaf_order[s] = id;
    slot_of[id] = s;
}

__device__ __forceinline__ void put_box(float* dst, const float* b) {
    for (int a = 0; a < 6; ++a) dst[a] = b[a];
}

// Bottom-up boxes: each leaf climbs; at a node the first arrival stops, the
// second (which sees the other child's box after the device-scope fence)
// merges both and climbs on.  Heights give the depth (internal levels).
__global__ __launch_bounds__(kBlock) void k_boxes(const uint64_t* k, int n, const float* boxes,
                                                  NodeRec* nodes, const int32_t* parent_int,
                                                  const int32_t* parent_leaf, uint32_t* flags, int32_t* height,
                                                  float* root_box, int32_t* depth_out) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    float b[6];
    const float* pb = boxes + (size_t)(uint32_t)k[s] * 6;
    for (int a = 0; a < 6; ++a) b[a] = pb[a];
    int32_t child = ~s;
    int h = 0;
    int node = parent_leaf[s];
    // unique 64-bit keys: every level down fixes at least one more key bit,
    // so no path is longer than 64 internal nodes
    for (int guard = 0; guard < 65; ++guard) {
        NodeRec& nd = nodes[node];
        float* slot = (nd.left == child) ? nd.lb : nd.rb;
        put_box(slot, b);
        height[node * 2 + ((nd.left == child) ? 0 : 1)] = h;
        __threadfence();
        const uint32_t prev = __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == 0) return;
        __threadfence();
        const volatile float* L = nd.lb;
        const volatile float* R = nd.rb;
        for (int a = 0; a < 3; ++a) {
            b[a] = fminf(L[a], R[a]);
            b[3 + a] = fmaxf(L[3 + a], R[3 + a]);
        }
        const volatile int32_t* hv = height;
        const int hl = hv[node * 2], hr = hv[node * 2 + 1];
        h = (hl > hr ? hl : hr) + 1;
        if (node == 0) {
            put_box(root_box, b);
            *depth_out = h;
            return;
        }
        child = node;
        node = parent_int[node];
    }
    *depth_out = -1;   // unreachable for a well-formed tree; reported as an error
}

}  // namespace

size_t temp_bytes(int n) {
    size_t sort = 0;
    hipcub::DeviceRadixSort::SortKeys(nullptr, sort, (const uint64_t*)nullptr, (uint64_t*)nullptr, n > 0 ? n : 1);
    return sort + 256;
}

hipError_t build(hipStream_t st, const PrimRec* d_prims_in, const float* d_boxes, int n, PrimRec* d_prims_out,
                 NodeRec* d_nodes, int32_t* d_leaf_order, int32_t* d_slot_of, void* d_temp, size_t temp_size,
                 uint64_t* d_keys_a, uint64_t* d_keys_b, int32_t* d_parent_int, int32_t* d_parent_leaf,
                 uint32_t* d_flags, int32_t* d_height, float* d_out7) {
    if (n <= 0) return hipSuccess;
    const int gb = (n + kBlock - 1) / kBlock;
    float* cb = d_out7 + 8;   // centers' bounds scratch after root box (6) + depth (1)
    hipLaunchKernelGGL(k_center_bounds, dim3(1), dim3(1024), 0, st, d_boxes, n, cb);
    hipLaunchKernelGGL(k_keys, dim3(gb), dim3(kBlock), 0, st, d_boxes, n, cb, d_keys_a);
    size_t tb = temp_size;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(d_temp, tb, d_keys_a, d_keys_b, n, 0, 64, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_leaves, dim3(gb), dim3(kBlock), 0, st, d_keys_b, n, d_prims_in, d_prims_out, d_leaf_order,
                       d_slot_of);
    if (n == 1) {
        // a single leaf: no internal node; the caller takes the root box from the primitive
        return hipGetLastError();
    }
    const int gi = (n - 1 + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_tree, dim3(gi), dim3(kBlock), 0, st, d_keys_b, n, d_nodes, d_parent_int, d_parent_leaf);
    e = hipMemsetAsync(d_flags, 0, sizeof(uint32_t) * (size_t)(n - 1), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_boxes, dim3(gb), dim3(kBlock), 0, st, d_keys_b, n, d_boxes, d_nodes, d_parent_int,
                       d_parent_leaf, d_flags, d_height, d_out7, reinterpret_cast<int32_t*>(d_out7 + 6));
    return hipGetLastError();
}

}  // namespace lbvh
}  // namespace pcl
