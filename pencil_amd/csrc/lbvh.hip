// lbvh.hip - BVH built on the device (the fast builder behind use_bvh = 2).
//
// A linear BVH (Morton-ordered radix tree, after Karras, "Maximizing
// parallelism in the construction of BVHs, octrees, and k-d trees", HPG
// 2012): primitive box centers are quantized to 10 bits per axis inside the
// centers' bounds, interleaved into a 30-bit Morton code, made unique by the
// insertion id in the low 32 bits of a 64-bit key, and radix-sorted; every
// internal node then finds its key range and split from longest common
// prefixes, independently; boxes are filled bottom-up, one node depth per
// launch.  The binary tree is then collapsed into the 4-wide nodes the
// traversal reads: every internal node at an even depth heads one, whose
// children are its own children with each internal child replaced by that
// child's two children (the host collapse picks cuts by surface area; depth
// parity needs no top-down pass here).  The topology is not the reference's
// SAH tree: hits are the same except where equal t ties (or a grazing box
// test) resolve by DFS order.
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

// Depth of every internal node (root = 0) by walking its parent chain:
// unique 64-bit keys fix at least one more key bit per level, so no chain is
// longer than 64.
__global__ __launch_bounds__(kBlock) void k_depth(int n, const int32_t* parent_int, int32_t* ndepth) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    int d = 0, v = i;
    while (v != 0 && d < 65) { v = parent_int[v]; ++d; }
    ndepth[i] = d;
}

__device__ __forceinline__ void child_box(const NodeRec* nodes, const float* sorted_boxes, int32_t ref, float* b) {
    if (ref < 0) {
        const float* p = sorted_boxes + (size_t)(~ref) * 6;
        for (int a = 0; a < 6; ++a) b[a] = p[a];
    } else {
        const NodeRec& c = nodes[ref];
        for (int a = 0; a < 3; ++a) {
            b[a] = fminf(c.lb[a], c.rb[a]);
            b[3 + a] = fmaxf(c.lb[3 + a], c.rb[3 + a]);
        }
    }
}

// Boxes of the internal nodes at one depth, from their children (leaves, or
// nodes one level deeper whose boxes the previous launch wrote).
__global__ __launch_bounds__(kBlock) void k_level(int n, int depth, const int32_t* ndepth, NodeRec* nodes,
                                                  const float* sorted_boxes) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1 || ndepth[i] != depth) return;
    NodeRec& nd = nodes[i];
    child_box(nodes, sorted_boxes, nd.left, nd.lb);
    child_box(nodes, sorted_boxes, nd.right, nd.rb);
}

__global__ void k_root(const NodeRec* nodes, float* out7) {
    if (threadIdx.x != 0) return;
    const NodeRec& r = nodes[0];
    for (int a = 0; a < 3; ++a) {
        out7[a] = fminf(r.lb[a], r.rb[a]);
        out7[3 + a] = fmaxf(r.lb[3 + a], r.rb[3 + a]);
    }
}

__global__ __launch_bounds__(kBlock) void k_sort_boxes(const uint64_t* k, int n, const float* boxes, float* sorted) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const float* p = boxes + (size_t)(uint32_t)k[s] * 6;
    for (int a = 0; a < 6; ++a) sorted[(size_t)s * 6 + a] = p[a];
}

__global__ __launch_bounds__(kBlock) void k_wide_flags(int n, const int32_t* ndepth, int32_t* flag) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n - 1) flag[i] = (ndepth[i] & 1) == 0 ? 1 : 0;
}

__device__ __forceinline__ int32_t wide_ref(const int32_t* idx, int32_t ref) { return ref >= 0 ? idx[ref] : ref; }

// One 4-wide node per even-depth internal node i, at idx[i] (exclusive scan
// of the flags).  Internal grandchildren are two levels deeper, so they head
// 4-wide nodes of their own.
__global__ __launch_bounds__(kBlock) void k_wide(int n, const NodeRec* nodes, const int32_t* ndepth,
                                                 const int32_t* idx, Node4* out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1 || (ndepth[i] & 1) != 0) return;
    const NodeRec& nd = nodes[i];
    int32_t ref[4];
    const float* box[4];
    int cnt = 0;
    const int32_t kids[2] = {nd.left, nd.right};
    const float* kbox[2] = {nd.lb, nd.rb};
    for (int c = 0; c < 2; ++c) {
        if (kids[c] >= 0) {
            const NodeRec& m = nodes[kids[c]];
            ref[cnt] = wide_ref(idx, m.left); box[cnt++] = m.lb;
            ref[cnt] = wide_ref(idx, m.right); box[cnt++] = m.rb;
        } else {
            ref[cnt] = kids[c]; box[cnt++] = kbox[c];
        }
    }
    Node4 o;
    for (int k = 0; k < 4; ++k) {
        const bool used = k < cnt;
        o.lx[k] = used ? box[k][0] : INFINITY;
        o.ly[k] = used ? box[k][1] : INFINITY;
        o.lz[k] = used ? box[k][2] : INFINITY;
        o.hx[k] = used ? box[k][3] : -INFINITY;
        o.hy[k] = used ? box[k][4] : -INFINITY;
        o.hz[k] = used ? box[k][5] : -INFINITY;
        o.child[k] = used ? ref[k] : kEmptyRoot;
        o.pad[k] = 0;
    }
    out[idx[i]] = o;
}

}  // namespace

size_t temp_bytes(int n) {
    size_t sort = 0, red = 0, scan = 0;
    const int m = n > 1 ? n : 2;
    (void)hipcub::DeviceRadixSort::SortKeys(nullptr, sort, (const uint64_t*)nullptr, (uint64_t*)nullptr, m);
    (void)hipcub::DeviceReduce::Max(nullptr, red, (const int32_t*)nullptr, (int32_t*)nullptr, m - 1);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const int32_t*)nullptr, (int32_t*)nullptr, m - 1);
    size_t b = sort > red ? sort : red;
    return (b > scan ? b : scan) + 256;
}

hipError_t wide_count(hipStream_t st, const int32_t* d_ndepth, int n, int32_t* d_flag, int32_t* d_idx,
                      void* d_temp, size_t temp_size, int* count) {
    *count = 0;
    if (n < 2) return hipSuccess;
    const int gi = (n - 1 + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_wide_flags, dim3(gi), dim3(kBlock), 0, st, n, d_ndepth, d_flag);
    size_t tb = temp_size;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(d_temp, tb, d_flag, d_idx, n - 1, st);
    if (e != hipSuccess) return e;
    int32_t last[2] = {0, 0};
    e = hipMemcpyAsync(&last[0], d_idx + (n - 2), sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&last[1], d_flag + (n - 2), sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    *count = last[0] + last[1];
    return hipSuccess;
}

hipError_t wide_emit(hipStream_t st, const NodeRec* d_nodes, int n, const int32_t* d_ndepth, const int32_t* d_idx,
                     Node4* d_out) {
    if (n < 2) return hipSuccess;
    const int gi = (n - 1 + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_wide, dim3(gi), dim3(kBlock), 0, st, n, d_nodes, d_ndepth, d_idx, d_out);
    return hipGetLastError();
}

hipError_t build(hipStream_t st, const PrimRec* d_prims_in, const float* d_boxes, int n, PrimRec* d_prims_out,
                 NodeRec* d_nodes, int32_t* d_leaf_order, int32_t* d_slot_of, void* d_temp, size_t temp_size,
                 uint64_t* d_keys_a, uint64_t* d_keys_b, int32_t* d_parent_int, int32_t* d_parent_leaf,
                 int32_t* d_ndepth, float* d_sorted_boxes, float* d_out16, int* depth) {
    *depth = 0;
    if (n <= 0) return hipSuccess;
    const int gb = (n + kBlock - 1) / kBlock;
    float* cb = d_out16 + 8;   // centers' bounds, after root box (6) and the depth slot
    hipLaunchKernelGGL(k_center_bounds, dim3(1), dim3(1024), 0, st, d_boxes, n, cb);
    hipLaunchKernelGGL(k_keys, dim3(gb), dim3(kBlock), 0, st, d_boxes, n, cb, d_keys_a);
    size_t tb = temp_size;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(d_temp, tb, d_keys_a, d_keys_b, n, 0, 64, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_leaves, dim3(gb), dim3(kBlock), 0, st, d_keys_b, n, d_prims_in, d_prims_out, d_leaf_order,
                       d_slot_of);
    if (n == 1) return hipGetLastError();   // a single leaf: no internal node
    const int gi = (n - 1 + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_tree, dim3(gi), dim3(kBlock), 0, st, d_keys_b, n, d_nodes, d_parent_int, d_parent_leaf);
    hipLaunchKernelGGL(k_sort_boxes, dim3(gb), dim3(kBlock), 0, st, d_keys_b, n, d_boxes, d_sorted_boxes);
    hipLaunchKernelGGL(k_depth, dim3(gi), dim3(kBlock), 0, st, n, d_parent_int, d_ndepth);
    int32_t* d_max = reinterpret_cast<int32_t*>(d_out16 + 6);
    tb = temp_size;
    e = hipcub::DeviceReduce::Max(d_temp, tb, d_ndepth, d_max, n - 1, st);
    if (e != hipSuccess) return e;
    int32_t maxd = 0;
    e = hipMemcpyAsync(&maxd, d_max, sizeof(maxd), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (maxd < 0 || maxd > 64) return hipErrorUnknown;   // impossible for unique keys
    // boxes bottom-up, one depth per launch: kernel order makes the children's
    // boxes visible to their parents, no atomics or fences needed
    for (int d = maxd; d >= 0; --d)
        hipLaunchKernelGGL(k_level, dim3(gi), dim3(kBlock), 0, st, n, d, d_ndepth, d_nodes, d_sorted_boxes);
    hipLaunchKernelGGL(k_root, dim3(1), dim3(64), 0, st, d_nodes, d_out16);
    *depth = maxd + 1;   // internal nodes on the longest root-to-leaf path
    return hipGetLastError();
}

}  // namespace lbvh
}  // namespace pcl
