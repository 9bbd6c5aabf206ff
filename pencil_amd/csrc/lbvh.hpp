// lbvh.hpp - host side of the device BVH builder (lbvh.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "scene_layout.hpp"

namespace pcl {
namespace lbvh {

// Bytes of sort workspace for n primitives.
size_t temp_bytes(int n);

// Builds the tree for n >= 1 primitives on st (synchronizes once, to size
// the bottom-up pass).  Inputs: primitive records in insertion order and
// their boxes (n x 6 floats, min xyz, max xyz).  Outputs: records in leaf
// order, n - 1 internal nodes (root = node 0; a single primitive has none),
// slot -> insertion id and id -> slot maps, d_out16[0..5] = root box (n >= 2)
// and *depth = internal nodes on the longest path.  Work arrays: two n-key
// buffers, n parents each for nodes and leaves, n node depths, 6n floats of
// sorted boxes, 16 floats of d_out16, temp_bytes(n) of workspace.
hipError_t build(hipStream_t st, const PrimRec* d_prims_in, const float* d_boxes, int n, PrimRec* d_prims_out,
                 NodeRec* d_nodes, int32_t* d_leaf_order, int32_t* d_slot_of, void* d_temp, size_t temp_size,
                 uint64_t* d_keys_a, uint64_t* d_keys_b, int32_t* d_parent_int, int32_t* d_parent_leaf,
                 int32_t* d_ndepth, float* d_sorted_boxes, float* d_out16, int* depth);

// 4-wide nodes of the tree build() left in d_nodes (n >= 2): flags and an
// exclusive scan over the n - 1 internal nodes (d_flag, d_idx: n - 1 ints
// each; d_ndepth from build()), *count = number of 4-wide nodes (synchronizes
// once); wide_emit then writes them to d_out (count x 128 B, root = 0).
hipError_t wide_count(hipStream_t st, const int32_t* d_ndepth, int n, int32_t* d_flag, int32_t* d_idx,
                      void* d_temp, size_t temp_size, int* count);
hipError_t wide_emit(hipStream_t st, const NodeRec* d_nodes, int n, const int32_t* d_ndepth, const int32_t* d_idx,
                     Node4* d_out);

}  // namespace lbvh
}  // namespace pcl
