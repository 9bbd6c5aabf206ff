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

// Enqueues the build of n >= 1 primitives on st.  Inputs: primitive records
// in insertion order and their boxes (n x 6 floats, min xyz, max xyz).
// Outputs: records in leaf order, n - 1 internal nodes (root = node 0; a
// single primitive has none), slot -> insertion id and id -> slot maps, and
// d_out7[0..5] = root box, ((int32_t*)d_out7)[6] = depth (internal levels on
// the longest path; n >= 2).  d_out7 holds 16 floats (the rest is scratch).
// Work arrays: two n-key buffers, n - 1 internal parents, n leaf parents,
// n - 1 flags, 2 (n - 1) heights, and temp_bytes(n) of sort workspace.
hipError_t build(hipStream_t st, const PrimRec* d_prims_in, const float* d_boxes, int n, PrimRec* d_prims_out,
                 NodeRec* d_nodes, int32_t* d_leaf_order, int32_t* d_slot_of, void* d_temp, size_t temp_size,
                 uint64_t* d_keys_a, uint64_t* d_keys_b, int32_t* d_parent_int, int32_t* d_parent_leaf,
                 uint32_t* d_flags, int32_t* d_height, float* d_out7);

}  // namespace lbvh
}  // namespace pcl
