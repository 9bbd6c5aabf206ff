// scene_layout.hpp - device memory layout of a scene (shared by the host
// builder and the kernels).  Plain structs, no HIP types.
//
//   nodes   : the traversal reads 4-wide nodes (Node4, 128 B): each is one
//             binary internal node of the reference-topology tree together
//             with up to four of its descendants (a cut of its subtree), the
//             cut's boxes in SoA and their refs, in the binary tree's
//             left-first order.  NodeRec (BVH2, 64 B: both child boxes + both
//             refs) is what the builders produce first.
//   prims   : one 64-B record per primitive in BVH leaf order (the order of
//             the reference's BVH::storage_ after its build), so leaf slot =
//             leaf rank; records carry the insertion id for reporting.
//   mats    : 64-B material records; lights: 48-B light records.
//
// A child ref >= 0 is an internal node index; a ref < 0 is a leaf holding
// primitive slot ~ref (one primitive per leaf, as the reference builds).
#pragma once

#include <stdint.h>

namespace pcl {

struct alignas(16) NodeRec {
    float lb[6];      // left child box  (min xyz, max xyz)
    float rb[6];      // right child box
    int32_t left, right, pad0, pad1;
};
static_assert(sizeof(NodeRec) == 64, "node is one 64-byte line");

// Child k of a 4-wide node: box (lx..hz)[k], ref child[k] (>= 0 a Node4
// index, < 0 a leaf ~slot).  Unused children have the empty box (+inf, -inf),
// which every slab test rejects, and ref kEmptyRoot.  The traversal reads the
// first 112 bytes; the last 16 pad the node to one 128-byte line.
struct alignas(16) Node4 {
    float lx[4], ly[4], lz[4], hx[4], hy[4], hz[4];
    int32_t child[4];
    int32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "4-wide node is one 128-byte line");

// Primitive payloads (g[]):
//   sphere   : center(0-2) radius(3)
//   triangle : v0(0-2) e1=v1-v0(3-5) e2=v2-v0(6-8) normal(9-11)
//   cube     : lo(0-2) hi(3-5)
//   plane    : point(0-2) unit normal(3-5)
//   cylinder : center(0-2) unit axis(3-5) radius(6) half height(7)
//   cone     : apex(0-2) unit axis(3-5) radius(6) height(7) base center(8-10)
struct alignas(16) PrimRec {
    float g[12];
    int32_t kind, mat, id, pad;
};
static_assert(sizeof(PrimRec) == 64, "primitive record is 64 bytes");

struct alignas(16) MatRec {
    int32_t kind;
    float albedo[3];
    float refl, ior, spec, shin;
    float spec_color[3];
    float emission[3];
};
static_assert(sizeof(MatRec) == 64, "material record is 64 bytes");

struct alignas(16) LightRec {
    int32_t kind;
    float pos[3];
    float I[3];
    float radius;
    int32_t emissive;
    int32_t pad[3];
};
static_assert(sizeof(LightRec) == 48, "light record is 48 bytes");

// Camera frame derived on the host exactly as the reference's Camera ctor
// (src/camera.hpp:29-36,67-86) does, with the host libm's tanf.
struct CamRec {
    float origin[3], lower_left[3], horizontal[3], vertical[3];
    float u[3], v[3];
    float aperture, focus_dist;
    int32_t blades, pad;
};

static const int32_t kEmptyRoot = (int32_t)0x80000000;

}  // namespace pcl
