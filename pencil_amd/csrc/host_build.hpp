// host_build.hpp - host side of scene upload: reference-topology BVH build,
// primitive/material/light flattening and the camera frame.
#pragma once

#include "pencil_hip.h"
#include "scene_layout.hpp"

#include <string>
#include <vector>

namespace pcl {

struct HostLayout {
    std::vector<NodeRec> nodes;       // binary tree (reference topology, preorder)
    std::vector<Node4> wide;          // what goes to the device (collapse_wide)
    int32_t wide_root = kEmptyRoot;   // Node4 index 0, a leaf ~slot, or kEmptyRoot
    int32_t stack_need = 0;           // deepest traversal stack over wide
    int32_t n_wide = 0;               // 4-wide nodes built on the device (use_bvh == 2)
    std::vector<PrimRec> prims;       // leaf order
    std::vector<int32_t> leaf_order;  // slot -> insertion id
    std::vector<int32_t> slot_of;     // insertion id -> slot (linear scan order)
    std::vector<MatRec> mats;
    std::vector<LightRec> lights;
    int32_t root = kEmptyRoot;
    float root_box[6] = {0, 0, 0, 0, 0, 0};
    int32_t depth = 0;
    // use_bvh == 2 (device builder): prims stay in insertion order, no nodes
    // are built here, and boxes (6 floats per primitive) go to the device
    std::vector<float> boxes;
};

// t[k], k = 1..255: the smallest float v in [0, 1] whose 8-bit encode
// (uint8)(powf(v, 1/2.2f) * 255 + 0.5) (src/image_io.hpp:53-55) is >= k,
// found with this host's powf; t[0] = -inf.
void tonemap_thresholds(float t[256]);

// Device material records of a pcl_material table; PCL_OK or an error code.
int pack_materials(const pcl_material* mats, int n, std::vector<MatRec>& out, std::string& err);

// 4-wide nodes from the binary tree: starting from a node's two children,
// the child with the largest box surface area is replaced by its two
// children until the cut holds four nodes or only leaves can be expanded no
// further.  Cut members stay in the binary tree's left-first order, so a
// traversal that visits them in slot order and defers each box's t_max check
// to the moment it is visited meets the leaves in the binary recursion's
// order with the same t_max at every check.  Output in preorder (first child
// right after its parent); need = the deepest stack such a traversal reaches
// (children 1..3 pushed before descending into child 0).
void collapse_wide(const std::vector<NodeRec>& bin, int32_t root, std::vector<Node4>& out, int32_t& out_root,
                   int32_t& need);

// Returns PCL_OK or an error code with a message in err.
int build_layout(const pcl_scene_desc* desc, HostLayout& out, std::string& err);

// Camera frame of src/camera.hpp:29-86 (focus_dist <= 0 -> |eye-look|, as
// ParsedScene::finalize does, src/pecl_parser.hpp:136-140).
CamRec make_camera(const pcl_camera& c);

}  // namespace pcl
