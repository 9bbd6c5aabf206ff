// host_build.cpp - reference-topology SAH BVH build and scene flattening.
//
// Compiled with g++ -ffp-contract=off so every float is the reference's.
//
// The BVH must have exactly the reference's topology (src/bvh.hpp:54-197):
// the closest hit among equal-t candidates is the last leaf reached in
// left-first DFS order, so primitive ids only match if leaves sit in the same
// order.  The build therefore follows the reference's steps on an index array:
// the range box as an fmin/fmax merge of per-primitive boxes, the longest
// axis, std::sort by doubled centroid for <= 4 primitives, 12-bucket binned
// SAH with cost 1 + (nL*SA_L + nR*SA_R)/SA_P otherwise, std::partition by
// bucket, median fallback, one primitive per leaf.  std::sort / std::partition
// from the same libstdc++ make the same element moves whatever the element
// type, since they only consult the comparator.
//
// What differs is only the representation: boxes are computed once per
// primitive instead of by a virtual aabb() call per visit, and the result is
// emitted as flat 64-byte nodes in preorder (left child right after its
// parent) instead of shared_ptr nodes.
#include "host_build.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

namespace pcl {
namespace {

struct V {
    float x = 0, y = 0, z = 0;
    V() = default;
    V(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    V operator+(const V& o) const { return {x + o.x, y + o.y, z + o.z}; }
    V operator-(const V& o) const { return {x - o.x, y - o.y, z - o.z}; }
    V operator*(float s) const { return {x * s, y * s, z * s}; }
    V operator/(float s) const { return {x / s, y / s, z / s}; }
    float dot(const V& o) const { return x * o.x + y * o.y + z * o.z; }
    V cross(const V& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
    float length() const { return std::sqrt(x * x + y * y + z * z); }
    V normalized() const { float l = length(); return (l > 1e-12f) ? (*this / l) : V{0, 0, 0}; }
};

struct Box {
    V mn{1e30f, 1e30f, 1e30f}, mx{-1e30f, -1e30f, -1e30f};
    void grow(const Box& b) {
        mn = {std::fmin(mn.x, b.mn.x), std::fmin(mn.y, b.mn.y), std::fmin(mn.z, b.mn.z)};
        mx = {std::fmax(mx.x, b.mx.x), std::fmax(mx.y, b.mx.y), std::fmax(mx.z, b.mx.z)};
    }
    int longest_axis() const {
        V e = mx - mn;
        if (e.x >= e.y && e.x >= e.z) return 0;
        if (e.y >= e.z) return 1;
        return 2;
    }
    float area() const {
        V e = mx - mn;
        return 2.0f * (e.x * e.y + e.y * e.z + e.x * e.z);
    }
};

V vec(const float* p) { return {p[0], p[1], p[2]}; }
void put(float* d, const V& v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

// Per-kind aabb() (src/primitives.hpp:122-125,172-181,240-243,275-278,424-432,532-545)
Box prim_box(const PrimRec& r) {
    Box b;
    const float* g = r.g;
    switch (r.kind) {
    case PCL_PRIM_SPHERE: {
        float s = g[3];
        b.mn = vec(g) - V{s, s, s}; b.mx = vec(g) + V{s, s, s};
        break;
    }
    case PCL_PRIM_TRIANGLE:
        // needs the raw vertices (v0 + e1 may round differently from v1):
        // computed from the constructor arguments in build_layout
        break;
    case PCL_PRIM_CUBE: b.mn = vec(g); b.mx = vec(g + 3); break;
    case PCL_PRIM_PLANE: b.mn = {-1e6f, -1e6f, -1e6f}; b.mx = {1e6f, 1e6f, 1e6f}; break;
    case PCL_PRIM_CYLINDER: {
        V h = vec(g + 3) * g[7];
        V e = {std::fabs(h.x) + g[6], std::fabs(h.y) + g[6], std::fabs(h.z) + g[6]};
        b.mn = vec(g) - e; b.mx = vec(g) + e;
        break;
    }
    case PCL_PRIM_CONE: {
        V apex = vec(g), base = vec(g + 8);
        float r = g[6];
        b.mn = {std::fmin(apex.x, base.x) - r, std::fmin(apex.y, base.y) - r, std::fmin(apex.z, base.z) - r};
        b.mx = {std::fmax(apex.x, base.x) + r, std::fmax(apex.y, base.y) + r, std::fmax(apex.z, base.z) + r};
        break;
    }
    }
    return b;
}

struct Builder {
    std::vector<int32_t> storage;   // insertion ids, permuted as the build goes
    std::vector<Box> pbox;          // per insertion id
    std::vector<NodeRec>* nodes;
    int last_split = 0;
    int max_depth = 0;

    // BVH::median_split (src/bvh.hpp:106-115)
    void median_split(int start, int end, int axis) {
        const std::vector<Box>& pb = pbox;
        std::sort(storage.begin() + start, storage.begin() + end, [axis, &pb](int32_t a, int32_t b) {
            const Box& ba = pb[a];
            const Box& bb = pb[b];
            return (ba.mn[axis] + ba.mx[axis]) < (bb.mn[axis] + bb.mx[axis]);
        });
        last_split = start + (end - start) / 2;
    }

    // BVH::sah_split (src/bvh.hpp:119-192)
    void sah_split(int start, int end, const Box& parent, int axis) {
        constexpr int NB = 12;
        int cnt[NB] = {0};
        Box bnd[NB];
        const float lo = parent.mn[axis], hi = parent.mx[axis];
        if (hi <= lo) { median_split(start, end, axis); return; }
        auto bucket_of = [&](const Box& pb) {
            float center = (pb.mn[axis] + pb.mx[axis]) * 0.5f;
            int b = (int)((center - lo) / (hi - lo) * NB);
            if (b < 0) b = 0;
            if (b >= NB) b = NB - 1;
            return b;
        };
        for (int i = start; i < end; ++i) {
            const Box& pb = pbox[storage[i]];
            int b = bucket_of(pb);
            cnt[b]++;
            bnd[b].grow(pb);
        }
        Box pre[NB + 1], suf[NB + 1];
        int prec[NB + 1] = {0}, sufc[NB + 1] = {0};
        for (int i = 0; i < NB; ++i) { pre[i + 1] = pre[i]; pre[i + 1].grow(bnd[i]); prec[i + 1] = prec[i] + cnt[i]; }
        for (int i = NB - 1; i >= 0; --i) { suf[i] = suf[i + 1]; suf[i].grow(bnd[i]); sufc[i] = sufc[i + 1] + cnt[i]; }
        float best_cost = FLT_MAX;
        int best = -1;
        const float sa_p = parent.area();
        for (int i = 1; i < NB; ++i) {
            int nL = prec[i], nR = sufc[i];
            if (nL == 0 || nR == 0) continue;
            float saL = pre[i].area(), saR = suf[i].area();
            float cost = 1.0f + (nL * saL + nR * saR) / (sa_p > 0 ? sa_p : 1.0f);
            if (cost < best_cost) { best_cost = cost; best = i; }
        }
        if (best < 0) { median_split(start, end, axis); return; }
        auto pivot = std::partition(storage.begin() + start, storage.begin() + end,
                                    [&](int32_t p) { return bucket_of(pbox[p]) < best; });
        last_split = (int)(pivot - storage.begin());
        if (last_split <= start || last_split >= end) median_split(start, end, axis);
    }

    // BVH::build_range (src/bvh.hpp:74-104); returns the child ref and box.
    int32_t build(int start, int end, int depth, Box& box_out) {
        Box box;
        for (int i = start; i < end; ++i) box.grow(pbox[storage[i]]);
        box_out = box;
        if (depth > max_depth) max_depth = depth;
        int count = end - start;
        if (count == 1) return ~(int32_t)start;   // leaf: slot = storage index
        int axis = box.longest_axis();
        if (count <= 4) median_split(start, end, axis);
        else sah_split(start, end, box, axis);
        int split = start + count / 2;
        if (count > 4) split = last_split;
        int32_t idx = (int32_t)nodes->size();
        nodes->push_back(NodeRec{});
        Box lb, rb;
        int32_t l = build(start, split, depth + 1, lb);
        int32_t r = build(split, end, depth + 1, rb);
        NodeRec& n = (*nodes)[idx];
        put(n.lb, lb.mn); put(n.lb + 3, lb.mx);
        put(n.rb, rb.mn); put(n.rb + 3, rb.mx);
        n.left = l; n.right = r;
        return idx;
    }
};

}  // namespace

namespace {

struct CutItem {
    int32_t ref;
    const float* box;   // 6 floats in a NodeRec (lb or rb)
};

float surface_area(const float* b) {
    const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

struct Collapser {
    const std::vector<NodeRec>& bin;
    std::vector<Node4>& out;

    int32_t emit(int32_t r, int32_t& need) {
        CutItem cut[4];
        int cnt = 2;
        cut[0] = {bin[r].left, bin[r].lb};
        cut[1] = {bin[r].right, bin[r].rb};
        while (cnt < 4) {
            int j = -1;
            float best = -1.0f;
            for (int k = 0; k < cnt; ++k)
                if (cut[k].ref >= 0) {
                    const float a = surface_area(cut[k].box);
                    if (a > best) { best = a; j = k; }
                }
            if (j < 0) break;
            const NodeRec& m = bin[cut[j].ref];
            for (int k = cnt; k > j + 1; --k) cut[k] = cut[k - 1];
            cut[j] = {m.left, m.lb};
            cut[j + 1] = {m.right, m.rb};
            ++cnt;
        }
        const int32_t idx = (int32_t)out.size();
        out.push_back(Node4{});
        int32_t ch[4] = {kEmptyRoot, kEmptyRoot, kEmptyRoot, kEmptyRoot};
        need = 0;
        for (int k = 0; k < cnt; ++k) {
            int32_t sub = 0;
            ch[k] = cut[k].ref >= 0 ? emit(cut[k].ref, sub) : cut[k].ref;
            need = std::max(need, (cnt - 1 - k) + sub);
        }
        Node4& o = out[idx];
        for (int k = 0; k < 4; ++k) {
            const bool used = k < cnt;
            o.lx[k] = used ? cut[k].box[0] : INFINITY;
            o.ly[k] = used ? cut[k].box[1] : INFINITY;
            o.lz[k] = used ? cut[k].box[2] : INFINITY;
            o.hx[k] = used ? cut[k].box[3] : -INFINITY;
            o.hy[k] = used ? cut[k].box[4] : -INFINITY;
            o.hz[k] = used ? cut[k].box[5] : -INFINITY;
            o.child[k] = ch[k];
            o.pad[k] = 0;
        }
        return idx;
    }
};

}  // namespace

void collapse_wide(const std::vector<NodeRec>& bin, int32_t root, std::vector<Node4>& out, int32_t& out_root,
                   int32_t& need) {
    out.clear();
    need = 0;
    if (root < 0) { out_root = root; return; }   // a leaf or an empty scene
    out.reserve(bin.size() / 2 + 1);
    Collapser c{bin, out};
    out_root = c.emit(root, need);
}

void tonemap_thresholds(float t[256]) {
    auto enc = [](float v) { return (int)(uint8_t)(std::pow(v, 1.0f / 2.2f) * 255.0f + 0.5f); };
    t[0] = -INFINITY;
    for (int k = 1; k < 256; ++k) {
        uint32_t lo = 0, hi = 0x3f800000u;   // bit patterns of [0, 1]; enc(1) = 255
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            float v;
            std::memcpy(&v, &mid, 4);
            if (enc(v) >= k) hi = mid; else lo = mid + 1;
        }
        std::memcpy(&t[k], &lo, 4);
    }
}

int pack_materials(const pcl_material* mats, int n, std::vector<MatRec>& out, std::string& err) {
    out.clear();
    for (int i = 0; i < n; ++i) {
        const pcl_material& m = mats[i];
        if (m.kind < PCL_MAT_DIFFUSE || m.kind > PCL_MAT_DIELECTRIC) {
            err = "material " + std::to_string(i) + " needs a feature outside the HIP core (principled BSDF, "
                  "texture or dispersion)";
            return PCL_ERR_UNSUPPORTED;
        }
        MatRec r{};
        r.kind = m.kind;
        for (int c = 0; c < 3; ++c) { r.albedo[c] = m.albedo[c]; r.spec_color[c] = m.specular_color[c]; r.emission[c] = m.emission[c]; }
        r.refl = m.reflectivity; r.ior = m.ior; r.spec = m.specular; r.shin = m.shininess;
        out.push_back(r);
    }
    return PCL_OK;
}

int build_layout(const pcl_scene_desc* d, HostLayout& out, std::string& err) {
    if (!d || d->n_prims < 0 || d->n_materials < 0 || d->n_lights < 0 ||
        (d->n_prims > 0 && !d->prims) || (d->n_materials > 0 && !d->materials) || (d->n_lights > 0 && !d->lights)) {
        err = "invalid scene description";
        return PCL_ERR_INVALID;
    }
    const int n = d->n_prims;
    std::vector<PrimRec> recs((size_t)n);
    std::vector<Box> boxes((size_t)n);
    for (int i = 0; i < n; ++i) {
        const pcl_prim& p = d->prims[i];
        if (p.material_id < 0 || p.material_id >= d->n_materials) {
            err = "primitive " + std::to_string(i) + " has material id " + std::to_string(p.material_id) +
                  " outside [0, " + std::to_string(d->n_materials) + ")";
            return PCL_ERR_INVALID;
        }
        const int kind = p.kind & 0xFF;
        const bool stored = (p.kind & PCL_PRIM_STORED) != 0;   // unit vectors already normalized
        if (p.kind & ~(0xFF | PCL_PRIM_STORED)) {
            err = "primitive " + std::to_string(i) + " has unknown kind flags";
            return PCL_ERR_INVALID;
        }
        auto unit = [stored](const V& v) { return stored ? v : v.normalized(); };
        PrimRec r{};
        r.kind = kind; r.mat = p.material_id; r.id = i;
        V p0 = vec(p.p), p1 = vec(p.p + 3), p2 = vec(p.p + 6);
        Box b;
        switch (kind) {
        case PCL_PRIM_SPHERE:
            put(r.g, p0); r.g[3] = p.p[3];
            b = prim_box(r);
            break;
        case PCL_PRIM_TRIANGLE: {
            V e1 = p1 - p0, e2 = p2 - p0;
            put(r.g, p0); put(r.g + 3, e1); put(r.g + 6, e2);
            put(r.g + 9, e1.cross(e2).normalized());   // Triangle ctor, primitives.hpp:140-143
            b.mn = {std::fmin(std::fmin(p0.x, p1.x), p2.x), std::fmin(std::fmin(p0.y, p1.y), p2.y),
                    std::fmin(std::fmin(p0.z, p1.z), p2.z)};
            b.mx = {std::fmax(std::fmax(p0.x, p1.x), p2.x), std::fmax(std::fmax(p0.y, p1.y), p2.y),
                    std::fmax(std::fmax(p0.z, p1.z), p2.z)};
            break;
        }
        case PCL_PRIM_CUBE: {
            float h = p.p[3];
            put(r.g, p0 - V{h, h, h}); put(r.g + 3, p0 + V{h, h, h});
            b = prim_box(r);
            break;
        }
        case PCL_PRIM_PLANE:
            put(r.g, p0); put(r.g + 3, unit(p1));
            b = prim_box(r);
            break;
        case PCL_PRIM_CYLINDER:
            put(r.g, p0); put(r.g + 3, unit(p1)); r.g[6] = p.p[6]; r.g[7] = p.p[7] * 0.5f;
            b = prim_box(r);
            break;
        case PCL_PRIM_CONE: {
            V ax = unit(p1);
            put(r.g, p0); put(r.g + 3, ax); r.g[6] = p.p[6]; r.g[7] = p.p[7];
            put(r.g + 8, p0 - ax * p.p[7]);   // Cone ctor, primitives.hpp:511-515
            b = prim_box(r);
            break;
        }
        default:
            err = "primitive " + std::to_string(i) + " has unknown kind " + std::to_string(p.kind);
            return PCL_ERR_INVALID;
        }
        recs[i] = r;
        boxes[i] = b;
    }
    {
        const int rc = pack_materials(d->materials, d->n_materials, out.mats, err);
        if (rc != PCL_OK) return rc;
    }
    out.lights.clear();
    for (int i = 0; i < d->n_lights; ++i) {
        const pcl_light& l = d->lights[i];
        if (l.kind != PCL_LIGHT_POINT && l.kind != PCL_LIGHT_DIRECTIONAL) {
            err = "light " + std::to_string(i) + " has unknown kind";
            return PCL_ERR_INVALID;
        }
        LightRec r{};
        r.kind = l.kind;
        for (int c = 0; c < 3; ++c) { r.pos[c] = l.pos_or_dir[c]; r.I[c] = l.intensity[c]; }
        r.radius = l.radius; r.emissive = l.is_emissive ? 1 : 0;
        out.lights.push_back(r);
    }

    if (d->use_bvh == 2) {   // the device builds the tree (lbvh.hip)
        out.nodes.clear();
        out.boxes.resize((size_t)n * 6);
        for (int i = 0; i < n; ++i) {
            put(&out.boxes[(size_t)i * 6], boxes[i].mn);
            put(&out.boxes[(size_t)i * 6 + 3], boxes[i].mx);
        }
        out.prims = std::move(recs);
        out.leaf_order.resize((size_t)n);
        out.slot_of.resize((size_t)n);
        for (int i = 0; i < n; ++i) out.leaf_order[i] = out.slot_of[i] = i;
        out.root = n > 0 ? (n == 1 ? ~0 : 0) : kEmptyRoot;
        // n >= 2: the device builds the nodes and sets the wide root; n <= 1
        // needs no node (a single leaf, or nothing)
        out.wide_root = n >= 2 ? 0 : out.root;
        if (n > 0) { put(out.root_box, boxes[0].mn); put(out.root_box + 3, boxes[0].mx); }
        out.depth = 0;
        return PCL_OK;
    }
    if (d->use_bvh != 0 && d->use_bvh != 1) {
        err = "use_bvh must be 0 (linear scan), 1 (reference SAH build) or 2 (device LBVH build)";
        return PCL_ERR_INVALID;
    }
    Builder B;
    B.nodes = &out.nodes;
    B.storage.resize((size_t)n);
    for (int i = 0; i < n; ++i) B.storage[i] = i;
    B.pbox = std::move(boxes);
    out.nodes.clear();
    out.nodes.reserve(n > 1 ? (size_t)n - 1 : 0);
    if (n > 0) {
        Box rb;
        out.root = B.build(0, n, 0, rb);
        put(out.root_box, rb.mn); put(out.root_box + 3, rb.mx);
    } else {
        out.root = kEmptyRoot;
    }
    out.depth = B.max_depth;
    collapse_wide(out.nodes, out.root, out.wide, out.wide_root, out.stack_need);
    out.prims.resize((size_t)n);
    out.leaf_order.resize((size_t)n);
    out.slot_of.resize((size_t)n);
    for (int s = 0; s < n; ++s) {
        int32_t id = B.storage[s];
        out.prims[s] = recs[id];
        out.leaf_order[s] = id;
        out.slot_of[id] = s;
    }
    return PCL_OK;
}

CamRec make_camera(const pcl_camera& c) {
    V eye = vec(c.eye), look = vec(c.look), up = vec(c.up);
    float fd = c.focus_dist;
    if (fd <= 0.0f) fd = (eye - look).length();
    V w = (eye - look).normalized();
    V u = up.cross(w).normalized();
    V v = w.cross(u);
    float theta = c.vfov_deg * 3.14159265358979323846f / 180.0f;
    float half_h = std::tan(theta * 0.5f);
    float half_w = c.aspect * half_h;
    V horizontal = u * (2.0f * half_w * fd);
    V vertical = v * (2.0f * half_h * fd);
    V lower_left = eye - horizontal * 0.5f - vertical * 0.5f - w * fd;
    CamRec r{};
    put(r.origin, eye); put(r.lower_left, lower_left); put(r.horizontal, horizontal); put(r.vertical, vertical);
    put(r.u, u); put(r.v, v);
    r.aperture = c.aperture; r.focus_dist = fd; r.blades = c.bokeh_blades;
    return r;
}

}  // namespace pcl
