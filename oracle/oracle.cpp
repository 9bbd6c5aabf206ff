// oracle.cpp - CPU oracle for the Pencil rendering hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain C++17 restatement of the
// reference renderer's hot path, written from its algorithm (the reference
// files and lines are cited per function).  It is the checker that the HIP
// core is compared against, and the CPU baseline bench.py may time.  Nothing in
// the product links or calls it.
//
// Pinning (what fixes each part independently of this restatement):
//   - Rng, Camera, primitive hits, BVH build + traversal, trace(), ACES/gamma:
//     the reference's own golden PNGs (tests/*.png of the reference) rendered
//     through this oracle match bit-exactly or within 1 LSB, and fixtures from a
//     harness that links the reference headers pin HDR values, primary-hit
//     (t, prim id), BVH leaf order and ray counts.
//   - trace_path(): harness HDR fixtures at small sizes (the reference's
//     path-trace goldens are stale).
//   - xorshift32 first outputs checked by hand (Marsaglia 2003 shifts 13/17/5).
//
// Build: g++ -O2 -std=c++17 -ffp-contract=off (no FMA contraction, like the
// reference's x86-64 SSE build), so float results are the reference's.

#include "pencil_hip.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace orc {

// ---- vec3.hpp:24-92 -------------------------------------------------------
struct V3 {
    float x = 0, y = 0, z = 0;
    V3() = default;
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator-() const { return {-x, -y, -z}; }
    V3 operator*(float s) const { return {x * s, y * s, z * s}; }
    V3 operator/(float s) const { return {x / s, y / s, z / s}; }
    V3 operator*(const V3& o) const { return {x * o.x, y * o.y, z * o.z}; }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    float dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    V3 cross(const V3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
    float length_squared() const { return x * x + y * y + z * z; }
    float length() const { return std::sqrt(length_squared()); }
    V3 normalized() const { float l = length(); return (l > 1e-12f) ? (*this / l) : V3{0, 0, 0}; }
    V3 reflect(const V3& n) const { return *this - n * (2.0f * this->dot(n)); }
};
static inline float clampf(float v, float lo, float hi) { return (v < lo) ? lo : (v > hi) ? hi : v; }

// Transcendentals of the shading path.  Mode 0 (default): the float libm
// functions the reference calls (glibc cosf / sinf / powf / atan2f,
// src/rng.hpp:42-43, src/vec3.hpp:86, src/renderer.hpp:564,941,959,
// src/camera.cpp:123-139).  Mode 1: the same function evaluated in double and
// rounded once to float, i.e. the correctly rounded result except in rare
// double-rounding cases -- the arithmetic the HIP core uses.  Mode 1 lets a
// test compare the core with this restatement bit for bit where glibc's
// not-correctly-rounded float results would otherwise decorrelate a path.
static int g_trig_mode = 0;
static inline float o_cos(float x) { return g_trig_mode ? (float)std::cos((double)x) : std::cos(x); }
static inline float o_sin(float x) { return g_trig_mode ? (float)std::sin((double)x) : std::sin(x); }
static inline float o_pow(float x, float y) { return g_trig_mode ? (float)std::pow((double)x, (double)y) : std::pow(x, y); }
static inline float o_atan2(float y, float x) { return g_trig_mode ? (float)std::atan2((double)y, (double)x) : std::atan2(y, x); }
// vec3.hpp:83-87
static inline float schlick(float c, float r) {
    float r0 = (1.0f - r) / (1.0f + r);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * o_pow(1.0f - c, 5.0f);
}

// ---- rng.hpp:13-45 --------------------------------------------------------
struct Rng {
    uint32_t s;
    explicit Rng(uint32_t v) : s(v) {}
    uint32_t u32() { s ^= s << 13u; s ^= s >> 17u; s ^= s << 5u; return s; }
    float unit() { return u32() / 4294967296.0f; }
    float sym() { return unit() * 2.0f - 1.0f; }
    void unit_disc(float& dx, float& dy) {
        float r1 = unit(), r2 = unit();
        float sx = 2.0f * r1 - 1.0f, sy = 2.0f * r2 - 1.0f;
        if (sx == 0.0f && sy == 0.0f) { dx = 0; dy = 0; return; }
        float theta, r;
        if (std::fabs(sx) > std::fabs(sy)) { r = sx; theta = (3.14159265f / 4.0f) * (sy / sx); }
        else { r = sy; theta = (3.14159265f / 2.0f) - (3.14159265f / 4.0f) * (sx / sy); }
        dx = r * o_cos(theta);
        dy = r * o_sin(theta);
    }
};

struct Ray { V3 o, d; V3 at(float t) const { return o + d * t; } };

// ---- primitives.hpp:31-81 -------------------------------------------------
struct AABB {
    V3 mn{1e30f, 1e30f, 1e30f}, mx{-1e30f, -1e30f, -1e30f};
    bool hit(const Ray& r, float t_min, float t_max) const {
        for (int a = 0; a < 3; ++a) {
            float inv_d = 1.0f / r.d[a];
            float t0 = (mn[a] - r.o[a]) * inv_d;
            float t1 = (mx[a] - r.o[a]) * inv_d;
            if (inv_d < 0.0f) std::swap(t0, t1);
            t_min = (t0 > t_min) ? t0 : t_min;
            t_max = (t1 < t_max) ? t1 : t_max;
            if (t_max < t_min) return false;
        }
        return true;
    }
    static AABB merge(const AABB& a, const AABB& b) {
        AABB o;
        o.mn = {std::fmin(a.mn.x, b.mn.x), std::fmin(a.mn.y, b.mn.y), std::fmin(a.mn.z, b.mn.z)};
        o.mx = {std::fmax(a.mx.x, b.mx.x), std::fmax(a.mx.y, b.mx.y), std::fmax(a.mx.z, b.mx.z)};
        return o;
    }
    int longest_axis() const {
        V3 e = mx - mn;
        if (e.x >= e.y && e.x >= e.z) return 0;
        if (e.y >= e.z) return 1;
        return 2;
    }
};

struct Hit {
    V3 p{0, 0, 0}, normal{0, 1, 0};
    float t = -1.0f;
    int mat = 0;
    bool front = true;
    void set_face_normal(const Ray& r, const V3& n) {
        front = r.d.dot(n) < 0.0f;
        normal = front ? n : -n;
    }
};

// One primitive; fields follow each reference constructor.
struct Prim {
    int kind = 0, mat = 0, id = 0;
    V3 a, b, c;       // sphere: a=center; tri: v0 v1 v2; cube: a=center; plane: a=point b=normal;
                      // cyl: a=center b=axis; cone: a=apex b=axis c=base_center
    float f0 = 0, f1 = 0;  // sphere radius; cube half; cyl radius, half_h; cone radius, height
    V3 n;             // triangle normal (primitives.hpp:142)

    static Prim make(const pcl_prim& p, int id) {
        Prim q; q.kind = p.kind; q.mat = p.material_id; q.id = id;
        V3 p0{p.p[0], p.p[1], p.p[2]}, p1{p.p[3], p.p[4], p.p[5]}, p2{p.p[6], p.p[7], p.p[8]};
        switch (p.kind) {
        case PCL_PRIM_SPHERE: q.a = p0; q.f0 = p.p[3]; break;
        case PCL_PRIM_TRIANGLE: q.a = p0; q.b = p1; q.c = p2; q.n = (p1 - p0).cross(p2 - p0).normalized(); break;
        case PCL_PRIM_CUBE: q.a = p0; q.f0 = p.p[3]; break;
        case PCL_PRIM_PLANE: q.a = p0; q.b = p1.normalized(); break;
        case PCL_PRIM_CYLINDER: q.a = p0; q.b = p1.normalized(); q.f0 = p.p[6]; q.f1 = p.p[7] * 0.5f; break;
        case PCL_PRIM_CONE: q.a = p0; q.b = p1.normalized(); q.f0 = p.p[6]; q.f1 = p.p[7];
            q.c = q.a - q.b * q.f1; break;
        }
        return q;
    }

    // Sphere::hit, primitives.hpp:98-120
    bool hit_sphere(const Ray& r, float t_min, float t_max, Hit& out) const {
        V3 oc = r.o - a;
        float A = r.d.dot(r.d);
        float half_b = oc.dot(r.d);
        float C = oc.dot(oc) - f0 * f0;
        float disc = half_b * half_b - A * C;
        if (disc < 0.0f) return false;
        float sq = std::sqrt(disc);
        float t = (-half_b - sq) / A;
        if (t < t_min || t > t_max) {
            t = (-half_b + sq) / A;
            if (t < t_min || t > t_max) return false;
        }
        out.t = t; out.p = r.at(t);
        V3 on = (out.p - a) / f0;
        out.set_face_normal(r, on); out.mat = mat;
        return true;
    }
    // Triangle::hit (Moller-Trumbore), primitives.hpp:145-170
    bool hit_tri(const Ray& r, float t_min, float t_max, Hit& out) const {
        V3 e1 = b - a, e2 = c - a;
        V3 h = r.d.cross(e2);
        float A = e1.dot(h);
        if (std::fabs(A) < 1e-8f) return false;
        float f = 1.0f / A;
        V3 s = r.o - a;
        float u = f * s.dot(h);
        if (u < 0.0f || u > 1.0f) return false;
        V3 q = s.cross(e1);
        float v = f * r.d.dot(q);
        if (v < 0.0f || u + v > 1.0f) return false;
        float t = f * e2.dot(q);
        if (t < t_min || t > t_max) return false;
        out.t = t; out.p = r.at(t); out.set_face_normal(r, n); out.mat = mat;
        return true;
    }
    // Cube::hit, primitives.hpp:201-238
    bool hit_cube(const Ray& r, float t_min, float t_max, Hit& out) const {
        V3 lo = a - V3{f0, f0, f0}, hi = a + V3{f0, f0, f0};
        float t_near = -INFINITY, t_far = INFINITY;
        int near_axis = -1, far_axis = -1;
        float near_sign = 1.0f, far_sign = -1.0f;
        for (int ax = 0; ax < 3; ++ax) {
            const float d = r.d[ax];
            if (std::fabs(d) < 1e-12f) {
                if (r.o[ax] < lo[ax] || r.o[ax] > hi[ax]) return false;
                continue;
            }
            float tn = (lo[ax] - r.o[ax]) / d, tf = (hi[ax] - r.o[ax]) / d;
            float ns = -1.0f, fs = 1.0f;
            if (tn > tf) { std::swap(tn, tf); std::swap(ns, fs); }
            if (tn > t_near) { t_near = tn; near_axis = ax; near_sign = ns; }
            if (tf < t_far) { t_far = tf; far_axis = ax; far_sign = fs; }
            if (t_far < t_near) return false;
        }
        const bool use_near = t_near >= t_min;
        const float t_hit = use_near ? t_near : t_far;
        const int hit_axis = use_near ? near_axis : far_axis;
        const float hit_sign = use_near ? near_sign : far_sign;
        if (hit_axis < 0 || t_hit < t_min || t_hit > t_max) return false;
        out.t = t_hit; out.p = r.at(t_hit);
        V3 nrm{0, 0, 0}; nrm[hit_axis] = hit_sign;
        out.set_face_normal(r, nrm); out.mat = mat;
        return true;
    }
    // Plane::hit, primitives.hpp:261-272
    bool hit_plane(const Ray& r, float t_min, float t_max, Hit& out) const {
        float denom = b.dot(r.d);
        if (std::fabs(denom) < 1e-6f) return false;
        float t = (a - r.o).dot(b) / denom;
        if (t < t_min || t > t_max) return false;
        out.t = t; out.p = r.at(t); out.set_face_normal(r, b); out.mat = mat;
        return true;
    }
    // Cylinder::hit_side, primitives.hpp:441-486
    bool cyl_side(const Ray& r, float t_min, float t_max, Hit& out) const {
        V3 oc = r.o - a;
        float dda = r.d.dot(b), oca = oc.dot(b);
        V3 dp = r.d - b * dda, ocp = oc - b * oca;
        float A = dp.dot(dp), B = 2.0f * ocp.dot(dp), C = ocp.dot(ocp) - f0 * f0;
        if (A < 1e-9f) {
            if (std::fabs(B) < 1e-9f) return false;
            const float t = -C / B;
            if (t < t_min || t > t_max) return false;
            const float h = oca + t * dda;
            if (std::fabs(h) > f1) return false;
            out.t = t; out.p = r.at(t);
            V3 axis_pt = a + b * h;
            V3 on = (out.p - axis_pt) / f0;
            out.set_face_normal(r, on); out.mat = mat;
            return true;
        }
        float disc = B * B - 4.0f * A * C;
        if (disc < 0.0f) return false;
        float sq = std::sqrt(disc);
        float ts[2] = {(-B - sq) / (2.0f * A), (-B + sq) / (2.0f * A)};
        for (float t : ts) {
            if (t < t_min || t > t_max) continue;
            float h = (oca + t * dda);
            if (std::fabs(h) > f1) continue;
            out.t = t; out.p = r.at(t);
            V3 axis_pt = a + b * h;
            V3 on = (out.p - axis_pt) / f0;
            out.set_face_normal(r, on); out.mat = mat;
            return true;
        }
        return false;
    }
    // Cylinder::hit_cap, primitives.hpp:488-504
    bool cyl_cap(const Ray& r, float t_min, float t_max, Hit& out, float sign) const {
        V3 cc = a + b * (sign * f1);
        V3 nrm = b * sign;
        float denom = nrm.dot(r.d);
        if (std::fabs(denom) < 1e-6f) return false;
        float t = (cc - r.o).dot(nrm) / denom;
        if (t < t_min || t > t_max) return false;
        V3 p = r.at(t);
        V3 d = p - cc;
        if (d.dot(d) > f0 * f0) return false;
        out.t = t; out.p = p; out.set_face_normal(r, nrm); out.mat = mat;
        return true;
    }
    // Cylinder::hit, primitives.hpp:406-422
    bool hit_cyl(const Ray& r, float t_min, float t_max, Hit& out) const {
        bool found = false; float closest = t_max; Hit best;
        if (cyl_side(r, t_min, closest, best)) { closest = best.t; out = best; found = true; }
        if (cyl_cap(r, t_min, closest, best, +1.0f)) { closest = best.t; out = best; found = true; }
        if (cyl_cap(r, t_min, closest, best, -1.0f)) { closest = best.t; out = best; found = true; }
        return found;
    }
    // Cone::hit_side, primitives.hpp:555-589
    bool cone_side(const Ray& r, float t_min, float t_max, Hit& out) const {
        V3 oc = r.o - a;
        float dda = r.d.dot(b), oca = oc.dot(b);
        float k = f0 / f1, k2 = k * k;
        V3 dp = r.d - b * dda, ocp = oc - b * oca;
        float A = dp.dot(dp) - k2 * dda * dda;
        float B = 2.0f * (dp.dot(ocp) - k2 * dda * oca);
        float C = ocp.dot(ocp) - k2 * oca * oca;
        float disc = B * B - 4.0f * A * C;
        if (disc < 0.0f || std::fabs(A) < 1e-12f) return false;
        float sq = std::sqrt(disc);
        float ts[2] = {(-B - sq) / (2.0f * A), (-B + sq) / (2.0f * A)};
        for (float t : ts) {
            if (t < t_min || t > t_max) continue;
            float h = oca + t * dda;
            if (h < 0.0f || h > f1) continue;
            out.t = t; out.p = r.at(t);
            V3 axis_pt = a + b * h;
            V3 radial = out.p - axis_pt;
            V3 on = (radial.normalized() - b * k).normalized();
            out.set_face_normal(r, on); out.mat = mat;
            return true;
        }
        return false;
    }
    // Cone::hit_base, primitives.hpp:591-605
    bool cone_base(const Ray& r, float t_min, float t_max, Hit& out) const {
        V3 nrm = -b;
        float denom = nrm.dot(r.d);
        if (std::fabs(denom) < 1e-6f) return false;
        float t = (c - r.o).dot(nrm) / denom;
        if (t < t_min || t > t_max) return false;
        V3 p = r.at(t);
        V3 d = p - c;
        if (d.dot(d) > f0 * f0) return false;
        out.t = t; out.p = p; out.set_face_normal(r, nrm); out.mat = mat;
        return true;
    }
    // Cone::hit, primitives.hpp:517-530
    bool hit_cone(const Ray& r, float t_min, float t_max, Hit& out) const {
        bool found = false; float closest = t_max; Hit best;
        if (cone_side(r, t_min, closest, best)) { closest = best.t; out = best; found = true; }
        if (cone_base(r, t_min, closest, best)) { closest = best.t; out = best; found = true; }
        return found;
    }
    bool hit(const Ray& r, float t_min, float t_max, Hit& out) const {
        switch (kind) {
        case PCL_PRIM_SPHERE: return hit_sphere(r, t_min, t_max, out);
        case PCL_PRIM_TRIANGLE: return hit_tri(r, t_min, t_max, out);
        case PCL_PRIM_CUBE: return hit_cube(r, t_min, t_max, out);
        case PCL_PRIM_PLANE: return hit_plane(r, t_min, t_max, out);
        case PCL_PRIM_CYLINDER: return hit_cyl(r, t_min, t_max, out);
        case PCL_PRIM_CONE: return hit_cone(r, t_min, t_max, out);
        }
        return false;
    }
    // aabb() of each type: primitives.hpp:122-125,172-181,240-243,275-278,424-432,532-545
    AABB aabb() const {
        AABB o;
        switch (kind) {
        case PCL_PRIM_SPHERE: o.mn = a - V3{f0, f0, f0}; o.mx = a + V3{f0, f0, f0}; break;
        case PCL_PRIM_TRIANGLE:
            o.mn = {std::fmin(std::fmin(a.x, b.x), c.x), std::fmin(std::fmin(a.y, b.y), c.y), std::fmin(std::fmin(a.z, b.z), c.z)};
            o.mx = {std::fmax(std::fmax(a.x, b.x), c.x), std::fmax(std::fmax(a.y, b.y), c.y), std::fmax(std::fmax(a.z, b.z), c.z)};
            break;
        case PCL_PRIM_CUBE: o.mn = a - V3{f0, f0, f0}; o.mx = a + V3{f0, f0, f0}; break;
        case PCL_PRIM_PLANE: o.mn = {-1e6f, -1e6f, -1e6f}; o.mx = {1e6f, 1e6f, 1e6f}; break;
        case PCL_PRIM_CYLINDER: {
            V3 h = b * f1;
            V3 e = {std::fabs(h.x) + f0, std::fabs(h.y) + f0, std::fabs(h.z) + f0};
            o.mn = a - e; o.mx = a + e; break;
        }
        case PCL_PRIM_CONE:
            o.mn = {std::fmin(a.x, c.x) - f0, std::fmin(a.y, c.y) - f0, std::fmin(a.z, c.z) - f0};
            o.mx = {std::fmax(a.x, c.x) + f0, std::fmax(a.y, c.y) + f0, std::fmax(a.z, c.z) + f0};
            break;
        }
        return o;
    }
};

// Per-query work counters (node visits = internal nodes whose box passed,
// prim tests per kind), used for the algorithmic-bytes roofline.
struct Counters {
    uint64_t nodes = 0;   // 4-wide nodes entered (heads, see BVH::mark_wide)
    uint64_t tests[6] = {0, 0, 0, 0, 0, 0};
};

// ---- bvh.hpp:22-202 -------------------------------------------------------
struct BVH {
    struct Node { AABB box; int leaf = -1; int left = -1, right = -1; };
    std::vector<int> storage;   // prim indices (insertion ids), reordered by the build
    std::vector<Node> nodes;
    std::vector<AABB> pbox;     // aabb() per insertion id (pure function of the prim)
    int root = -1;
    int last_split = 0;
    std::vector<int> splits;    // preorder split points, for topology fixtures

    void build(const std::vector<Prim>& prims) {
        storage.resize(prims.size());
        pbox.resize(prims.size());
        for (size_t i = 0; i < prims.size(); ++i) { storage[i] = (int)i; pbox[i] = prims[i].aabb(); }
        nodes.clear(); splits.clear();
        nodes.reserve(prims.size() * 2);
        root = storage.empty() ? -1 : build_range(0, (int)storage.size());
        mark_wide();
    }
    // Work counts only (not a reference algorithm): the binary nodes that head
    // a 4-wide node of the device layout.  The root heads one; a head's cut
    // starts as its two children and grows by splitting the internal member
    // with the largest box surface area (the first such on ties) until it has
    // four members or only leaves; every internal member of the cut heads a
    // node of its own.  A 4-wide node is entered exactly when the recursion
    // enters its head, so counting entered heads gives the device's count.
    std::vector<char> wide_head;
    void mark_wide() {
        wide_head.assign(nodes.size(), 0);
        if (root < 0 || nodes[root].leaf >= 0) return;
        std::vector<int> todo{root};
        while (!todo.empty()) {
            const int h = todo.back();
            todo.pop_back();
            wide_head[h] = 1;
            std::vector<int> cut{nodes[h].left, nodes[h].right};
            while (cut.size() < 4) {
                int j = -1;
                float best = -1.0f;
                for (int k = 0; k < (int)cut.size(); ++k) {
                    const Node& m = nodes[cut[k]];
                    if (m.leaf < 0 && surface_area(m.box) > best) { best = surface_area(m.box); j = k; }
                }
                if (j < 0) break;
                const int a = nodes[cut[j]].left, b = nodes[cut[j]].right;
                cut[j] = a;
                cut.insert(cut.begin() + j + 1, b);
            }
            for (int m : cut) if (nodes[m].leaf < 0) todo.push_back(m);
        }
    }
    // bvh.hpp:74-104
    int build_range(int start, int end) {
        int idx = (int)nodes.size();
        nodes.push_back(Node{});
        AABB box;
        for (int i = start; i < end; ++i) box = AABB::merge(box, pbox[storage[i]]);
        nodes[idx].box = box;
        int count = end - start;
        if (count == 1) { nodes[idx].leaf = storage[start]; return idx; }
        int axis = box.longest_axis();
        if (count <= 4) median_split(start, end, axis);
        else sah_split(start, end, box, axis);
        int split = start + count / 2;
        if (count > 4) split = last_split;
        splits.push_back(split);
        int l = build_range(start, split);
        int r = build_range(split, end);
        nodes[idx].left = l; nodes[idx].right = r;
        return idx;
    }
    // bvh.hpp:106-115
    void median_split(int start, int end, int axis) {
        const std::vector<AABB>& pb = pbox;
        std::sort(storage.begin() + start, storage.begin() + end, [axis, &pb](int x, int y) {
            const AABB& ba = pb[x]; const AABB& bb = pb[y];
            return (ba.mn[axis] + ba.mx[axis]) < (bb.mn[axis] + bb.mx[axis]);
        });
        last_split = start + (end - start) / 2;
    }
    static float surface_area(const AABB& b) {
        V3 e = b.mx - b.mn;
        return 2.0f * (e.x * e.y + e.y * e.z + e.x * e.z);
    }
    // bvh.hpp:119-192
    void sah_split(int start, int end, const AABB& parent, int axis) {
        constexpr int NB = 12;
        struct Bucket { int count = 0; AABB bounds; };
        Bucket bk[NB];
        float lo = parent.mn[axis], hi = parent.mx[axis];
        if (hi <= lo) { median_split(start, end, axis); return; }
        for (int i = start; i < end; ++i) {
            const AABB& pb = pbox[storage[i]];
            float center = (pb.mn[axis] + pb.mx[axis]) * 0.5f;
            int b = (int)((center - lo) / (hi - lo) * NB);
            if (b < 0) b = 0;
            if (b >= NB) b = NB - 1;
            bk[b].count++;
            bk[b].bounds = AABB::merge(bk[b].bounds, pb);
        }
        AABB pre[NB + 1]; int prec[NB + 1] = {0};
        for (int i = 0; i < NB; ++i) { pre[i + 1] = AABB::merge(pre[i], bk[i].bounds); prec[i + 1] = prec[i] + bk[i].count; }
        AABB suf[NB + 1]; int sufc[NB + 1] = {0};
        for (int i = NB - 1; i >= 0; --i) { suf[i] = AABB::merge(suf[i + 1], bk[i].bounds); sufc[i] = sufc[i + 1] + bk[i].count; }
        float best_cost = FLT_MAX; int best = -1;
        float sa_p = surface_area(parent);
        for (int i = 1; i < NB; ++i) {
            int nL = prec[i], nR = sufc[i];
            if (nL == 0 || nR == 0) continue;
            float saL = surface_area(pre[i]), saR = surface_area(suf[i]);
            float cost = 1.0f + (nL * saL + nR * saR) / (sa_p > 0 ? sa_p : 1.0f);
            if (cost < best_cost) { best_cost = cost; best = i; }
        }
        if (best < 0) { median_split(start, end, axis); return; }
        const std::vector<AABB>& pbv = pbox;
        auto pivot = std::partition(storage.begin() + start, storage.begin() + end, [&](int p) {
            const AABB& pb = pbv[p];
            float center = (pb.mn[axis] + pb.mx[axis]) * 0.5f;
            int b = (int)((center - lo) / (hi - lo) * NB);
            if (b < 0) b = 0;
            if (b >= NB) b = NB - 1;
            return b < best;
        });
        last_split = (int)(pivot - storage.begin());
        if (last_split <= start || last_split >= end) median_split(start, end, axis);
    }
    // BVHNode::hit, bvh.hpp:31-37
    bool hit(const std::vector<Prim>& P, int ni, const Ray& r, float t_min, float t_max, Hit& out, Counters* c) const {
        const Node& n = nodes[ni];
        if (!n.box.hit(r, t_min, t_max)) return false;
        if (n.leaf >= 0) { if (c) c->tests[P[n.leaf].kind]++; return P[n.leaf].hit(r, t_min, t_max, out); }
        if (c && wide_head[ni]) c->nodes++;
        bool hl = n.left >= 0 && hit(P, n.left, r, t_min, t_max, out, c);
        bool hr = n.right >= 0 && hit(P, n.right, r, t_min, hl ? out.t : t_max, out, c);
        return hl || hr;
    }
    // BVHNode::occluded, bvh.hpp:39-48
    bool occluded(const std::vector<Prim>& P, int ni, const Ray& r, float t_min, float t_max, Counters* c) const {
        const Node& n = nodes[ni];
        if (!n.box.hit(r, t_min, t_max)) return false;
        if (n.leaf >= 0) { if (c) c->tests[P[n.leaf].kind]++; Hit tmp; return P[n.leaf].hit(r, t_min, t_max, tmp); }
        if (c && wide_head[ni]) c->nodes++;
        if (n.left >= 0 && occluded(P, n.left, r, t_min, t_max, c)) return true;
        if (n.right >= 0 && occluded(P, n.right, r, t_min, t_max, c)) return true;
        return false;
    }
};

struct Mat { int kind; V3 albedo; float refl, ior, spec, shin; V3 spec_color, emission; };
struct Light { int kind; V3 pos; V3 I; float radius; bool emissive; };

// ---- scene.hpp:21-105 -----------------------------------------------------
struct Scene {
    std::vector<Prim> prims;
    std::vector<Mat> mats;
    std::vector<Light> lights;
    V3 bg, amb;
    bool use_bvh = true;
    BVH bvh;
    // Scene::hit, scene.hpp:55-70 ("last_id" reports which prim produced out).
    bool hit(const Ray& r, float t_min, float t_max, Hit& out, int* last_id, Counters* c) const {
        if (use_bvh && !prims.empty()) {
            Hit h; int id = -1;
            bool ok = hit_id(bvh.root, r, t_min, t_max, h, id, c);
            if (ok) { out = h; if (last_id) *last_id = id; }
            return ok;
        }
        bool found = false; float closest = t_max;
        for (const auto& p : prims) {
            Hit tmp;
            if (c) c->tests[p.kind]++;
            if (p.hit(r, t_min, closest, tmp)) { closest = tmp.t; out = tmp; found = true; if (last_id) *last_id = p.id; }
        }
        return found;
    }
    // bvh recursion carrying the id of the last successful leaf.
    bool hit_id(int ni, const Ray& r, float t_min, float t_max, Hit& out, int& id, Counters* c) const {
        const BVH::Node& n = bvh.nodes[ni];
        if (!n.box.hit(r, t_min, t_max)) return false;
        if (n.leaf >= 0) {
            if (c) c->tests[prims[n.leaf].kind]++;
            if (prims[n.leaf].hit(r, t_min, t_max, out)) { id = n.leaf; return true; }
            return false;
        }
        if (c && bvh.wide_head[ni]) c->nodes++;
        bool hl = n.left >= 0 && hit_id(n.left, r, t_min, t_max, out, id, c);
        bool hr = n.right >= 0 && hit_id(n.right, r, t_min, hl ? out.t : t_max, out, id, c);
        return hl || hr;
    }
    // Scene::occluded, scene.hpp:72-81
    bool occluded(const Ray& r, float t_min, float t_max, Counters* c) const {
        if (use_bvh && !prims.empty()) return bvh.occluded(prims, bvh.root, r, t_min, t_max, c);
        for (const auto& p : prims) { Hit tmp; if (c) c->tests[p.kind]++; if (p.hit(r, t_min, t_max, tmp)) return true; }
        return false;
    }
};

// ---- camera.hpp:29-86, camera.cpp:14-68 -----------------------------------
struct Camera {
    V3 origin, lower_left, horizontal, vertical, u, v, w;
    float aperture = 0, focus_dist = 1;
    int blades = 0;
    void init(const pcl_camera& c) {
        V3 eye{c.eye[0], c.eye[1], c.eye[2]}, look{c.look[0], c.look[1], c.look[2]}, up{c.up[0], c.up[1], c.up[2]};
        origin = eye; aperture = c.aperture; blades = c.bokeh_blades;
        float fd = c.focus_dist;
        if (fd <= 0.0f) fd = (eye - look).length();   // ParsedScene::finalize, pecl_parser.hpp:136-140
        focus_dist = fd;
        w = (eye - look).normalized();
        u = up.cross(w).normalized();
        v = w.cross(u);
        float theta = c.vfov_deg * 3.14159265358979323846f / 180.0f;
        float half_h = std::tan(theta * 0.5f);
        float half_w = c.aspect * half_h;
        horizontal = u * (2.0f * half_w * focus_dist);
        vertical = v * (2.0f * half_h * focus_dist);
        lower_left = origin - horizontal * 0.5f - vertical * 0.5f - w * focus_dist;
    }
    static void sample_aperture(Rng& rng, int blades, float& dx, float& dy) {
        if (blades <= 2) { rng.unit_disc(dx, dy); return; }
        for (int attempt = 0; attempt < 16; ++attempt) {
            rng.unit_disc(dx, dy);
            float angle_step = 2.0f * 3.14159265358979f / blades;
            float r = std::sqrt(dx * dx + dy * dy);
            float theta = o_atan2(dy, dx);
            if (theta < 0) theta += 2.0f * 3.14159265358979f;
            int sector = (int)(theta / angle_step);
            float local_angle = theta - sector * angle_step;
            float half_step = angle_step * 0.5f;
            float max_r = o_cos(half_step) / o_cos(local_angle - half_step);
            if (r <= max_r) return;
        }
        rng.unit_disc(dx, dy);
    }
    Ray ray_at(float uu, float vv, Rng* rng) const {
        V3 dir = lower_left + horizontal * uu + vertical * vv - origin;
        V3 o = origin;
        if (rng && aperture > 0.0f) {
            float fd = focus_dist;
            const V3 forward = (lower_left + horizontal * 0.5f + vertical * 0.5f - origin).normalized();
            const float denom = dir.normalized().dot(forward);
            const float focus_t = (std::fabs(denom) > 1e-6f) ? fd / denom : fd;
            V3 focal = origin + dir.normalized() * focus_t;
            float dx, dy;
            sample_aperture(*rng, blades, dx, dy);
            V3 off = u * dx + v * dy;
            off = off * aperture;
            o = origin + off;
            dir = focal - o;
        }
        return Ray{o, dir.normalized()};
    }
};

// ---- renderer.hpp --------------------------------------------------------
struct Renderer {
    const Scene& S;
    Camera cam;
    pcl_settings st;
    mutable std::atomic<uint64_t> rays{0};
    mutable std::atomic<uint64_t> closest_q{0}, any_q{0};
    // work counters of the bytes model (node visits, tests per kind), when asked
    bool count = false;
    mutable std::atomic<uint64_t> nodes{0};
    mutable std::atomic<uint64_t> tests[6] = {{0}, {0}, {0}, {0}, {0}, {0}};
    Renderer(const Scene& s, const Camera& c, const pcl_settings& se) : S(s), cam(c), st(se) {}

    void add(const Counters& c) const {
        nodes += c.nodes;
        for (int k = 0; k < 6; ++k) tests[k] += c.tests[k];
    }
    bool occl(const Ray& r, float a, float b) const {
        any_q++;
        if (!count) return S.occluded(r, a, b, nullptr);
        Counters c;
        const bool o = S.occluded(r, a, b, &c);
        add(c);
        return o;
    }
    bool hit(const Ray& r, float a, float b, Hit& h) const {
        closest_q++;
        if (!count) return S.hit(r, a, b, h, nullptr, nullptr);
        Counters c;
        const bool x = S.hit(r, a, b, h, nullptr, &c);
        add(c);
        return x;
    }

    // Renderer::trace, renderer.hpp:403-822 (branches without textures, fog,
    // SSS, caustics, dispersion and principled: rejected at scene creation).
    V3 trace(const Ray& r, int depth, Rng& rng) const {
        rays++;
        if (depth <= 0) return {0, 0, 0};
        Hit rec;
        if (!hit(r, 1e-4f, 1e9f, rec)) return S.bg;
        const Mat& m = S.mats[rec.mat];
        V3 albedo = m.albedo;
        V3 shade_normal = rec.normal;
        V3 local = m.emission;
        for (const auto& L : S.lights) {
            V3 contrib{0, 0, 0};
            int nsamp = (L.radius > 0.0f) ? st.soft_shadow_samples : 1;
            for (int s = 0; s < nsamp; ++s) {
                V3 L_pos = L.pos;
                if (L.kind == PCL_LIGHT_POINT && L.radius > 0.0f) {
                    V3 to_light = L.pos - rec.p;
                    V3 axis = to_light.normalized();
                    V3 up = (std::fabs(axis.y) < 0.9f) ? V3{0, 1, 0} : V3{1, 0, 0};
                    V3 t1 = axis.cross(up).normalized();
                    V3 t2 = axis.cross(t1).normalized();
                    float dx, dy;
                    rng.unit_disc(dx, dy);
                    L_pos = L.pos + (t1 * dx + t2 * dy) * L.radius;
                }
                V3 Ld; float dist = 1.0f;
                if (L.kind == PCL_LIGHT_POINT) { Ld = L_pos - rec.p; dist = Ld.length(); Ld = Ld / dist; }
                else { Ld = -L.pos; dist = 1e9f; }
                Ray sr{rec.p + rec.normal * st.shadow_bias, Ld};
                if (!L.emissive && occl(sr, st.shadow_bias, dist - st.shadow_bias)) continue;
                float ndotl = clampf(shade_normal.dot(Ld), 0.0f, 1.0f);
                V3 c = albedo * L.I * ndotl;
                if (m.spec > 0.0f && ndotl > 0.0f) {
                    V3 Vv = -r.d.normalized();
                    V3 H = (Ld + Vv).normalized();
                    float ndoth = clampf(shade_normal.dot(H), 0.0f, 1.0f);
                    float sp = o_pow(ndoth, m.shin);
                    c += m.spec_color * L.I * (sp * m.spec);
                }
                contrib += c;
            }
            local += contrib / (float)nsamp;
        }
        local += albedo * S.amb;
        if (m.kind == PCL_MAT_METAL && m.refl > 0.0f) {
            V3 rd = r.d.normalized().reflect(rec.normal).normalized();
            Ray rr{rec.p + rec.normal * st.shadow_bias, rd};
            V3 inc = trace(rr, depth - 1, rng);
            V3 rc = inc * albedo * m.refl;
            local = local * (1.0f - m.refl) + rc;
        } else if (m.kind == PCL_MAT_DIELECTRIC) {
            local = trace_dielectric(r, rec, m, albedo, depth, rng);
        }
        return local;
    }

    // Renderer::trace_dielectric, renderer.hpp:1582-1627
    V3 trace_dielectric(const Ray& r, const Hit& rec, const Mat& m, const V3& albedo, int depth, Rng& rng) const {
        float ior = m.ior;
        bool entering = rec.front;
        float safe = std::fmax(ior, 0.01f);
        float eta = entering ? (1.0f / safe) : safe;
        V3 nf = rec.normal;
        float cos_i = std::fabs(r.d.normalized().dot(rec.normal));
        float sin_t2 = eta * eta * (1.0f - cos_i * cos_i);
        V3 rd = r.d.normalized().reflect(rec.normal).normalized();
        Ray rr{rec.p + rec.normal * st.shadow_bias, rd};
        V3 refl = trace(rr, depth - 1, rng);
        if (sin_t2 > 1.0f) return refl * albedo;
        float cos_t = std::sqrt(1.0f - sin_t2);
        V3 td = (r.d.normalized() * eta + nf * (eta * cos_i - cos_t)).normalized();
        Ray tr{rec.p - nf * st.shadow_bias, td};
        V3 refr = trace(tr, depth - 1, rng);
        float cs = entering ? cos_i : cos_t;
        float kr = schlick(cs, eta);
        float kt = 1.0f - kr;
        V3 g = refl * kr + refr * kt;
        return g * albedo;
    }

    // Renderer::trace_path, renderer.hpp:829-972; with mis, trace_path_mis
    // (renderer.hpp:1205-1335), which differs only in the light loop
    V3 trace_path(const Ray& r0, int bounces, Rng& rng, bool mis = false) const {
        rays++;
        V3 T{1, 1, 1}, R{0, 0, 0};
        Ray cur = r0;
        int depth = bounces;
        while (depth > 0) {
            Hit rec;
            if (!hit(cur, 1e-4f, 1e9f, rec)) { R += T * S.bg; break; }
            const Mat& m = S.mats[rec.mat];
            R += T * m.emission;
            V3 albedo = m.albedo;
            V3 sn = rec.normal;
            for (const auto& L : S.lights) {
                V3 Ld; float dist = 1.0f;
                V3 L_pos = L.pos;
                if (L.kind == PCL_LIGHT_POINT) {
                    if (L.radius > 0.0f) {
                        V3 to_light = L.pos - rec.p;
                        V3 axis = to_light.normalized();
                        V3 up = (std::fabs(axis.y) < 0.9f) ? V3{0, 1, 0} : V3{1, 0, 0};
                        V3 t1 = axis.cross(up).normalized();
                        V3 t2 = axis.cross(t1).normalized();
                        float dx, dy;
                        rng.unit_disc(dx, dy);
                        L_pos = L.pos + (t1 * dx + t2 * dy) * L.radius;
                    }
                    Ld = L_pos - rec.p; dist = Ld.length(); Ld = Ld / dist;
                } else { Ld = -L.pos; dist = 1e9f; }
                Ray sr{rec.p + rec.normal * st.shadow_bias, Ld};
                if (!L.emissive && occl(sr, st.shadow_bias, dist - st.shadow_bias)) continue;
                float ndotl = clampf(sn.dot(Ld), 0.0f, 1.0f);
                if (!mis) {
                    R += T * albedo * L.I * ndotl;
                    continue;
                }
                // renderer.hpp:1264-1302: power-heuristic weights of the light
                // sample against a cosine-hemisphere BRDF sample
                if (ndotl <= 0.0f) continue;
                float pdf_brdf = ndotl / 3.14159265f;
                float pdf_light;
                if (L.kind == PCL_LIGHT_POINT && L.radius > 0.0f) {
                    float rr = L.radius;
                    float cos_at_light = (-Ld).dot(V3{0, 1, 0});
                    if (cos_at_light < 0.01f) cos_at_light = 0.01f;
                    pdf_light = (dist * dist) / (3.14159265f * rr * rr * cos_at_light);
                } else {
                    pdf_light = 1.0f;
                }
                float w_light = (pdf_light * pdf_light) / (pdf_light * pdf_light + pdf_brdf * pdf_brdf + 1e-9f);
                float w_brdf = 1.0f - w_light;
                V3 contrib_nee = albedo * L.I * ndotl;
                V3 contrib_brdf = albedo * L.I;
                R += T * (contrib_nee * w_light + contrib_brdf * w_brdf);
            }
            R += T * albedo * S.amb;
            if (bounces - depth > 3) {
                float p = std::fmax(T.x, std::fmax(T.y, T.z));
                if (rng.unit() > p) break;
                T = T / p;
            }
            V3 nd;
            V3 NT = T * albedo;
            if (m.kind == PCL_MAT_METAL && m.refl > 0.0f) {
                V3 refl = cur.d.normalized().reflect(sn).normalized();
                V3 up = (std::fabs(refl.y) < 0.9f) ? V3{0, 1, 0} : V3{1, 0, 0};
                V3 t1 = refl.cross(up).normalized();
                V3 t2 = refl.cross(t1).normalized();
                float r1 = rng.unit() * 0.1f;
                float r2 = rng.unit() * 6.28f;
                nd = (refl + t1 * (r1 * o_cos(r2)) + t2 * (r1 * o_sin(r2))).normalized();
                NT = T * albedo * m.refl;
            } else if (m.kind == PCL_MAT_DIELECTRIC) {
                V3 g = trace_dielectric(cur, rec, m, albedo, depth, rng);
                R += T * g - T * m.emission;
                break;
            } else {
                float r1 = rng.unit() * 6.2831853f;
                float r2 = rng.unit();
                float r2s = std::sqrt(r2);
                V3 N = sn;
                V3 up = (std::fabs(N.y) < 0.9f) ? V3{0, 1, 0} : V3{1, 0, 0};
                V3 Tt = up.cross(N).normalized();
                V3 B = N.cross(Tt).normalized();
                nd = (Tt * o_cos(r1) * r2s + B * o_sin(r1) * r2s + N * std::sqrt(1.0f - r2)).normalized();
            }
            cur = Ray{rec.p + rec.normal * st.shadow_bias, nd};
            T = NT;
            depth--;
        }
        return R;
    }

    // render_tile_region per pixel, renderer.hpp:309-390 (non-adaptive branch)
    V3 pixel(int i, int j) const {
        const int W = st.width, H = st.height;
        Rng rng((uint32_t)(i * 97861u + j * 1973u + 1u));
        V3 acc{0, 0, 0};
        int n = st.samples;
        for (int sy = 0; sy < n; ++sy) {
            for (int sx = 0; sx < n; ++sx) {
                float uu = (i + (sx + 0.5f) / n) / (float)W;
                float vv = 1.0f - (j + (sy + 0.5f) / n) / (float)H;
                uu += rng.sym() * (0.5f / (W * n));
                vv += rng.sym() * (0.5f / (H * n));
                Ray r = cam.ray_at(uu, vv, &rng);
                if (st.pathtrace) acc += trace_path(r, st.pathtrace_bounces, rng, st.pathtrace == 2);
                else acc += trace(r, st.max_depth, rng);
            }
        }
        return acc / (float)(n * n);
    }
};

struct SceneHandle {
    Scene scene;
};

static int build_scene(const pcl_scene_desc* d, Scene& S) {
    S.prims.clear(); S.mats.clear(); S.lights.clear();
    for (int i = 0; i < d->n_prims; ++i) S.prims.push_back(Prim::make(d->prims[i], i));
    for (int i = 0; i < d->n_materials; ++i) {
        const pcl_material& m = d->materials[i];
        if (m.kind == PCL_MAT_PRINCIPLED) return -1;
        S.mats.push_back(Mat{m.kind, {m.albedo[0], m.albedo[1], m.albedo[2]}, m.reflectivity, m.ior, m.specular,
                             m.shininess, {m.specular_color[0], m.specular_color[1], m.specular_color[2]},
                             {m.emission[0], m.emission[1], m.emission[2]}});
    }
    for (int i = 0; i < d->n_lights; ++i) {
        const pcl_light& l = d->lights[i];
        S.lights.push_back(Light{l.kind, {l.pos_or_dir[0], l.pos_or_dir[1], l.pos_or_dir[2]},
                                 {l.intensity[0], l.intensity[1], l.intensity[2]}, l.radius, l.is_emissive != 0});
    }
    S.bg = {d->background[0], d->background[1], d->background[2]};
    S.amb = {d->ambient[0], d->ambient[1], d->ambient[2]};
    S.use_bvh = d->use_bvh != 0;
    S.bvh.build(S.prims);
    return 0;
}

template <class F>
static void parallel_for(int n, int threads, F&& f) {
    if (threads <= 1 || n < 2) { for (int i = 0; i < n; ++i) f(i); return; }
    std::atomic<int> next{0};
    std::vector<std::thread> ws;
    const int chunk = 64;
    for (int t = 0; t < threads; ++t) ws.emplace_back([&]() {
        while (true) {
            int b = next.fetch_add(chunk);
            if (b >= n) break;
            int e = std::min(n, b + chunk);
            for (int i = b; i < e; ++i) f(i);
        }
    });
    for (auto& w : ws) w.join();
}

}  // namespace orc

extern "C" {

typedef struct orc_counts {
    uint64_t rays_traced, closest_queries, anyhit_queries;
    uint64_t node_visits;
    uint64_t prim_tests[6];
} orc_counts;

// 0: libm float transcendentals (the reference's); 1: double-evaluated,
// rounded once (the HIP core's).  Returns the previous mode.  Process-wide;
// set it before rendering, not while another thread renders.
int orc_set_trig_mode(int mode) {
    const int prev = orc::g_trig_mode;
    orc::g_trig_mode = mode ? 1 : 0;
    return prev;
}

void* orc_scene_create(const pcl_scene_desc* d) {
    auto* h = new orc::SceneHandle();
    if (orc::build_scene(d, h->scene) != 0) { delete h; return nullptr; }
    return h;
}
void orc_scene_destroy(void* h) { delete (orc::SceneHandle*)h; }

// out[slot] = insertion id of the prim stored at BVH leaf slot (BVH::storage_).
int orc_leaf_order(void* h, int32_t* out, int32_t n) {
    auto& S = ((orc::SceneHandle*)h)->scene;
    if (n < (int)S.bvh.storage.size()) return -1;
    for (size_t i = 0; i < S.bvh.storage.size(); ++i) out[i] = S.bvh.storage[i];
    return (int)S.bvh.storage.size();
}
// Preorder split points of build_range.
int orc_split_points(void* h, int32_t* out, int32_t n) {
    auto& S = ((orc::SceneHandle*)h)->scene;
    if (n < (int)S.bvh.splits.size()) return -1;
    for (size_t i = 0; i < S.bvh.splits.size(); ++i) out[i] = S.bvh.splits[i];
    return (int)S.bvh.splits.size();
}
// Number of binary nodes that head a 4-wide node (BVH::mark_wide).
int orc_wide_heads(void* h) {
    auto& S = ((orc::SceneHandle*)h)->scene;
    int n = 0;
    for (char c : S.bvh.wide_head) n += c ? 1 : 0;
    return n;
}

// Scene::hit / Scene::occluded over a ray batch; counts (optional) accumulate
// the work per query for the bytes model.
int orc_trace_rays(void* h, const pcl_ray* rays, int32_t n, int32_t any_hit, pcl_hit* hits,
                   orc_counts* counts, int32_t threads) {
    auto& S = ((orc::SceneHandle*)h)->scene;
    std::vector<orc::Counters> per(n);
    orc::parallel_for(n, threads, [&](int i) {
        const pcl_ray& q = rays[i];
        orc::Ray r{{q.origin[0], q.origin[1], q.origin[2]}, {q.dir[0], q.dir[1], q.dir[2]}};
        orc::Counters* c = counts ? &per[i] : nullptr;
        if (any_hit) {
            bool o = S.occluded(r, q.tmin, q.tmax, c);
            hits[i].t = o ? 1.0f : -1.0f;
            hits[i].prim_id = o ? 1 : -1;
        } else {
            orc::Hit rec; int id = -1;
            if (S.hit(r, q.tmin, q.tmax, rec, &id, c)) { hits[i].t = rec.t; hits[i].prim_id = id; }
            else { hits[i].t = -1.0f; hits[i].prim_id = -1; }
        }
    });
    if (counts) {
        for (int i = 0; i < n; ++i) {
            counts->node_visits += per[i].nodes;
            for (int k = 0; k < 6; ++k) counts->prim_tests[k] += per[i].tests[k];
        }
        if (any_hit) counts->anyhit_queries += n; else counts->closest_queries += n;
    }
    return 0;
}

// Full hit record for closest hits: p, normal, t, material, front_face.
int orc_hit_records(void* h, const pcl_ray* rays, int32_t n, float* out /* n*9 */, int32_t threads) {
    auto& S = ((orc::SceneHandle*)h)->scene;
    orc::parallel_for(n, threads, [&](int i) {
        const pcl_ray& q = rays[i];
        orc::Ray r{{q.origin[0], q.origin[1], q.origin[2]}, {q.dir[0], q.dir[1], q.dir[2]}};
        orc::Hit rec;
        float* o = out + (size_t)i * 9;
        if (S.hit(r, q.tmin, q.tmax, rec, nullptr, nullptr)) {
            o[0] = rec.p.x; o[1] = rec.p.y; o[2] = rec.p.z;
            o[3] = rec.normal.x; o[4] = rec.normal.y; o[5] = rec.normal.z;
            o[6] = rec.t; o[7] = (float)rec.mat; o[8] = rec.front ? 1.0f : 0.0f;
        } else {
            for (int k = 0; k < 9; ++k) o[k] = 0.0f;
            o[6] = -1.0f;
        }
    });
    return 0;
}

// Camera::ray_at for pixel-center rays (no jitter, no aperture): the rays
// tests feed to the traversal entry points.
int orc_camera_rays(const pcl_camera* c, int32_t W, int32_t H, pcl_ray* out) {
    orc::Camera cam; cam.init(*c);
    for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            float u = (i + 0.5f) / (float)W;
            float v = 1.0f - (j + 0.5f) / (float)H;
            orc::Ray r = cam.ray_at(u, v, nullptr);
            pcl_ray& q = out[(size_t)j * W + i];
            q.origin[0] = r.o.x; q.origin[1] = r.o.y; q.origin[2] = r.o.z;
            q.dir[0] = r.d.x; q.dir[1] = r.d.y; q.dir[2] = r.d.z;
            q.tmin = 1e-4f; q.tmax = 1e9f;
        }
    return 0;
}

// Renderer::render restated; rows [row0, row1) only (row1 <= 0 means H), so a
// bounded sample of a large frame can be timed.
int orc_render(void* h, const pcl_camera* c, const pcl_settings* s, float* out, orc_counts* counts,
               int32_t threads, int32_t row0, int32_t row1) {
    auto& S = ((orc::SceneHandle*)h)->scene;
    orc::Camera cam; cam.init(*c);
    orc::Renderer R(S, cam, *s);
    R.count = counts != nullptr;
    const int W = s->width;
    int r1 = (row1 <= 0) ? s->height : row1;
    int npx = (r1 - row0) * W;
    orc::parallel_for(npx, threads, [&](int k) {
        int j = row0 + k / W, i = k % W;
        orc::V3 px = R.pixel(i, j);
        float* o = out + ((size_t)j * W + i) * 3;
        o[0] = px.x; o[1] = px.y; o[2] = px.z;
    });
    if (counts) {
        counts->rays_traced += R.rays.load();
        counts->closest_queries += R.closest_q.load();
        counts->anyhit_queries += R.any_q.load();
        counts->node_visits += R.nodes.load();
        for (int k = 0; k < 6; ++k) counts->prim_tests[k] += R.tests[k].load();
    }
    return 0;
}

// tonemap_color(ACES) + color_to_rgb8, tonemap.hpp:27-61 / image_io.hpp:46-59
int orc_tonemap_rgb8(const float* hdr, int64_t npx, float exposure, uint8_t* out) {
    float e = std::pow(2.0f, exposure);
    auto aces = [](float x) {
        float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, ee = 0.14f;
        return std::max(0.0f, (x * (a * x + b)) / (x * (cc * x + d) + ee));
    };
    for (int64_t i = 0; i < npx; ++i) {
        for (int k = 0; k < 3; ++k) {
            float x = hdr[i * 3 + k] * e;
            float m = orc::clampf(aces(x), 0.0f, 1.0f);
            m = std::pow(m, 1.0f / 2.2f);
            out[i * 3 + k] = (uint8_t)(m * 255.0f + 0.5f);
        }
    }
    return 0;
}

// Checks a gamma threshold table against color_to_rgb8's encode step
// (image_io.hpp:53-55) on every float in [0, 1]: returns the number of v with
// (uint8)(powf(v, 1/2.2f) * 255 + 0.5) != #{k >= 1 : t[k] <= v}.
int64_t orc_check_thresholds(const float* t, int threads) {
    const uint32_t end = 0x3f800000u + 1;   // bit patterns 0 .. 1.0f inclusive
    std::atomic<int64_t> bad{0};
    std::vector<std::thread> pool;
    if (threads < 1) threads = 1;
    for (int w = 0; w < threads; ++w) {
        pool.emplace_back([&, w]() {
            int64_t local = 0;
            for (uint32_t u = (uint32_t)w; u < end; u += (uint32_t)threads) {
                float v;
                std::memcpy(&v, &u, 4);
                const int e = (int)(uint8_t)(std::pow(v, 1.0f / 2.2f) * 255.0f + 0.5f);
                int lo = 0;
                for (int step = 128; step >= 1; step >>= 1)
                    if (t[lo + step] <= v) lo += step;
                if (lo != e) ++local;
            }
            bad += local;
        });
    }
    for (auto& th : pool) th.join();
    return bad.load();
}

uint32_t orc_rng_u32(uint32_t seed, int32_t k) {
    orc::Rng r(seed);
    uint32_t v = 0;
    for (int i = 0; i <= k; ++i) v = r.u32();
    return v;
}

}  // extern "C"
