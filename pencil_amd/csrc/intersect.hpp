// intersect.hpp - ray/primitive tests and BVH traversal on the device.
//
// Each primitive test restates the reference's Primitive::hit (cited per
// function) with the same float operations in the same order, so t, the hit
// point and the normal are the reference's bit for bit.  One template serves
// both the traversal (REC = false: only t) and the final hit record
// (REC = true), so the two can never disagree.
#pragma once

#include "device_math.hpp"
#include "scene_layout.hpp"
#include "pencil_hip.h"

namespace pcl {

struct HitRec {
    f3 p, normal;
    float t;
    int mat;
    bool front;
};

PCL_D void set_face_normal(HitRec* h, const Ray& r, const f3& n) {
    h->front = r.d.dot(n) < 0.0f;
    h->normal = h->front ? n : -n;
}

// Record loaded from one 64-byte PrimRec line.
struct PrimView {
    float4 a, b, c;
    int kind, mat;
    PCL_D f3 v(int i) const {   // three consecutive floats starting at g[i]
        const float g[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
        return f3(g[i], g[i + 1], g[i + 2]);
    }
};

PCL_D PrimView load_prim(const PrimRec* P, int slot) {
    const float4* q = reinterpret_cast<const float4*>(P + slot);
    PrimView v;
    v.a = gload_f4(q); v.b = gload_f4(q + 1); v.c = gload_f4(q + 2);
    int4 d = gload_i4(q + 3);
    v.kind = d.x; v.mat = d.y;
    return v;
}

// Sphere::hit, primitives.hpp:98-120
template <bool REC>
PCL_D bool hit_sphere(const PrimView& P, const Ray& r, float t_min, float t_max, float& t_out, HitRec* rec) {
    const f3 c(P.a.x, P.a.y, P.a.z);
    const float rad = P.a.w;
    f3 oc = r.o - c;
    float a = r.d.dot(r.d);
    float half_b = oc.dot(r.d);
    float cc = oc.dot(oc) - rad * rad;
    float disc = half_b * half_b - a * cc;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float t = (-half_b - sq) / a;
    if (t < t_min || t > t_max) {
        t = (-half_b + sq) / a;
        if (t < t_min || t > t_max) return false;
    }
    t_out = t;
    if (REC) {
        rec->t = t; rec->p = r.at(t);
        f3 on = (rec->p - c) / rad;
        set_face_normal(rec, r, on); rec->mat = P.mat;
    }
    return true;
}

// Triangle::hit (Moller-Trumbore), primitives.hpp:145-170
template <bool REC>
PCL_D bool hit_triangle(const PrimView& P, const Ray& r, float t_min, float t_max, float& t_out, HitRec* rec) {
    const f3 v0(P.a.x, P.a.y, P.a.z), e1(P.a.w, P.b.x, P.b.y), e2(P.b.z, P.b.w, P.c.x);
    f3 h = r.d.cross(e2);
    float a = e1.dot(h);
    if (fabsf(a) < 1e-8f) return false;
    float f = 1.0f / a;
    f3 s = r.o - v0;
    float u = f * s.dot(h);
    if (u < 0.0f || u > 1.0f) return false;
    f3 q = s.cross(e1);
    float v = f * r.d.dot(q);
    if (v < 0.0f || u + v > 1.0f) return false;
    float t = f * e2.dot(q);
    if (t < t_min || t > t_max) return false;
    t_out = t;
    if (REC) {
        rec->t = t; rec->p = r.at(t);
        set_face_normal(rec, r, f3(P.c.y, P.c.z, P.c.w)); rec->mat = P.mat;
    }
    return true;
}

// Cube::hit (slab with near/far axis), primitives.hpp:201-238
template <bool REC>
PCL_D bool hit_cube(const PrimView& P, const Ray& r, float t_min, float t_max, float& t_out, HitRec* rec) {
    const float lo[3] = {P.a.x, P.a.y, P.a.z}, hi[3] = {P.a.w, P.b.x, P.b.y};
    const float o[3] = {r.o.x, r.o.y, r.o.z}, dd[3] = {r.d.x, r.d.y, r.d.z};
    float t_near = -__builtin_inff(), t_far = __builtin_inff();
    int near_axis = -1, far_axis = -1;
    float near_sign = 1.0f, far_sign = -1.0f;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float d = dd[ax];
        if (fabsf(d) < 1e-12f) {
            if (o[ax] < lo[ax] || o[ax] > hi[ax]) return false;
            continue;
        }
        float tn = (lo[ax] - o[ax]) / d, tf = (hi[ax] - o[ax]) / d;
        float ns = -1.0f, fs = 1.0f;
        if (tn > tf) { float x = tn; tn = tf; tf = x; x = ns; ns = fs; fs = x; }
        if (tn > t_near) { t_near = tn; near_axis = ax; near_sign = ns; }
        if (tf < t_far) { t_far = tf; far_axis = ax; far_sign = fs; }
        if (t_far < t_near) return false;
    }
    const bool use_near = t_near >= t_min;
    const float t_hit = use_near ? t_near : t_far;
    const int hit_axis = use_near ? near_axis : far_axis;
    const float hit_sign = use_near ? near_sign : far_sign;
    if (hit_axis < 0 || t_hit < t_min || t_hit > t_max) return false;
    t_out = t_hit;
    if (REC) {
        rec->t = t_hit; rec->p = r.at(t_hit);
        f3 n(hit_axis == 0 ? hit_sign : 0.0f, hit_axis == 1 ? hit_sign : 0.0f, hit_axis == 2 ? hit_sign : 0.0f);
        set_face_normal(rec, r, n); rec->mat = P.mat;
    }
    return true;
}

// Plane::hit, primitives.hpp:261-272
template <bool REC>
PCL_D bool hit_plane(const PrimView& P, const Ray& r, float t_min, float t_max, float& t_out, HitRec* rec) {
    const f3 pt(P.a.x, P.a.y, P.a.z), n(P.a.w, P.b.x, P.b.y);
    float denom = n.dot(r.d);
    if (fabsf(denom) < 1e-6f) return false;
    float t = (pt - r.o).dot(n) / denom;
    if (t < t_min || t > t_max) return false;
    t_out = t;
    if (REC) {
        rec->t = t; rec->p = r.at(t);
        set_face_normal(rec, r, n); rec->mat = P.mat;
    }
    return true;
}

// Cylinder::hit, primitives.hpp:406-422 (side :441-486, caps :488-504)
template <bool REC>
PCL_D bool hit_cylinder(const PrimView& P, const Ray& r, float t_min, float t_max, float& t_out, HitRec* rec) {
    const f3 c(P.a.x, P.a.y, P.a.z), ax(P.a.w, P.b.x, P.b.y);
    const float rad = P.b.z, half_h = P.b.w;
    bool found = false;
    float closest = t_max;
    int part = -1;          // 0 side, 1 cap +, 2 cap -
    float side_h = 0.0f;
    // side
    {
        f3 oc = r.o - c;
        float dda = r.d.dot(ax), oca = oc.dot(ax);
        f3 dp = r.d - ax * dda, ocp = oc - ax * oca;
        float A = dp.dot(dp), B = 2.0f * ocp.dot(dp), C = ocp.dot(ocp) - rad * rad;
        if (A < 1e-9f) {
            if (!(fabsf(B) < 1e-9f)) {
                const float t = -C / B;
                if (!(t < t_min || t > closest)) {
                    const float h = oca + t * dda;
                    if (!(fabsf(h) > half_h)) { closest = t; found = true; part = 0; side_h = h; }
                }
            }
        } else {
            float disc = B * B - 4.0f * A * C;
            if (!(disc < 0.0f)) {
                float sq = sqrtf(disc);
                float t0 = (-B - sq) / (2.0f * A), t1 = (-B + sq) / (2.0f * A);
                for (int k = 0; k < 2; ++k) {
                    float t = k == 0 ? t0 : t1;
                    if (t < t_min || t > closest) continue;
                    float h = (oca + t * dda);
                    if (fabsf(h) > half_h) continue;
                    closest = t; found = true; part = 0; side_h = h;
                    break;
                }
            }
        }
    }
    // caps, sign +1 then -1
    for (int k = 0; k < 2; ++k) {
        const float sign = k == 0 ? 1.0f : -1.0f;
        f3 cc = c + ax * (sign * half_h);
        f3 nrm = ax * sign;
        float denom = nrm.dot(r.d);
        if (fabsf(denom) < 1e-6f) continue;
        float t = (cc - r.o).dot(nrm) / denom;
        if (t < t_min || t > closest) continue;
        f3 p = r.at(t);
        f3 d = p - cc;
        if (d.dot(d) > rad * rad) continue;
        closest = t; found = true; part = 1 + k;
    }
    if (!found) return false;
    t_out = closest;
    if (REC) {
        rec->t = closest; rec->mat = P.mat;
        if (part == 0) {
            rec->p = r.at(closest);
            f3 axis_pt = c + ax * side_h;
            f3 on = (rec->p - axis_pt) / rad;
            set_face_normal(rec, r, on);
        } else {
            const float sign = part == 1 ? 1.0f : -1.0f;
            rec->p = r.at(closest);
            set_face_normal(rec, r, ax * sign);
        }
    }
    return true;
}

// Cone::hit, primitives.hpp:517-530 (side :555-589, base :591-605)
template <bool REC>
PCL_D bool hit_cone(const PrimView& P, const Ray& r, float t_min, float t_max, float& t_out, HitRec* rec) {
    const f3 apex(P.a.x, P.a.y, P.a.z), ax(P.a.w, P.b.x, P.b.y), base(P.c.x, P.c.y, P.c.z);
    const float rad = P.b.z, height = P.b.w;
    bool found = false;
    float closest = t_max;
    int part = -1;
    float side_h = 0.0f, k = 0.0f;
    {
        f3 oc = r.o - apex;
        float dda = r.d.dot(ax), oca = oc.dot(ax);
        k = rad / height;
        float k2 = k * k;
        f3 dp = r.d - ax * dda, ocp = oc - ax * oca;
        float A = dp.dot(dp) - k2 * dda * dda;
        float B = 2.0f * (dp.dot(ocp) - k2 * dda * oca);
        float C = ocp.dot(ocp) - k2 * oca * oca;
        float disc = B * B - 4.0f * A * C;
        if (!(disc < 0.0f || fabsf(A) < 1e-12f)) {
            float sq = sqrtf(disc);
            float t0 = (-B - sq) / (2.0f * A), t1 = (-B + sq) / (2.0f * A);
            for (int q = 0; q < 2; ++q) {
                float t = q == 0 ? t0 : t1;
                if (t < t_min || t > closest) continue;
                float h = oca + t * dda;
                if (h < 0.0f || h > height) continue;
                closest = t; found = true; part = 0; side_h = h;
                break;
            }
        }
    }
    {
        f3 nrm = -ax;
        float denom = nrm.dot(r.d);
        if (!(fabsf(denom) < 1e-6f)) {
            float t = (base - r.o).dot(nrm) / denom;
            if (!(t < t_min || t > closest)) {
                f3 p = r.at(t);
                f3 d = p - base;
                if (!(d.dot(d) > rad * rad)) { closest = t; found = true; part = 1; }
            }
        }
    }
    if (!found) return false;
    t_out = closest;
    if (REC) {
        rec->t = closest; rec->mat = P.mat;
        rec->p = r.at(closest);
        if (part == 0) {
            f3 axis_pt = apex + ax * side_h;
            f3 radial = rec->p - axis_pt;
            f3 on = (radial.normalized() - ax * k).normalized();
            set_face_normal(rec, r, on);
        } else {
            set_face_normal(rec, r, -ax);
        }
    }
    return true;
}

template <bool REC>
PCL_D bool hit_prim(const PrimView& P, const Ray& r, float t_min, float t_max, float& t, HitRec* rec) {
    switch (P.kind) {
    case PCL_PRIM_TRIANGLE: return hit_triangle<REC>(P, r, t_min, t_max, t, rec);
    case PCL_PRIM_SPHERE: return hit_sphere<REC>(P, r, t_min, t_max, t, rec);
    case PCL_PRIM_CUBE: return hit_cube<REC>(P, r, t_min, t_max, t, rec);
    case PCL_PRIM_PLANE: return hit_plane<REC>(P, r, t_min, t_max, t, rec);
    case PCL_PRIM_CYLINDER: return hit_cylinder<REC>(P, r, t_min, t_max, t, rec);
    case PCL_PRIM_CONE: return hit_cone<REC>(P, r, t_min, t_max, t, rec);
    default: return false;
    }
}

// ---------------------------------------------------------------------------
// BVH traversal

struct DevScene {
    const Node4* nodes;
    const PrimRec* prims;
    const MatRec* mats;
    const LightRec* lights;
    const int32_t* slot_of;    // insertion id -> slot, for the linear scan
    int32_t n_prims, n_lights, root, use_bvh;
    float root_box[6];
    float bg[3], amb[3];
};

// Per-ray slab setup: 1/d per axis as AABB::hit computes it (1.0f / d[a]).
// The slab's t0/t1 swap when inv_d < 0 (primitives.hpp:41): the entry face
// of axis a is then the box's max face.  A Node4 holds the four children's
// min faces lx, ly, lz at bytes 0, 16, 32 and max faces hx, hy, hz at 48,
// 64, 80; sw[a] (0 or 48) moves axis a's entry-face load from the min to the
// max face and its exit-face load the other way, so a node step loads each
// child's entry and exit faces directly instead of swapping two computed
// distances per axis and child.
struct SlabRay {
    float ox, oy, oz, ix, iy, iz;
    bool nx, ny, nz;   // inv_d < 0: the slab's t0/t1 swap
    uint32_t sw[3];
};

PCL_D SlabRay make_slab_ray(const Ray& r) {
    SlabRay s;
    s.ox = r.o.x; s.oy = r.o.y; s.oz = r.o.z;
    s.ix = 1.0f / r.d.x; s.iy = 1.0f / r.d.y; s.iz = 1.0f / r.d.z;
    s.nx = s.ix < 0.0f; s.ny = s.iy < 0.0f; s.nz = s.iz < 0.0f;
    s.sw[0] = s.nx ? 48u : 0u;
    s.sw[1] = s.ny ? 48u : 0u;
    s.sw[2] = s.nz ? 48u : 0u;
    return s;
}

// AABB::hit (primitives.hpp:38-49).  The running t_min only grows and t_max
// only shrinks, so the early exit never changes the answer: the box passes
// iff the final t_max >= final t_min.  `near` is the final t_min; a box
// that passed with some t_max passes with a smaller t_max' iff near <= t_max'
// (used to re-check deferred right children exactly as the recursion would).
// The reference's "t0 > t_min ? t0 : t_min" / "t1 < t_max ? t1 : t_max"
// (src/primitives.hpp:40-46) as IEEE maxNum / minNum (v_max_f32 / v_min_f32,
// no compare + select).  Same values here: a NaN face (0 * inf, origin on the
// plane) leaves the bound unchanged in both forms; t_min starts at a positive
// t_min of the query and only grows, so no signed zero reaches it; t_max can
// become a zero of either sign, which only meets comparisons with a positive
// t_min, where the signs compare equal.  near (= t_min) is therefore
// bit-identical and every decision the same.
PCL_D bool slab_test(float lx, float ly, float lz, float hx, float hy, float hz, const SlabRay& s,
                     float t_min, float t_max, float& near) {
    float t0 = (lx - s.ox) * s.ix, t1 = (hx - s.ox) * s.ix;
    if (s.nx) { float x = t0; t0 = t1; t1 = x; }
    t_min = fmaxf(t0, t_min);
    t_max = fminf(t1, t_max);
    t0 = (ly - s.oy) * s.iy; t1 = (hy - s.oy) * s.iy;
    if (s.ny) { float x = t0; t0 = t1; t1 = x; }
    t_min = fmaxf(t0, t_min);
    t_max = fminf(t1, t_max);
    t0 = (lz - s.oz) * s.iz; t1 = (hz - s.oz) * s.iz;
    if (s.nz) { float x = t0; t0 = t1; t1 = x; }
    t_min = fmaxf(t0, t_min);
    t_max = fminf(t1, t_max);
    near = t_min;
    return !(t_max < t_min);
}

// slab_test with the faces already ordered: (nx, ny, nz) are the entry faces
// and (fx, fy, fz) the exit faces of the ray's slabs.  t0 = (entry - o) *
// inv_d and t1 = (exit - o) * inv_d are the swapped (t0, t1) of slab_test
// bit for bit: the same operands, the same operations.
PCL_D bool slab_test_ordered(float nx, float ny, float nz, float fx, float fy, float fz, const SlabRay& s,
                             float t_min, float t_max, float& near) {
    t_min = fmaxf((nx - s.ox) * s.ix, t_min);
    t_max = fminf((fx - s.ox) * s.ix, t_max);
    t_min = fmaxf((ny - s.oy) * s.iy, t_min);
    t_max = fminf((fy - s.oy) * s.iy, t_max);
    t_min = fmaxf((nz - s.oz) * s.iz, t_min);
    t_max = fminf((fz - s.oz) * s.iz, t_max);
    near = t_min;
    return !(t_max < t_min);
}

// Per-lane traversal stack: the top K entries in LDS (a ring, one column per
// lane), older entries spilled to a private array when the path is deeper
// than K.  Entry = (child ref, entry distance of its box).
template <int K, int SPILL>
struct LaneStack {
    PCL_LDS unsigned long long* lds;   // this lane's column base (ref | near << 32)
    int shift;                         // log2 of the lanes per block (the column stride)
    int sp;
    unsigned long long spill[SPILL];
    // lanes per block, a power of two: entries are addressed with a shift, not
    // an integer multiply (a quarter-rate instruction on every push and pop)
    PCL_D void set_stride(int lanes) { shift = __builtin_ctz((unsigned)lanes); }
    PCL_D void push(int32_t ref, float near) {
        const int slot = (sp & (K - 1)) << shift;
        if (sp >= K) spill[sp - K] = lds[slot];
        lds[slot] = (unsigned long long)(uint32_t)ref | ((unsigned long long)__float_as_uint(near) << 32);
        ++sp;
    }
    PCL_D int32_t pop(float& near) {
        --sp;
        const int slot = (sp & (K - 1)) << shift;
        const unsigned long long e = lds[slot];
        if (sp >= K) lds[slot] = spill[sp - K];
        near = __uint_as_float((uint32_t)(e >> 32));
        return (int32_t)(uint32_t)e;
    }
};

// Work counters of the algorithmic-bytes model: 4-wide nodes entered and
// primitive tests per kind.  ON = false compiles them away.
template <bool ON>
struct TravCount {
    uint32_t nodes = 0, t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0;
    PCL_D void node() { if (ON) nodes++; }
    PCL_D void test(int kind) {
        if (!ON) return;
        switch (kind) {
        case 0: t0++; break;
        case 1: t1++; break;
        case 2: t2++; break;
        case 3: t3++; break;
        case 4: t4++; break;
        default: t5++; break;
        }
    }
};

// Scene::hit (ANY = false) / Scene::occluded (ANY = true), src/scene.hpp:55-81,
// src/bvh.hpp:31-48, over 4-wide nodes that collapse the reference-topology
// binary tree.  The recursion tests a node's box when it calls the node, with
// the t_max of that moment; here all children of a 4-wide node get their slab
// distances at once, children 1..3 are pushed with their box's entry distance
// and each is re-checked against the current t_max when popped, so every box
// of the cut is checked at the moment the recursion would check it.  The
// binary nodes the cut skips need no test of their own: a skipped node's box
// is the union of its children's boxes (min/max of the same floats), the slab
// test's rounding is monotone in the box faces and NaN slabs (0 * inf) impose
// no bound in either, so a child that passes implies its skipped ancestor
// passed at that ancestor's (earlier, larger) t_max; and a child that fails
// visits nothing, as the recursion below a failing ancestor would.  Hence the
// leaves are met in the recursion's order with the recursion's t_max.  A
// primitive is accepted when its test passes with t_max = best t so far
// (t == t_max accepted), so among equal t the last leaf in DFS order wins, as
// in the reference.  Returns the winning slot (-1 on a miss); for closest hits
// t_best is its t and t_acc the t_max it was accepted with (needed to rebuild
// its record: a cone's answer depends on t_max).
// One 4-wide node `ref` of the traversal: slab distances of the four child
// boxes with the current interval; children 3..1 that pass are pushed with
// their entry distance; returns true with ref = child 0 when child 0 passes.
template <int K, int SPILL, class TC>
PCL_D bool node_step(const DevScene& S, const SlabRay& s, float t_min, float t_max, int32_t& ref,
                     LaneStack<K, SPILL>& st, TC& cnt) {
    cnt.node();
    // 32-bit byte offsets from the node array's base (under 4 GiB, checked at
    // upload): with a uniform base the loads take it from scalar registers
    // and each address costs one 32-bit add
    const char* nb = reinterpret_cast<const char*>(S.nodes);
    const uint32_t off = (uint32_t)ref << 7;
    const float4 ex = gload_f4(nb + (off + s.sw[0])), ey = gload_f4(nb + (off + s.sw[1] + 16u)),
                 ez = gload_f4(nb + (off + s.sw[2] + 32u));
    const float4 fx = gload_f4(nb + (off - s.sw[0] + 48u)), fy = gload_f4(nb + (off - s.sw[1] + 64u)),
                 fz = gload_f4(nb + (off - s.sw[2] + 80u));
    const int4 ch = gload_i4(nb + (off + 96u));
    float n0, n1, n2, n3;
    const bool p0 = slab_test_ordered(ex.x, ey.x, ez.x, fx.x, fy.x, fz.x, s, t_min, t_max, n0);
    const bool p1 = slab_test_ordered(ex.y, ey.y, ez.y, fx.y, fy.y, fz.y, s, t_min, t_max, n1);
    const bool p2 = slab_test_ordered(ex.z, ey.z, ez.z, fx.z, fy.z, fz.z, s, t_min, t_max, n2);
    const bool p3 = slab_test_ordered(ex.w, ey.w, ez.w, fx.w, fy.w, fz.w, s, t_min, t_max, n3);
    if (p3) st.push(ch.w, n3);
    if (p2) st.push(ch.z, n2);
    if (p1) st.push(ch.y, n1);
    if (p0) ref = ch.x;
    return p0;
}

template <bool ANY, int K, int SPILL, class TC>
PCL_D int traverse(const DevScene& S, const Ray& r, float t_min, float t_max, LaneStack<K, SPILL>& st,
                   float& t_best, float& t_acc, TC& cnt) {
    t_best = t_max; t_acc = t_max;
    int best = -1;
    if (S.root == kEmptyRoot) return -1;
    if (!S.use_bvh) {   // linear scan in insertion order, scene.hpp:59-69,77-80
        float closest = t_max;
        for (int i = 0; i < S.n_prims; ++i) {
            const int slot = gptr(S.slot_of)[i];
            PrimView P = load_prim(S.prims, slot);
            float t;
            cnt.test(P.kind);
            if (hit_prim<false>(P, r, t_min, closest, t, nullptr)) {
                if (ANY) return slot;
                best = slot; t_acc = closest; closest = t;
            }
        }
        t_best = closest;
        return best;
    }
    const SlabRay s = make_slab_ray(r);
    float near;
    if (!slab_test(S.root_box[0], S.root_box[1], S.root_box[2], S.root_box[3], S.root_box[4], S.root_box[5],
                   s, t_min, t_max, near))
        return -1;
    int32_t ref = S.root;
    st.sp = 0;
    while (true) {
        if (ref >= 0) {
            if (node_step(S, s, t_min, t_max, ref, st, cnt)) continue;
        } else {
            const int slot = ~ref;
            PrimView P = load_prim(S.prims, slot);
            float t;
            cnt.test(P.kind);
            if (hit_prim<false>(P, r, t_min, t_max, t, nullptr)) {
                if (ANY) return slot;
                best = slot; t_acc = t_max; t_max = t;
            }
        }
        bool next = false;
        while (st.sp > 0) {
            float nn;
            const int32_t rr = st.pop(nn);
            if (ANY || nn <= t_max) { ref = rr; next = true; break; }
        }
        if (!next) break;
    }
    t_best = t_max;
    return best;
}

}  // namespace pcl
