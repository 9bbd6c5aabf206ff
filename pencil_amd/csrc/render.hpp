// render.hpp - camera, Whitted trace() and trace_path() on the device.
//
// One lane owns one pixel: the reference shares one xorshift stream across
// all n*n samples of a pixel and the number of draws per sample depends on
// the path (src/renderer.hpp:313), so a pixel's samples run in order in its
// lane.  Whitted recursion (metal: one child, glass: reflect then refract) is
// an explicit frame stack that combines child results exactly the way the
// recursive return expressions do, so HDR values are the reference's own.
#pragma once

#include "intersect.hpp"

namespace pcl {

struct RSettings {
    int32_t width, height, samples, max_depth;
    float shadow_bias;
    int32_t soft_samples, pathtrace, bounces;
};

// Per-lane work counters; COUNT adds the per-node / per-test counts.
template <bool COUNT>
struct LaneCount {
    uint32_t rays = 0;      // reference convention: trace()/trace_path() calls
    uint32_t closest = 0;   // closest-hit queries
    uint32_t anyhit = 0;    // any-hit (shadow) queries
    TravCount<COUNT> trav;
};

constexpr int kStackK = 16;       // LDS entries per lane
constexpr int kStackSpill = 96;   // private overflow entries per lane
constexpr int kMaxFrames = 16;    // Whitted recursion depth the frame stack holds

using Stack = LaneStack<kStackK, kStackSpill>;

// Scene::hit with the full record of the winner.
template <class LC>
PCL_D bool scene_hit(const DevScene& S, const Ray& r, float t_min, float t_max, HitRec& rec, Stack& st, LC& c) {
    c.closest++;
    float tb, ta;
    int slot = traverse<false>(S, r, t_min, t_max, st, tb, ta, c.trav);
    if (slot < 0) return false;
    PrimView P = load_prim(S.prims, slot);
    float t;
    hit_prim<true>(P, r, t_min, ta, t, &rec);
    return true;
}

template <class LC>
PCL_D bool scene_occluded(const DevScene& S, const Ray& r, float t_min, float t_max, Stack& st, LC& c) {
    c.anyhit++;
    float tb, ta;
    return traverse<true>(S, r, t_min, t_max, st, tb, ta, c.trav) >= 0;
}

PCL_D f3 ld3(const float* p) { return f3(p[0], p[1], p[2]); }

// Field-wise 16-byte global loads of a 48-byte light record.
PCL_D LightRec load_light(const LightRec* L, int i) {
    const float4* q = reinterpret_cast<const float4*>(L + i);
    const float4 a = gload_f4(q), b = gload_f4(q + 1), c = gload_f4(q + 2);
    LightRec r;
    r.kind = __float_as_int(a.x);
    r.pos[0] = a.y; r.pos[1] = a.z; r.pos[2] = a.w;
    r.I[0] = b.x; r.I[1] = b.y; r.I[2] = b.z;
    r.radius = b.w;
    r.emissive = __float_as_int(c.x);
    return r;
}

// Camera::ray_at with DOF (camera.cpp:14-68)
PCL_D Ray camera_ray(const CamRec& C, float u, float v, Rng& rng) {
    const f3 ll = ld3(C.lower_left), hz = ld3(C.horizontal), vt = ld3(C.vertical), org = ld3(C.origin);
    f3 dir = ll + hz * u + vt * v - org;
    f3 o = org;
    if (C.aperture > 0.0f) {
        float fd = C.focus_dist;
        const f3 forward = (ll + hz * 0.5f + vt * 0.5f - org).normalized();
        const float denom = dir.normalized().dot(forward);
        const float focus_t = (fabsf(denom) > 1e-6f) ? fd / denom : fd;
        f3 focal = org + dir.normalized() * focus_t;
        float dx, dy;
        if (C.blades <= 2) {
            rng.unit_disc(dx, dy);
        } else {   // polygonal bokeh by rejection, camera.cpp:114-146
            bool inside = false;
            for (int attempt = 0; attempt < 16 && !inside; ++attempt) {
                rng.unit_disc(dx, dy);
                float angle_step = 2.0f * 3.14159265358979f / C.blades;
                float rr = sqrtf(dx * dx + dy * dy);
                float theta = atan2f_cr(dy, dx);
                if (theta < 0) theta += 2.0f * 3.14159265358979f;
                int sector = (int)(theta / angle_step);
                float local_angle = theta - sector * angle_step;
                float half_step = angle_step * 0.5f;
                float max_r = cosf_cr(half_step) / cosf_cr(local_angle - half_step);
                inside = rr <= max_r;
            }
            if (!inside) rng.unit_disc(dx, dy);
        }
        f3 off = ld3(C.u) * dx + ld3(C.v) * dy;
        off = off * C.aperture;
        o = org + off;
        dir = focal - o;
    }
    Ray r;
    r.o = o; r.d = dir.normalized();
    return r;
}

PCL_D void load_mat(const MatRec* M, int id, int& kind, f3& albedo, f3& emission, float& refl, float& ior,
                    float& spec, float& shin, f3& spec_color) {
    // 64-byte record as four 16-byte global loads
    const float4* q = reinterpret_cast<const float4*>(M + id);
    const float4 a = gload_f4(q), b = gload_f4(q + 1), c = gload_f4(q + 2), d = gload_f4(q + 3);
    kind = __float_as_int(a.x);
    albedo = f3(a.y, a.z, a.w);
    refl = b.x; ior = b.y; spec = b.z; shin = b.w;
    spec_color = f3(c.x, c.y, c.z);
    emission = f3(c.w, d.x, d.y);
}

// Disc-jittered point light position (renderer.hpp:524-535 / :891-900)
PCL_D f3 jitter_light(const f3& lpos, float radius, const f3& p, Rng& rng) {
    f3 to_light = lpos - p;
    f3 axis = to_light.normalized();
    f3 up = (fabsf(axis.y) < 0.9f) ? f3(0, 1, 0) : f3(1, 0, 0);
    f3 t1 = axis.cross(up).normalized();
    f3 t2 = axis.cross(t1).normalized();
    float dx, dy;
    rng.unit_disc(dx, dy);
    return lpos + (t1 * dx + t2 * dy) * radius;
}

// The light loop + ambient of Renderer::trace (renderer.hpp:514-577).  With
// dq, also d local / d albedo (per channel): the lights' unoccluded I*ndotl
// sample means plus the ambient term.
template <class LC>
PCL_D f3 whitted_local(const DevScene& S, const RSettings& st, const Ray& r, const HitRec& rec, const f3& albedo,
                       const f3& emission, float spec, float shin, const f3& spec_color, Rng& rng, Stack& stk,
                       LC& c, f3* dq = nullptr) {
    const f3 sn = rec.normal;
    f3 local = emission;
    f3 q(0, 0, 0);
    for (int li = 0; li < S.n_lights; ++li) {
        const LightRec L = load_light(S.lights, li);
        const f3 lpos = ld3(L.pos), I = ld3(L.I);
        f3 contrib(0, 0, 0), qc(0, 0, 0);
        const int nsamp = (L.radius > 0.0f) ? st.soft_samples : 1;
        for (int s = 0; s < nsamp; ++s) {
            f3 L_pos = lpos;
            if (L.kind == PCL_LIGHT_POINT && L.radius > 0.0f) L_pos = jitter_light(lpos, L.radius, rec.p, rng);
            f3 Ld;
            float dist = 1.0f;
            if (L.kind == PCL_LIGHT_POINT) { Ld = L_pos - rec.p; dist = Ld.length(); Ld = Ld / dist; }
            else { Ld = -lpos; dist = 1e9f; }
            Ray sr;
            sr.o = rec.p + rec.normal * st.shadow_bias; sr.d = Ld;
            if (!L.emissive && scene_occluded(S, sr, st.shadow_bias, dist - st.shadow_bias, stk, c)) continue;
            float ndotl = clampf(sn.dot(Ld), 0.0f, 1.0f);
            f3 col = albedo * I * ndotl;
            if (dq) qc += I * ndotl;
            if (spec > 0.0f && ndotl > 0.0f) {
                f3 V = -r.d.normalized();
                f3 H = (Ld + V).normalized();
                float ndoth = clampf(sn.dot(H), 0.0f, 1.0f);
                float sp = powf_cr(ndoth, shin);
                col += spec_color * I * (sp * spec);
            }
            contrib += col;
        }
        local += contrib / (float)nsamp;
        if (dq) q += qc / (float)nsamp;
    }
    local += albedo * ld3(S.amb);
    if (dq) *dq = q + ld3(S.amb);
    return local;
}

enum : int { FR_METAL = 0, FR_GLASS_REFL = 1, FR_GLASS_REFR = 2, FR_GLASS_TIR = 3 };

struct Frame {
    f3 albedo, acc, ro, rd;
    float w0, w1;
    int type, depth;
};

// Pushes the glass frame of trace_dielectric (renderer.hpp:1582-1627) and
// returns the reflection ray; the refraction ray is kept in the frame.
template <class FR>
PCL_D Ray push_glass(FR* F, int& nf, const Ray& r, const HitRec& rec, const f3& albedo, float ior, int depth,
                     float bias) {
    const bool entering = rec.front;
    const float safe = fmaxf(ior, 0.01f);
    const float eta = entering ? (1.0f / safe) : safe;
    const f3 nf_ = rec.normal;
    const f3 dn = r.d.normalized();
    const float cos_i = fabsf(dn.dot(rec.normal));
    const float sin_t2 = eta * eta * (1.0f - cos_i * cos_i);
    Ray rr;
    rr.o = rec.p + rec.normal * bias;
    rr.d = dn.reflect(rec.normal).normalized();
    FR& f = F[nf++];
    f.albedo = albedo;
    f.depth = depth;
    if (sin_t2 > 1.0f) {
        f.type = FR_GLASS_TIR;
    } else {
        const float cos_t = sqrtf(1.0f - sin_t2);
        f.rd = (dn * eta + nf_ * (eta * cos_i - cos_t)).normalized();
        f.ro = rec.p - nf_ * bias;
        const float cs = entering ? cos_i : cos_t;
        const float kr = schlick(cs, eta);
        f.w0 = kr;
        f.w1 = 1.0f - kr;
        f.type = FR_GLASS_REFL;
    }
    return rr;
}

// Deferred glass vertices.  A path that hits glass ends there: its sample
// is R + (T * g - T * emission), g the value of trace_dielectric's Whitted
// subtree (renderer.hpp:941-949).  g depends on nothing the path's later
// samples use except through the generator: Whitted's only draws are the
// light jitter, two per unit_disc, soft_samples unit_discs per point light
// with a radius, at every node whose closest hit succeeds -- `draws` per
// shaded node, the same for every node of the scene.  The subtree's rays do
// not depend on those draws (reflection and refraction are deterministic),
// so the per-pixel program can record the glass vertex as a task with the
// generator state it starts from, walk the subtree's closest hits alone to
// count its shaded nodes (none of that when draws == 0), advance the
// generator past the subtree's draws and go on with its next sample; a
// separate launch (core.hip: k_glass) evaluates the subtrees of all tasks,
// from the recorded states, with full waves instead of one lane while its
// wave waits.  A task is 8 float4: the reflection ray, the glass frame
// (push_glass), T, R, emission, the generator state and the sample it
// completes.
struct GlassSink {
    float4* tasks;       // 8 float4 per task
    uint32_t* count;     // tasks appended; may pass cap (those were not deferred)
    uint32_t cap;
    int32_t draws;       // generator draws per shaded Whitted node
    int32_t pixel, k;    // the sample being traced
    int32_t item = 0;    // batch item of the pixel (the batched tracer's finish launch)
    // when set: per task, the walk's count of the subtree's shaded nodes
    // (draws > 0), else the subtree's depth limit: the glass launch's order
    uint32_t* keys = nullptr;
    // Wave-aggregated append (every lane of the branch calls it); the task's
    // index, or -1 when the task array is full and the caller evaluates the
    // vertex itself.
    PCL_D int append(const Frame& f, const Ray& rr, const f3& T, const f3& R, const f3& e, uint32_t rng) {
        const unsigned long long m = __ballot(1);
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((long long)m) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
        base = __shfl(base, leader, 64);
        const uint32_t idx = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (idx >= cap) return -1;
        float4* q = tasks + (size_t)idx * 8;
        q[0] = make_float4(rr.o.x, rr.o.y, rr.o.z, __int_as_float(f.depth));
        q[1] = make_float4(rr.d.x, rr.d.y, rr.d.z, __int_as_float(f.type));
        q[2] = make_float4(f.albedo.x, f.albedo.y, f.albedo.z, f.w0);
        q[3] = make_float4(f.ro.x, f.ro.y, f.ro.z, f.w1);
        q[4] = make_float4(f.rd.x, f.rd.y, f.rd.z, __int_as_float(pixel));
        q[5] = make_float4(T.x, T.y, T.z, __int_as_float(k));
        q[6] = make_float4(R.x, R.y, R.z, __uint_as_float(rng));
        q[7] = make_float4(e.x, e.y, e.z, __int_as_float(item));
        if (keys && draws == 0) keys[idx] = (uint32_t)f.depth;   // no walk: the subtree's depth limit
        return (int)idx;
    }
};

// State of one Renderer::trace recursion (renderer.hpp:403-822, in-core
// branches) between trace() calls: the ray and depth of the next call and the
// frames of the calls waiting for their children.
struct WhittedState {
    Ray r;
    int depth, nf;
};

// One trace() call of the recursion -- its closest-hit query, light loop and
// children -- followed by the return path as far as it goes.  Returns true
// when the outermost call has returned, with its value in `result`; false
// when another trace() call (a child or a refraction sibling) is pending in W.
template <class LC>
PCL_D bool whitted_step(const DevScene& S, const RSettings& st, WhittedState& W, Rng& rng, Frame* F, Stack& stk,
                        LC& c, f3& result) {
    c.rays++;
    if (W.depth <= 0) {
        result = f3(0, 0, 0);
    } else {
        HitRec rec;
        if (!scene_hit(S, W.r, 1e-4f, 1e9f, rec, stk, c)) {
            result = ld3(S.bg);
        } else {
            int kind; f3 albedo, emission, spec_color; float refl, ior, spec, shin;
            load_mat(S.mats, rec.mat, kind, albedo, emission, refl, ior, spec, shin, spec_color);
            f3 local = whitted_local(S, st, W.r, rec, albedo, emission, spec, shin, spec_color, rng, stk, c);
            if (kind == PCL_MAT_METAL && refl > 0.0f) {
                Frame& f = F[W.nf++];
                f.type = FR_METAL; f.albedo = albedo; f.acc = local; f.w0 = refl; f.depth = W.depth;
                Ray rr;
                rr.o = rec.p + rec.normal * st.shadow_bias;
                rr.d = W.r.d.normalized().reflect(rec.normal).normalized();
                W.r = rr; W.depth = W.depth - 1;
                return false;
            } else if (kind == PCL_MAT_DIELECTRIC) {
                W.r = push_glass(F, W.nf, W.r, rec, albedo, ior, W.depth, st.shadow_bias);
                W.depth = W.depth - 1;
                return false;
            } else {
                result = local;
            }
        }
    }
    // return path: fold child results into their parents
    while (W.nf > 0) {
        Frame& f = F[W.nf - 1];
        if (f.type == FR_METAL) {
            // local * (1 - rho) + (incoming * albedo) * rho   (:652-657)
            result = f.acc * (1.0f - f.w0) + result * f.albedo * f.w0;
            --W.nf;
        } else if (f.type == FR_GLASS_TIR) {
            result = result * f.albedo;   // :1609-1611
            --W.nf;
        } else if (f.type == FR_GLASS_REFL) {
            f.acc = result;
            f.type = FR_GLASS_REFR;
            W.r.o = f.ro; W.r.d = f.rd;
            W.depth = f.depth - 1;
            return false;
        } else {
            // (refl * kr + refr * kt) * albedo   (:1622-1626)
            f3 g = f.acc * f.w0 + result * f.w1;
            result = g * f.albedo;
            --W.nf;
        }
    }
    return true;
}

// Renderer::trace from ray r with `depth`, continuing below frames F[0..nf)
// already pushed by the caller (trace_path's glass case).  Returns the value
// the outermost call returns.
template <class LC>
PCL_D f3 whitted(const DevScene& S, const RSettings& st, Ray r, int depth, Rng& rng, Frame* F, int nf, Stack& stk,
                 LC& c) {
    WhittedState W;
    W.r = r; W.depth = depth; W.nf = nf;
    f3 result;
    while (!whitted_step(S, st, W, rng, F, stk, c, result)) {
    }
    return result;
}

// Renderer::trace for max_depth <= 1 (ray casting with direct light): the
// children of a metal or glass hit are depth-0 calls, which count a trace()
// and return black.  Same float operations as whitted() on these inputs --
// the same light loop (and its RNG draws), and the children's black folded
// with the same expressions -- without the frame stack.
template <class LC>
PCL_D f3 whitted_shallow(const DevScene& S, const RSettings& st, const Ray& r, int depth, Rng& rng, Stack& stk,
                         LC& c) {
    c.rays++;
    if (depth <= 0) return f3(0, 0, 0);
    HitRec rec;
    if (!scene_hit(S, r, 1e-4f, 1e9f, rec, stk, c)) return ld3(S.bg);
    int kind; f3 albedo, emission, spec_color; float refl, ior, spec, shin;
    load_mat(S.mats, rec.mat, kind, albedo, emission, refl, ior, spec, shin, spec_color);
    const f3 local = whitted_local(S, st, r, rec, albedo, emission, spec, shin, spec_color, rng, stk, c);
    const f3 zero(0.0f, 0.0f, 0.0f);   // a depth-0 child's value
    if (kind == PCL_MAT_METAL && refl > 0.0f) {
        c.rays++;                                          // trace(reflected, depth - 1)
        return local * (1.0f - refl) + zero * albedo * refl;
    }
    if (kind == PCL_MAT_DIELECTRIC) {   // trace_dielectric: push_glass's TIR test and Fresnel weights
        const bool entering = rec.front;
        const float safe = fmaxf(ior, 0.01f);
        const float eta = entering ? (1.0f / safe) : safe;
        const f3 dn = r.d.normalized();
        const float cos_i = fabsf(dn.dot(rec.normal));
        const float sin_t2 = eta * eta * (1.0f - cos_i * cos_i);
        c.rays++;                                          // reflection child
        if (sin_t2 > 1.0f) return zero * albedo;
        c.rays++;                                          // refraction child
        const float cos_t = sqrtf(1.0f - sin_t2);
        const float kr = schlick(entering ? cos_i : cos_t, eta);
        const f3 g = zero * kr + zero * (1.0f - kr);
        return g * albedo;
    }
    return local;
}

// Power-heuristic light-sample term of trace_path_mis (renderer.hpp:1264-1302)
// for an unoccluded sample with ndotl > 0: albedo * I * (ndotl * w_light)
// + albedo * I * w_brdf, in the reference's operation order.
PCL_D f3 mis_light_term(const LightRec& L, const f3& albedo, const f3& I, const f3& Ld, float dist, float ndotl) {
    const float pdf_brdf = ndotl / 3.14159265f;
    float pdf_light;
    if (L.kind == PCL_LIGHT_POINT && L.radius > 0.0f) {
        const float rr = L.radius;
        float cos_at_light = (-Ld).dot(f3(0, 1, 0));
        if (cos_at_light < 0.01f) cos_at_light = 0.01f;
        pdf_light = (dist * dist) / (3.14159265f * rr * rr * cos_at_light);
    } else {
        pdf_light = 1.0f;
    }
    const float w_light = (pdf_light * pdf_light) / (pdf_light * pdf_light + pdf_brdf * pdf_brdf + 1e-9f);
    const float w_brdf = 1.0f - w_light;
    const f3 contrib_nee = albedo * I * ndotl;
    const f3 contrib_brdf = albedo * I;
    return contrib_nee * w_light + contrib_brdf * w_brdf;
}

// State of one Renderer::trace_path (renderer.hpp:829-972) walk between
// bounces: the next ray, throughput, radiance so far and bounces left.
struct PathState {
    Ray cur;
    f3 T, R;
    int depth;
    bool deferred = false;   // the walk ended at a deferred glass vertex
    int walk_task = -1;      // that vertex's task, when its subtree is still to be walked
};

struct NoSink {
    int32_t draws = 0;
    PCL_D int append(const Frame&, const Ray&, const f3&, const f3&, const f3&, uint32_t) { return -1; }
};

// A node of a deferred subtree's walk: ray, depth of the trace() call, the
// lane whose glass vertex the subtree belongs to, and (lanes of different
// scenes in one wave: the batched tracer's finish launch) that lane's scene
// and shadow bias.
struct WalkNode {
    f3 o, d;
    int depth, owner;
};
template <bool MULTI>
struct WalkEntry : WalkNode {};
template <>
struct WalkEntry<true> : WalkNode {
    const DevScene* scene;
    float bias;
};

// One walked node: the closest hit of trace(r, depth) and the trace() calls
// it makes -- a metal child, or a glass node's reflection and (unless total
// internal reflection) refraction, as whitted() spawns them -- when they
// query (depth - 1 > 0).  Returns the number of children (0..2).
template <bool MULTI>
PCL_D int walk_node(const DevScene& S, float bias, const WalkEntry<MULTI>& e, Stack& stk, bool& hit,
                    WalkEntry<MULTI>& c1, WalkEntry<MULTI>& c2) {
    LaneCount<false> none;   // the rays are counted where the subtree is evaluated
    HitRec rec;
    Ray r;
    r.o = e.o; r.d = e.d;
    hit = scene_hit(S, r, 1e-4f, 1e9f, rec, stk, none);
    if (!hit || e.depth - 1 <= 0) return 0;
    int kind; f3 albedo, emission, spec_color; float refl, ior, spec, shin;
    load_mat(S.mats, rec.mat, kind, albedo, emission, refl, ior, spec, shin, spec_color);
    if constexpr (MULTI) {
        c1.scene = c2.scene = e.scene;
        c1.bias = c2.bias = e.bias;
    }
    if (kind == PCL_MAT_METAL && refl > 0.0f) {
        c1.o = rec.p + rec.normal * bias;
        c1.d = r.d.normalized().reflect(rec.normal).normalized();
        c1.depth = e.depth - 1; c1.owner = e.owner;
        return 1;
    }
    if (kind != PCL_MAT_DIELECTRIC) return 0;
    Frame f[1];
    int nf = 0;
    const Ray rr = push_glass(f, nf, r, rec, albedo, ior, e.depth, bias);
    c1.o = rr.o; c1.d = rr.d; c1.depth = e.depth - 1; c1.owner = e.owner;
    if (f[0].type != FR_GLASS_REFL) return 1;
    c2.o = f[0].ro; c2.d = f[0].rd; c2.depth = e.depth - 1; c2.owner = e.owner;
    return 2;
}

// Index of the n-th set bit of m (n < popcount(m)).
PCL_D int nth_bit(unsigned long long m, uint32_t n) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const uint32_t c = (uint32_t)__popcll(m & ((1ull << w) - 1ull));
        if (n >= c) { n -= c; m >>= w; pos += w; }
    }
    return pos;
}

template <bool MULTI>
PCL_D WalkEntry<MULTI> shfl_entry(const WalkEntry<MULTI>& e, int src) {
    WalkEntry<MULTI> r;
    r.o = f3(__shfl(e.o.x, src, 64), __shfl(e.o.y, src, 64), __shfl(e.o.z, src, 64));
    r.d = f3(__shfl(e.d.x, src, 64), __shfl(e.d.y, src, 64), __shfl(e.d.z, src, 64));
    r.depth = __shfl(e.depth, src, 64);
    r.owner = __shfl(e.owner, src, 64);
    if constexpr (MULTI) {
        const unsigned long long a = (unsigned long long)(uintptr_t)e.scene;
        const uint32_t lo = __shfl((uint32_t)a, src, 64), hi = __shfl((uint32_t)(a >> 32), src, 64);
        r.scene = reinterpret_cast<const DevScene*>((uintptr_t)(((unsigned long long)hi << 32) | lo));
        r.bias = __shfl(e.bias, src, 64);
    }
    return r;
}

// The walks of the wave's deferred glass vertices of this sample, together
// (every active lane of the wave calls it at the sample's end, where they
// reconverge).  The number of nodes a subtree shades does not depend on the
// order its nodes are visited in, so each lane walks depth-first from its
// current node, keeps one sibling as a spare (further ones on a private
// stack: at most one per depth level), and a lane out of work takes another
// lane's spare; the shaded nodes are counted per owning lane.  An owner then
// skips its generator past draws x count values, as the subtree's
// evaluation will draw them.  MULTI: the wave's lanes may render different
// scenes; a node is traced in its owner's scene.
template <bool MULTI>
PCL_D void walk_wave(const DevScene& S, const RSettings& st, const GlassSink& sink, int task, Rng& rng, Stack& stk) {
    const bool mine = task >= 0 && sink.draws > 0;   // a subtree that draws is pending
    const unsigned long long pend = __ballot(mine);
    if (!pend) return;
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = (1ull << lane) - 1ull;
    bool hc = false, hs = false;   // a current node, a spare node
    WalkEntry<MULTI> cur, spare;
    cur.o = cur.d = spare.o = spare.d = f3(0, 0, 0);
    cur.depth = spare.depth = 0; cur.owner = spare.owner = lane;
    if constexpr (MULTI) {
        cur.scene = spare.scene = &S;
        cur.bias = spare.bias = st.shadow_bias;
    }
    WalkEntry<MULTI> priv[kMaxFrames];
    int sp = 0;
    uint32_t count = 0;
    if (mine) {   // the glass frame's two calls: reflection, then refraction
        const float4* t = sink.tasks + (size_t)task * 8;
        const float4 a = t[0], b = t[1], e = t[3], g4 = t[4];
        const int cd = __float_as_int(a.w) - 1;
        if (cd > 0) {
            cur.o = f3(a.x, a.y, a.z); cur.d = f3(b.x, b.y, b.z); cur.depth = cd; hc = true;
            if (__float_as_int(b.w) == FR_GLASS_REFL) {
                spare.o = f3(e.x, e.y, e.z); spare.d = f3(g4.x, g4.y, g4.z); spare.depth = cd; hs = true;
            }
        }
    }
    while (true) {
        if (!hc && sp > 0) { cur = priv[--sp]; hc = true; }
        if (!hc && hs) { cur = spare; hc = true; hs = false; }
        if (!__ballot(hc || hs)) break;   // private stacks are empty once hc is false
        const unsigned long long idle = __ballot(!hc), donors = __ballot(hs);
        if (idle && donors) {
            const uint32_t m = min((uint32_t)__popcll(idle), (uint32_t)__popcll(donors));
            const uint32_t ri = (uint32_t)__popcll(idle & lt), rd = (uint32_t)__popcll(donors & lt);
            const int src = (!hc && ri < m) ? nth_bit(donors, ri) : lane;
            const WalkEntry<MULTI> got = shfl_entry<MULTI>(spare, src);
            if (!hc && ri < m) { cur = got; hc = true; }
            if (hs && rd < m) hs = false;
        }
        bool hit = false;
        const int owner = cur.owner;
        if (hc) {
            WalkEntry<MULTI> c1, c2;
            int nc;
            if constexpr (MULTI) nc = walk_node<true>(*cur.scene, cur.bias, cur, stk, hit, c1, c2);
            else nc = walk_node<false>(S, st.shadow_bias, cur, stk, hit, c1, c2);
            hc = nc >= 1;
            if (hc) cur = c1;
            if (nc == 2) {
                if (!hs) { spare = c2; hs = true; }
                else priv[sp++] = c2;
            }
        }
        for (unsigned long long po = pend; po; po &= po - 1) {
            const int o = __ffsll((long long)po) - 1;
            const uint32_t c = (uint32_t)__popcll(__ballot(hit && owner == o));
            if (lane == o) count += c;
        }
    }
    if (mine) {
        rng.skip((uint32_t)sink.draws * count);
        if (sink.keys) sink.keys[task] = count;
    }
}

// One iteration of trace_path's bounce loop (its caller checks depth > 0):
// closest hit, emission, NEE per light, ambient, Russian roulette, the next
// ray.  Returns false when the walk ends (miss, roulette, glass: the glass
// subtree runs here as trace_dielectric does).  With st.pathtrace == 2,
// trace_path_mis (renderer.hpp:1205-1335), which differs only in the light
// loop (its tangent frame is the diffuse branch's, computed earlier).
// `sink`: where a glass vertex may be deferred (GlassSink); P.deferred is
// then set and P.R is not the sample's value.
template <class LC, class SINK = NoSink, bool WALK = false>
PCL_D bool path_step(const DevScene& S, const RSettings& st, PathState& P, Rng& rng, Frame* F, Stack& stk, LC& c,
                     SINK* sink = nullptr) {
    const bool mis = st.pathtrace == 2;
    const int bounces = st.bounces;
    HitRec rec;
    if (!scene_hit(S, P.cur, 1e-4f, 1e9f, rec, stk, c)) { P.R += P.T * ld3(S.bg); return false; }
    int kind; f3 albedo, emission, spec_color; float refl, ior, spec, shin;
    load_mat(S.mats, rec.mat, kind, albedo, emission, refl, ior, spec, shin, spec_color);
    P.R += P.T * emission;
    const f3 sn = rec.normal;
    for (int li = 0; li < S.n_lights; ++li) {
        const LightRec L = load_light(S.lights, li);
        const f3 lpos = ld3(L.pos);
        f3 Ld;
        float dist = 1.0f;
        if (L.kind == PCL_LIGHT_POINT) {
            f3 L_pos = lpos;
            if (L.radius > 0.0f) L_pos = jitter_light(lpos, L.radius, rec.p, rng);
            Ld = L_pos - rec.p; dist = Ld.length(); Ld = Ld / dist;
        } else {
            Ld = -lpos; dist = 1e9f;
        }
        Ray sr;
        sr.o = rec.p + rec.normal * st.shadow_bias; sr.d = Ld;
        if (!L.emissive && scene_occluded(S, sr, st.shadow_bias, dist - st.shadow_bias, stk, c)) continue;
        float ndotl = clampf(sn.dot(Ld), 0.0f, 1.0f);
        if (!mis) P.R += P.T * albedo * ld3(L.I) * ndotl;
        else if (!(ndotl <= 0.0f)) P.R += P.T * mis_light_term(L, albedo, ld3(L.I), Ld, dist, ndotl);
    }
    P.R += P.T * albedo * ld3(S.amb);
    if (bounces - P.depth > 3) {   // Russian roulette, :922-926
        float p = fmaxf(P.T.x, fmaxf(P.T.y, P.T.z));
        if (rng.unit() > p) return false;
        P.T = P.T / p;
    }
    f3 nd;
    f3 NT = P.T * albedo;
    if (kind == PCL_MAT_METAL && refl > 0.0f) {
        f3 rf = P.cur.d.normalized().reflect(sn).normalized();
        f3 up = (fabsf(rf.y) < 0.9f) ? f3(0, 1, 0) : f3(1, 0, 0);
        f3 t1 = rf.cross(up).normalized();
        f3 t2 = rf.cross(t1).normalized();
        float r1 = rng.unit() * 0.1f;
        float r2 = rng.unit() * 6.28f;
        nd = (rf + t1 * (r1 * cosf_cr(r2)) + t2 * (r1 * sinf_cr(r2))).normalized();
        NT = P.T * albedo * refl;
    } else if (kind == PCL_MAT_DIELECTRIC) {
        int nf = 0;
        Ray rr = push_glass(F, nf, P.cur, rec, albedo, ior, P.depth, st.shadow_bias);
        const int ti = sink ? sink->append(F[0], rr, P.T, P.R, emission, rng.s) : -1;
        if (ti >= 0) {
            if constexpr (WALK) P.walk_task = ti;   // walked with the wave when the sample ends
            P.deferred = true;
            return false;
        }
        f3 g = whitted(S, st, rr, P.depth - 1, rng, F, nf, stk, c);
        P.R += P.T * g - P.T * emission;
        return false;
    } else {
        float r1 = rng.unit() * 6.2831853f;
        float r2 = rng.unit();
        float r2s = sqrtf(r2);
        f3 N = sn;
        f3 up = (fabsf(N.y) < 0.9f) ? f3(0, 1, 0) : f3(1, 0, 0);
        f3 Tt = up.cross(N).normalized();
        f3 B = N.cross(Tt).normalized();
        nd = (Tt * cosf_cr(r1) * r2s + B * sinf_cr(r1) * r2s + N * sqrtf(1.0f - r2)).normalized();
    }
    P.cur.o = rec.p + rec.normal * st.shadow_bias;
    P.cur.d = nd;
    P.T = NT;
    P.depth--;
    return true;
}

// Renderer::trace_path (renderer.hpp:829-972) / trace_path_mis.
template <class LC>
PCL_D f3 trace_path(const DevScene& S, const RSettings& st, Ray cur, Rng& rng, Frame* F, Stack& stk, LC& c) {
    c.rays++;
    PathState P;
    P.cur = cur; P.T = f3(1, 1, 1); P.R = f3(0, 0, 0); P.depth = st.bounces;
    while (P.depth > 0) {
        if (!path_step(S, st, P, rng, F, stk, c)) break;
    }
    return P.R;
}

// render_tile_region's per-pixel body (renderer.hpp:313-387, non-adaptive).
// PT: 0 = Whitted trace() only, 1 = trace_path() only (st.pathtrace 1 or 2),
// 2 = Whitted with max_depth <= 1 (whitted_shallow, no frame stack),
// -1 = chosen per call from st.pathtrace.
template <int PT = -1, class LC>
PCL_D f3 render_pixel(const DevScene& S, const CamRec& C, const RSettings& st, int i, int j, Frame* F, Stack& stk,
                      LC& c) {
    const int W = st.width, H = st.height, n = st.samples;
    Rng rng((uint32_t)(i * 97861u + j * 1973u + 1u));
    f3 acc(0, 0, 0);
    for (int sy = 0; sy < n; ++sy) {
        for (int sx = 0; sx < n; ++sx) {
            float u = (i + (sx + 0.5f) / n) / (float)W;
            float v = 1.0f - (j + (sy + 0.5f) / n) / (float)H;
            u += rng.sym() * (0.5f / (W * n));
            v += rng.sym() * (0.5f / (H * n));
            Ray r = camera_ray(C, u, v, rng);
            if (PT == 2) acc += whitted_shallow(S, st, r, st.max_depth, rng, stk, c);
            else if (PT == 1 || (PT < 0 && st.pathtrace)) acc += trace_path(S, st, r, rng, F, stk, c);
            else acc += whitted(S, st, r, st.max_depth, rng, F, 0, stk, c);
        }
    }
    return acc / (float)(n * n);
}

// Glass subtrees evaluated by a later launch (trace_path with a GlassSink),
// from sample k0 on: `rng` is the pixel's stream as sample k0 finds it and
// `acc` the sum of samples 0..k0-1.  render_pixel's loop, except that once a
// sample is deferred the sum stops: acc then holds the sum of the samples
// before the first deferred one (`first`), and every later sample evaluated
// here is stored in samp (sample k of this pixel at samp[k * stride]), the
// deferred ones being filled in by the glass launch; a last launch adds them
// to acc in sample order, so the pixel gets render_pixel's additions in
// render_pixel's order.  Returns true, with the pixel value in px, when no
// sample was deferred (first = n^2).  WALK: the scene's Whitted draws (soft
// lights), so a deferred subtree's closest hits are walked to skip them; a
// program without the walk is compiled for scenes whose Whitted draws nothing
// (the walk's code alone, never run, costs the megakernel 8 % on the hard-light
// Cornell + mesh frame).
template <bool WALK, bool MULTI, class LC>
PCL_D bool render_samples_defer(const DevScene& S, const CamRec& C, const RSettings& st, int i, int j, int k0,
                                Rng& rng, f3 acc, Frame* F, Stack& stk, LC& c, GlassSink& sink, float* samp,
                                size_t stride, f3& px, int& first) {
    const int W = st.width, H = st.height, n = st.samples;
    first = n * n;
    for (int k = k0; k < n * n; ++k) {
        const int sy = k / n, sx = k - sy * n;
        float u = (i + (sx + 0.5f) / n) / (float)W;
        float v = 1.0f - (j + (sy + 0.5f) / n) / (float)H;
        u += rng.sym() * (0.5f / (W * n));
        v += rng.sym() * (0.5f / (H * n));
        Ray r = camera_ray(C, u, v, rng);
        c.rays++;   // trace_path()
        PathState P;
        P.cur = r; P.T = f3(1, 1, 1); P.R = f3(0, 0, 0); P.depth = st.bounces;
        sink.k = k;
        while (P.depth > 0) {
            if (!path_step<LC, GlassSink, WALK>(S, st, P, rng, F, stk, c, &sink)) break;
        }
        if constexpr (WALK) walk_wave<MULTI>(S, st, sink, P.walk_task, rng, stk);
        if (P.deferred) {
            if (first == n * n) first = k;
        } else if (first == n * n) {
            acc += P.R;
        } else {
            float* o = samp + (size_t)k * stride;
            o[0] = P.R.x; o[1] = P.R.y; o[2] = P.R.z;
        }
    }
    px = (first == n * n) ? acc / (float)(n * n) : acc;
    return first == n * n;
}

// The whole pixel (render_pixel's generator seed, samples 0..n^2-1).
template <bool WALK, class LC>
PCL_D bool render_pixel_defer(const DevScene& S, const CamRec& C, const RSettings& st, int i, int j, Frame* F,
                              Stack& stk, LC& c, GlassSink& sink, float* samp, size_t stride, f3& px, int& first) {
    Rng rng((uint32_t)(i * 97861u + j * 1973u + 1u));
    return render_samples_defer<WALK, false>(S, C, st, i, j, 0, rng, f3(0, 0, 0), F, stk, c, sink, samp, stride,
                                             px, first);
}

// One deferred glass vertex (a GlassSink task, 8 float4 at q): the Whitted
// subtree from the recorded frame and generator state -- whitted() with the
// frame pushed, as path_step runs it, so the same draws -- and the sample's
// value R + (T * g - T * emission).  The task's pixel, sample and item are
// the ints in q[4].w, q[5].w and q[7].w.
template <class LC>
PCL_D f3 glass_task(const DevScene& S, const RSettings& st, const float4* q, Frame* F, Stack& stk, LC& c) {
    const float4 a = q[0], b = q[1], d = q[2], e = q[3], g4 = q[4], h = q[5], r4 = q[6], m = q[7];
    Frame& f = F[0];
    f.depth = __float_as_int(a.w);
    f.type = __float_as_int(b.w);
    f.albedo = f3(d.x, d.y, d.z);
    f.w0 = d.w;
    f.ro = f3(e.x, e.y, e.z);
    f.w1 = e.w;
    f.rd = f3(g4.x, g4.y, g4.z);
    Ray rr;
    rr.o = f3(a.x, a.y, a.z);
    rr.d = f3(b.x, b.y, b.z);
    Rng rng(__float_as_uint(r4.w));
    const f3 gv = whitted(S, st, rr, f.depth - 1, rng, F, 1, stk, c);
    const f3 T(h.x, h.y, h.z), E(m.x, m.y, m.z);
    f3 v(r4.x, r4.y, r4.z);
    v += T * gv - T * E;   // P.R += P.T * g - P.T * emission
    return v;
}

// render_pixel from sample k0 on: `rng` is the pixel's stream as sample k0
// finds it and `acc` the sum of samples 0..k0-1, so the result is
// render_pixel's bit for bit (same sample order, same additions).  Used by
// the batched tracer to finish pixels it has started.
template <int PT = -1, class LC>
PCL_D f3 render_samples(const DevScene& S, const CamRec& C, const RSettings& st, int i, int j, int k0, Rng& rng,
                        f3 acc, Frame* F, Stack& stk, LC& c) {
    const int W = st.width, H = st.height, n = st.samples;
    for (int k = k0; k < n * n; ++k) {
        const int sy = k / n, sx = k - sy * n;
        float u = (i + (sx + 0.5f) / n) / (float)W;
        float v = 1.0f - (j + (sy + 0.5f) / n) / (float)H;
        u += rng.sym() * (0.5f / (W * n));
        v += rng.sym() * (0.5f / (H * n));
        Ray r = camera_ray(C, u, v, rng);
        if (PT == 1 || (PT < 0 && st.pathtrace)) acc += trace_path(S, st, r, rng, F, stk, c);
        else acc += whitted(S, st, r, st.max_depth, rng, F, 0, stk, c);
    }
    return acc / (float)(n * n);
}

}  // namespace pcl
