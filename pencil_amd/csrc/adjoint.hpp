// adjoint.hpp - reverse-mode derivatives of a pixel's radiance with respect
// to every material's albedo and emission, by replaying the pixel's paths.
//
// The adjoint pass re-runs render_pixel's program (the same RNG draws, the
// same float operations, so the same paths and the same values) and pushes a
// weight W down each sample (k_adjoint pushes 1 and scales the pixel's sums
// by its image adjoint afterwards; every term is channel-diagonal):
//
//   Whitted trees (trace(), renderer.hpp:403-822): every node carries the
//   weight of its value in the root.  A frame stores it when pushed; leaf
//   terms (emission, the albedo-scaled light sum and ambient) are credited
//   where the node is shaded, and the terms that multiply a child's value
//   (metal rho*albedo*L, glass (kr*Lr + kt*Lt)*albedo) when the child's value
//   folds back into the frame.  Children get weight * d parent / d child.
//
//   Path chains (trace_path(), renderer.hpp:829-972): R = sum_k T_k (e_k +
//   a_k Q_k) (+ the background, or a glass subtree, at the end) with
//   T_{k+1} = T_k a_k s_k.  The vertices are taped on the forward walk
//   (material, T_k, Q_k, s_k; at most `bounces` of them) and a backward sweep
//   over the tape forms the suffix radiance S_k, giving
//   dR/da_k = T_k (Q_k + s_k S_{k+1}) and dR/de_k = T_k without dividing by
//   any albedo.  A glass vertex hands W*T to the Whitted adjoint of its
//   subtree.  trace_path_mis (renderer.hpp:1205-1335) differs in the light
//   term only, albedo * I * (ndotl w_light + w_brdf) with geometric weights:
//   its Q_k part is that term at unit albedo.
//
// Products are formed in float like the forward pass and accumulated in FP64
// per lane.
#pragma once

#include "render.hpp"

namespace pcl {

constexpr int kAdjMaxMats = PCL_ADJ_MAX_MATERIALS;

// Per-lane gradient accumulator for NM materials (6 parameters each).
template <int NM>
struct Grad {
    double g[NM * 6];
    PCL_D void clear() {
        for (int i = 0; i < NM * 6; ++i) g[i] = 0.0;
    }
    // d/d albedo_m += W * v (per channel)
    PCL_D void albedo(int m, const f3& W, const f3& v) {
        g[m * 6 + 0] += (double)W.x * (double)v.x;
        g[m * 6 + 1] += (double)W.y * (double)v.y;
        g[m * 6 + 2] += (double)W.z * (double)v.z;
    }
    // d/d emission_m += W
    PCL_D void emission(int m, const f3& W) {
        g[m * 6 + 3] += (double)W.x;
        g[m * 6 + 4] += (double)W.y;
        g[m * 6 + 5] += (double)W.z;
    }
};

// A Whitted frame with its node's weight in the root value and its material.
struct AdjFrame {
    f3 albedo, acc, ro, rd, W;
    float w0, w1;
    int type, depth, mat;
};

// whitted() of render.hpp with the adjoint: Wc is the weight of the value of
// the node being evaluated (the value of the ray r at `depth`).
template <class LC, int NM>
PCL_D f3 whitted_adj(const DevScene& S, const RSettings& st, Ray r, int depth, Rng& rng, AdjFrame* F, int nf,
                     Stack& stk, LC& c, f3 Wc, Grad<NM>& G) {
    f3 result;
    while (true) {
        c.rays++;
        bool descended = false;
        if (depth <= 0) {
            result = f3(0, 0, 0);
        } else {
            HitRec rec;
            if (!scene_hit(S, r, 1e-4f, 1e9f, rec, stk, c)) {
                result = ld3(S.bg);
            } else {
                int kind; f3 albedo, emission, spec_color; float refl, ior, spec, shin;
                load_mat(S.mats, rec.mat, kind, albedo, emission, refl, ior, spec, shin, spec_color);
                f3 q;
                f3 local = whitted_local(S, st, r, rec, albedo, emission, spec, shin, spec_color, rng, stk, c, &q);
                if (kind == PCL_MAT_METAL && refl > 0.0f) {
                    AdjFrame& f = F[nf++];
                    f.type = FR_METAL; f.albedo = albedo; f.acc = local; f.w0 = refl; f.depth = depth;
                    f.W = Wc; f.mat = rec.mat;
                    // value = local * (1 - rho) + L_refl * albedo * rho
                    const f3 Wl = Wc * (1.0f - refl);
                    G.emission(rec.mat, Wl);
                    G.albedo(rec.mat, Wl, q);
                    Wc = Wc * albedo * refl;
                    Ray rr;
                    rr.o = rec.p + rec.normal * st.shadow_bias;
                    rr.d = r.d.normalized().reflect(rec.normal).normalized();
                    r = rr; depth = depth - 1; descended = true;
                } else if (kind == PCL_MAT_DIELECTRIC) {
                    // the light loop ran for its RNG draws only: local is discarded
                    r = push_glass(F, nf, r, rec, albedo, ior, depth, st.shadow_bias);
                    AdjFrame& f = F[nf - 1];
                    f.W = Wc; f.mat = rec.mat;
                    Wc = (f.type == FR_GLASS_TIR) ? Wc * albedo : Wc * albedo * f.w0;
                    depth = depth - 1; descended = true;
                } else {
                    result = local;
                    G.emission(rec.mat, Wc);
                    G.albedo(rec.mat, Wc, q);
                }
            }
        }
        if (descended) continue;
        bool resumed = false;
        while (nf > 0) {
            AdjFrame& f = F[nf - 1];
            if (f.type == FR_METAL) {
                G.albedo(f.mat, f.W * f.w0, result);
                result = f.acc * (1.0f - f.w0) + result * f.albedo * f.w0;
                --nf;
            } else if (f.type == FR_GLASS_TIR) {
                G.albedo(f.mat, f.W, result);
                result = result * f.albedo;
                --nf;
            } else if (f.type == FR_GLASS_REFL) {
                f.acc = result;
                f.type = FR_GLASS_REFR;
                r.o = f.ro; r.d = f.rd;
                depth = f.depth - 1;
                Wc = f.W * f.albedo * f.w1;
                resumed = true;
                break;
            } else {
                f3 g = f.acc * f.w0 + result * f.w1;
                G.albedo(f.mat, f.W, g);
                result = g * f.albedo;
                --nf;
            }
        }
        if (!resumed) return result;
    }
}

// trace_path() of render.hpp with the adjoint for a sample of weight W
// (st.pathtrace == 2: the MIS light loop of trace_path_mis).
template <class LC, int NM>
PCL_D f3 trace_path_adj(const DevScene& S, const RSettings& st, Ray cur, Rng& rng, AdjFrame* F, Stack& stk, LC& c,
                        const f3& W, Grad<NM>& G) {
    const bool mis = st.pathtrace == 2;
    c.rays++;
    const int bounces = st.bounces;
    // tape: material, throughput, d/d albedo of the vertex's own terms, and
    // the factor s/p that carries the suffix radiance (metal rho, RR 1/p)
    int tm[kMaxFrames];
    f3 tT[kMaxFrames], tQ[kMaxFrames];
    float ts[kMaxFrames];
    int nv = 0;
    f3 S_end(0, 0, 0);   // radiance past the last taped vertex, per unit throughput there
    f3 T(1, 1, 1), R(0, 0, 0);
    int depth = bounces;
    while (depth > 0) {
        HitRec rec;
        if (!scene_hit(S, cur, 1e-4f, 1e9f, rec, stk, c)) {
            R += T * ld3(S.bg);
            S_end = ld3(S.bg);
            break;
        }
        int kind; f3 albedo, emission, spec_color; float refl, ior, spec, shin;
        load_mat(S.mats, rec.mat, kind, albedo, emission, refl, ior, spec, shin, spec_color);
        R += T * emission;
        const f3 sn = rec.normal;
        f3 q(0, 0, 0);
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
            if (!mis) {
                R += T * albedo * ld3(L.I) * ndotl;
                q += ld3(L.I) * ndotl;
            } else if (!(ndotl <= 0.0f)) {
                // the MIS term is albedo * (I ndotl w_light + I w_brdf) with weights
                // set by the geometry alone: its albedo derivative is the term at unit albedo
                R += T * mis_light_term(L, albedo, ld3(L.I), Ld, dist, ndotl);
                q += mis_light_term(L, f3(1, 1, 1), ld3(L.I), Ld, dist, ndotl);
            }
        }
        R += T * albedo * ld3(S.amb);
        q += ld3(S.amb);
        const f3 T_in = T;
        float inv_p = 1.0f;
        bool stop = false;
        if (bounces - depth > 3) {   // Russian roulette, :922-926
            float p = fmaxf(T.x, fmaxf(T.y, T.z));
            if (rng.unit() > p) stop = true;
            else { T = T / p; inv_p = 1.0f / p; }
        }
        if (stop) {   // nothing follows this vertex
            tm[nv] = rec.mat; tT[nv] = T_in; tQ[nv] = q; ts[nv] = 0.0f; ++nv;
            break;
        }
        f3 nd;
        f3 NT = T * albedo;
        float s = inv_p;
        if (kind == PCL_MAT_METAL && refl > 0.0f) {
            f3 rf = cur.d.normalized().reflect(sn).normalized();
            f3 up = (fabsf(rf.y) < 0.9f) ? f3(0, 1, 0) : f3(1, 0, 0);
            f3 t1 = rf.cross(up).normalized();
            f3 t2 = rf.cross(t1).normalized();
            float r1 = rng.unit() * 0.1f;
            float r2 = rng.unit() * 6.28f;
            nd = (rf + t1 * (r1 * cosf_cr(r2)) + t2 * (r1 * sinf_cr(r2))).normalized();
            NT = T * albedo * refl;
            s = inv_p * refl;
        } else if (kind == PCL_MAT_DIELECTRIC) {
            // R += T_in (e + a q) + T (glass - e); the subtree's weight is W * T
            int nf = 0;
            Ray rr = push_glass(F, nf, cur, rec, albedo, ior, depth, st.shadow_bias);
            F[0].W = W * T; F[0].mat = rec.mat;
            const f3 Wsub = (F[0].type == FR_GLASS_TIR) ? W * T * albedo : W * T * albedo * F[0].w0;
            f3 g = whitted_adj(S, st, rr, depth - 1, rng, F, nf, stk, c, Wsub, G);
            R += T * g - T * emission;
            G.emission(rec.mat, W * (T_in - T));
            G.albedo(rec.mat, W * T_in, q);
            S_end = emission + albedo * q + (g - emission) * inv_p;   // suffix at this vertex
            // the glass vertex is folded into S_end; earlier vertices see it as their S_{k+1}
            break;
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
        tm[nv] = rec.mat; tT[nv] = T_in; tQ[nv] = q; ts[nv] = s; ++nv;
        cur.o = rec.p + rec.normal * st.shadow_bias;
        cur.d = nd;
        T = NT;
        depth--;
    }
    // backward sweep: S = radiance after vertex k per unit throughput T_{k+1}
    f3 Sx = S_end;
    for (int k = nv - 1; k >= 0; --k) {
        int kind; f3 albedo, emission, spec_color; float refl, ior, spec, shin;
        load_mat(S.mats, tm[k], kind, albedo, emission, refl, ior, spec, shin, spec_color);
        const f3 WT = W * tT[k];
        G.emission(tm[k], WT);
        G.albedo(tm[k], WT, tQ[k] + Sx * ts[k]);
        Sx = emission + albedo * tQ[k] + albedo * Sx * ts[k];
    }
    return R;
}

}  // namespace pcl
