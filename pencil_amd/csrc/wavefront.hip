// wavefront.hip - batched, persistent wavefront tracer.
//
// Renders a job of many (scene, camera, settings) items in one pass over a
// fixed pool of pixel slots.  Each slot owns one pixel at a time and runs the
// reference's per-pixel program (src/renderer.hpp:309-390: n*n samples in
// order on the pixel's xorshift stream; Whitted trace() trees or
// trace_path() chains) as a state machine that stops at every BVH query.
// One iteration of the machine is five kernels:
//
//   closest  : persistent; pops the ray queue, Scene::hit -> (slot, t_max at
//              acceptance) per query                    [LDS ring stack]
//   shade    : hit record, material, every light sample's contribution and
//              shadow ray (pushed to the shadow queue), bounce / child ray;
//              every RNG draw of the vertex happens here, in the reference's
//              order (light jitter, Russian roulette, bounce)
//   anyhit   : persistent; pops the shadow queue, Scene::occluded -> one
//              bit per sample                           [LDS ring stack]
//   resolve  : folds the unoccluded contributions in the reference's order,
//              unwinds Whitted frames, finishes samples and pixels, pulls the
//              next pixel of the job, and pushes the slot's next query
//   swap     : flips the queues and clears the work heads
//
// Queues are compacted with wave ballots and prefix sums (one atomic per
// wave).  Kernel grids are fixed (persistent waves fetch 64-item chunks), so
// an iteration is captured once as a hipGraph and replayed until the ray
// queue drains.  Slot count, persistent blocks per CU and the LDS stack depth
// are the autotuned knobs.
//
// Floating-point results are the megakernel's (and so the reference's): every
// color is accumulated with the same operations in the same order.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "render.hpp"
#include "wavefront.hpp"

namespace pcl {
namespace wf {

constexpr int kWave = 64;
constexpr int kWfBlock = 128;

enum : int { M_PATH = 0, M_WHITTED = 1 };
// vertex outcome decided by shade, consumed by resolve
enum : int {
    V_NONE = 0,
    V_W_MISS = 1,      // Whitted miss: result = background
    V_W_LEAF = 2,      // Whitted hit, no child: result = local
    V_W_CHILD = 3,     // Whitted metal/glass: child ray spawned, frame pushed
    V_P_MISS = 4,      // path miss: R += T * bg, sample ends
    V_P_BOUNCE = 5,    // path bounce: continue with the new ray
    V_P_END = 6,       // path ends after this vertex (Russian roulette)
    V_P_GLASS = 7      // path glass: Whitted subtree from the pushed frame
};

__device__ __forceinline__ float3 ld(const float* a, int s) { return make_float3(a[3 * s], a[3 * s + 1], a[3 * s + 2]); }
__device__ __forceinline__ f3 ldf(const float* a, int s) { return f3(a[3 * s], a[3 * s + 1], a[3 * s + 2]); }
__device__ __forceinline__ void stf(float* a, int s, const f3& v) { a[3 * s] = v.x; a[3 * s + 1] = v.y; a[3 * s + 2] = v.z; }

// ---- wave-level queue helpers --------------------------------------------

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Appends one entry per lane with `pred`; returns the lane's index (or -1).
__device__ __forceinline__ int wave_append(bool pred, uint32_t* counter) {
    const unsigned long long m = __ballot(pred);
    if (m == 0) return -1;
    const int lane = lane_id();
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader, kWave);
    if (!pred) return -1;
    return (int)(base + __popcll(m & ((1ull << lane) - 1ull)));
}

// Appends `n` entries per lane; returns the lane's first index.
__device__ __forceinline__ uint32_t wave_append_n(uint32_t n, uint32_t* counter) {
    uint32_t incl = n;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t v = __shfl_up(incl, d, kWave);
        if (lane_id() >= d) incl += v;
    }
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    uint32_t base = 0;
    if (lane_id() == kWave - 1 && total) base = atomicAdd(counter, total);
    base = __shfl(base, kWave - 1, kWave);
    return base + incl - n;
}

// Persistent work fetch: one 64-item chunk per wave.
__device__ __forceinline__ int wave_fetch(uint32_t* head) {
    uint32_t b = 0;
    if (lane_id() == 0) b = atomicAdd(head, (uint32_t)kWave);
    return (int)__shfl(b, 0, kWave);
}

// Chunk fetch for the persistent loops: one 64-item chunk per wave from a
// global head.  (Splitting the queue into one contiguous segment per XCD, for
// L2 locality, measured slower: the segments' costs differ.)
__device__ __forceinline__ int seg_fetch(uint32_t* heads, int n, int& t, int& hi) {
    if (t) return -1;
    hi = n;
    const int b = wave_fetch(heads);
    if (b < n) return b;
    t = 1;
    return -1;
}

__device__ __forceinline__ void wave_add_u64(uint32_t v, unsigned long long* dst) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    if (lane_id() == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// ---- per-slot program pieces ----------------------------------------------

struct Ctx {
    const DevScene* scenes;
    const ItemRec* items;
    int32_t n_items;
    Slots sl;
    Queues q;
};

__device__ __forceinline__ int find_item(const ItemRec* items, int n, int64_t p) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (items[mid].pix_begin <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Camera ray of the slot's current sample (render_tile_region :361-367).
__device__ void start_sample(const Ctx& C, int s, uint32_t& rays) {
    const Slots& S = C.sl;
    const ItemRec& it = C.items[S.item[s]];
    const RSettings& st = it.st;
    const int W = st.width, H = st.height, n = st.samples;
    const int pix = S.pixel[s];
    const int i = pix % W, j = pix / W;
    const int smp = S.sample[s];
    const int sy = smp / n, sx = smp % n;
    Rng rng(S.rng[s]);
    float u = (i + (sx + 0.5f) / n) / (float)W;
    float v = 1.0f - (j + (sy + 0.5f) / n) / (float)H;
    u += rng.sym() * (0.5f / (W * n));
    v += rng.sym() * (0.5f / (H * n));
    Ray r = camera_ray(it.cam, u, v, rng);
    S.rng[s] = rng.s;
    stf(S.ray_o, s, r.o);
    stf(S.ray_d, s, r.d);
    S.nf[s] = 0;
    if (st.pathtrace) {
        rays++;                      // one per trace_path() call
        S.mode[s] = M_PATH;
        S.depth[s] = st.bounces;
        stf(S.T, s, f3(1, 1, 1));
        stf(S.R, s, f3(0, 0, 0));
    } else {
        S.mode[s] = M_WHITTED;
        S.depth[s] = st.max_depth;
    }
    S.glass_root[s] = 0;
}

// Slot program results.
enum : int { ADV_IDLE = 0, ADV_QUERY = 1, ADV_NEED_PIXEL = 2 };

// Gives slot s job pixel p (render_tile_region's per-pixel setup, :313).
// Job pixels are numbered in 8x8 tiles (row-major tiles, row-major inside a
// tile) so that a wave's 64 consecutive slots start on one coherent tile, as
// in the per-pixel megakernel.  Items are padded to whole tiles; a padding
// pixel returns false.
__device__ bool assign_pixel(const Ctx& C, int s, int64_t p) {
    const Slots& S = C.sl;
    const int it = find_item(C.items, C.n_items, p);
    const int r = (int)(p - C.items[it].pix_begin);
    const int W = C.items[it].st.width, H = C.items[it].st.height;
    const int tiles_x = (W + 7) >> 3;
    const int tile = r >> 6, in = r & 63;
    const int i = (tile % tiles_x) * 8 + (in & 7), j = (tile / tiles_x) * 8 + (in >> 3);
    if (i >= W || j >= H) return false;
    const int pix = j * W + i;
    S.item[s] = it;
    S.pixel[s] = pix;
    S.rng[s] = (uint32_t)(i * 97861u + j * 1973u + 1u);
    S.sample[s] = 0;
    stf(S.acc, s, f3(0, 0, 0));
    return true;
}

// acc += sample color; starts the next sample (ADV_QUERY pending) or writes
// the pixel (ADV_NEED_PIXEL).
__device__ int finish_sample(const Ctx& C, int s, const f3& color, uint32_t& rays) {
    const Slots& S = C.sl;
    f3 acc = ldf(S.acc, s);
    acc += color;
    const ItemRec& it = C.items[S.item[s]];
    const int n = it.st.samples;
    int smp = S.sample[s] + 1;
    if (smp < n * n) {
        stf(S.acc, s, acc);
        S.sample[s] = smp;
        start_sample(C, s, rays);
        return ADV_QUERY;
    }
    const f3 px = acc / (float)(n * n);
    float* o = it.out + (size_t)S.pixel[s] * 3;
    o[0] = px.x; o[1] = px.y; o[2] = px.z;
    return ADV_NEED_PIXEL;
}

// Folds a Whitted result into the frames (the recursion's return path).
// ADV_QUERY: a refraction child (ray/depth set) or a new sample is pending;
// ADV_NEED_PIXEL: the pixel is done.
__device__ int unwind(const Ctx& C, int s, f3 result, uint32_t& rays) {
    const Slots& S = C.sl;
    Frame* F = S.frames + (size_t)s * kMaxFrames;
    int nf = S.nf[s];
    while (nf > 0) {
        Frame& f = F[nf - 1];
        if (f.type == FR_METAL) {
            result = f.acc * (1.0f - f.w0) + result * f.albedo * f.w0;
            --nf;
        } else if (f.type == FR_GLASS_TIR) {
            result = result * f.albedo;
            --nf;
        } else if (f.type == FR_GLASS_REFL) {
            f.acc = result;
            f.type = FR_GLASS_REFR;
            S.nf[s] = nf;
            stf(S.ray_o, s, f.ro);
            stf(S.ray_d, s, f.rd);
            S.depth[s] = f.depth - 1;
            return ADV_QUERY;
        } else {
            f3 g = f.acc * f.w0 + result * f.w1;
            result = g * f.albedo;
            --nf;
        }
    }
    S.nf[s] = 0;
    if (S.glass_root[s]) {       // trace_path's glass: R += T*glass - T*emission
        const f3 T = ldf(S.gT, s), e = ldf(S.gE, s);
        f3 R = ldf(S.R, s);
        R += T * result - T * e;
        result = R;
    }
    return finish_sample(C, s, result, rays);
}

// Drives the slot until it needs a closest-hit query (ADV_QUERY, ray set) or
// a new pixel (ADV_NEED_PIXEL).  Whitted nodes with depth <= 0 return black
// at once.  status: ADV_QUERY = the slot's ray is a fresh query; 3 = fold
// `result` into the frames.
__device__ int advance(const Ctx& C, int s, int status, f3 result, uint32_t& rays) {
    const Slots& S = C.sl;
    for (int guard = 0; guard < 1 << 20; ++guard) {
        if (status == ADV_NEED_PIXEL || status == ADV_IDLE) return status;
        if (status == 3) { status = unwind(C, s, result, rays); continue; }
        if (S.mode[s] == M_WHITTED) {
            rays++;                                  // trace() entry (:404)
            if (S.depth[s] <= 0) { status = 3; result = f3(0, 0, 0); continue; }
            return ADV_QUERY;
        }
        if (S.depth[s] > 0) return ADV_QUERY;        // path: while (depth > 0)
        status = finish_sample(C, s, ldf(S.R, s), rays);
    }
    return ADV_IDLE;
}

// Hands new pixels to every lane whose slot finished one: one atomic per wave
// for the whole wave's requests.  Loops because a new pixel can finish
// without a query (zero depth / zero bounces).
__device__ int refill(const Ctx& C, int s, int a, uint32_t& rays) {
    while (true) {
        const unsigned long long need = __ballot(a == ADV_NEED_PIXEL);
        if (!need) return a;
        const int leader = __ffsll((long long)need) - 1;
        uint32_t base = 0;
        if (lane_id() == leader) base = atomicAdd(&C.q.ctr[Q_PIXEL], (uint32_t)__popcll(need));
        base = __shfl(base, leader, kWave);
        if (a == ADV_NEED_PIXEL) {
            const int64_t p = (int64_t)base + __popcll(need & ((1ull << lane_id()) - 1ull));
            if (p < C.q.total_pixels) {
                if (assign_pixel(C, s, p)) {
                    start_sample(C, s, rays);
                    a = advance(C, s, ADV_QUERY, f3(0, 0, 0), rays);
                }   // else: a padding pixel, ask again
            } else {
                C.sl.item[s] = -1;
                a = ADV_IDLE;
            }
        }
    }
}

template <int K>
__global__ __launch_bounds__(kWfBlock) void k_closest(Ctx C) {
    __shared__ unsigned long long lds[K * kWfBlock];
    LaneStack<K, kStackK + kStackSpill - K> st;
    st.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
    st.stride = kWfBlock;
    const int cur = C.q.ctr[Q_CUR];
    const int n = (int)C.q.ctr[Q_COUNT0 + cur];
    const int32_t* Q = C.q.ray_q[cur];
    const Slots& S = C.sl;
    TravCount<false> cnt;
    int seg_t = 0, seg_hi = 0;
    while (true) {
        const int base = seg_fetch(&C.q.ctr[Q_HEAD_CLOSEST], n, seg_t, seg_hi);
        if (base < 0) break;
        const int k = base + lane_id();
        if (k < seg_hi) {
            const int s = Q[k];
            const DevScene& D = C.scenes[C.items[S.item[s]].scene];
            Ray r;
            r.o = ldf(S.ray_o, s);
            r.d = ldf(S.ray_d, s);
            float tb, ta;
            const int slot = traverse<false>(D, r, 1e-4f, 1e9f, st, tb, ta, cnt);
            S.hit[s] = slot;
            S.tacc[s] = ta;
        }
    }
}

template <int K>
__global__ __launch_bounds__(kWfBlock) void k_anyhit(Ctx C) {
    __shared__ unsigned long long lds[K * kWfBlock];
    LaneStack<K, kStackK + kStackSpill - K> st;
    st.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
    st.stride = kWfBlock;
    const int n = (int)C.q.ctr[Q_SHADOW];
    const Slots& S = C.sl;
    TravCount<false> cnt;
    int seg_t = 0, seg_hi = 0;
    while (true) {
        const int base = seg_fetch(&C.q.ctr[Q_HEAD_ANY], n, seg_t, seg_hi);
        if (base < 0) break;
        const int k = base + lane_id();
        if (k < seg_hi) {
            const int2 who = C.q.sh_who[k];
            const float4 a = C.q.sh_ray[2 * k], b = C.q.sh_ray[2 * k + 1];
            const DevScene& D = C.scenes[C.items[S.item[who.x]].scene];
            Ray r;
            r.o = f3(a.x, a.y, a.z);
            r.d = f3(b.x, b.y, b.z);
            float tb, ta;
            const bool occ = traverse<true>(D, r, a.w, b.w, st, tb, ta, cnt) >= 0;
            S.occl[(size_t)who.x * kMaxSh + who.y] = occ ? 1 : 0;
        }
    }
}

// Shading of one vertex (see the file comment).  Light samples are stored in
// the slot's sample array in the reference's loop order; the shadow rays go
// straight into the shadow queue at an offset allocated per wave up front
// (the number of tested samples depends only on the lights and the settings).
__global__ __launch_bounds__(kWfBlock, 4) void k_shade(Ctx C) {
    const int cur = C.q.ctr[Q_CUR];
    const int n = (int)C.q.ctr[Q_COUNT0 + cur];
    const int32_t* Q = C.q.ray_q[cur];
    const Slots& S = C.sl;
    int seg_t = 0, seg_hi = 0;
    while (true) {
        const int base = seg_fetch(&C.q.ctr[Q_HEAD_SHADE], n, seg_t, seg_hi);
        if (base < 0) break;
        const int k = base + lane_id();
        const int s = (k < seg_hi) ? Q[k] : -1;
        const ItemRec* itp = (s >= 0) ? &C.items[S.item[s]] : nullptr;
        const DevScene* Dp = itp ? &C.scenes[itp->scene] : nullptr;
        const bool path = (s >= 0) && S.mode[s] == M_PATH;
        const int slot = (s >= 0) ? S.hit[s] : -1;
        uint32_t n_test = 0;
        if (slot >= 0) {
            for (int li = 0; li < Dp->n_lights; ++li) {
                const LightRec L = load_light(Dp->lights, li);
                if (!L.emissive) n_test += path ? 1 : ((L.radius > 0.0f) ? itp->st.soft_samples : 1);
            }
        }
        uint32_t o = wave_append_n(n_test, &C.q.ctr[Q_SHADOW]);
        if (s < 0) continue;
        const ItemRec& it = *itp;
        const DevScene& D = *Dp;
        const RSettings& st = it.st;
        Ray r;
        r.o = ldf(S.ray_o, s);
        r.d = ldf(S.ray_d, s);
        int outcome;
        S.nsh[s] = 0;
        if (slot < 0) {
            outcome = path ? V_P_MISS : V_W_MISS;
        } else {
            HitRec rec;
            PrimView P = load_prim(D.prims, slot);
            float t;
            hit_prim<true>(P, r, 1e-4f, S.tacc[s], t, &rec);
            int kind; f3 albedo, emission, spec_color; float refl, ior, spec, shin;
            load_mat(D.mats, rec.mat, kind, albedo, emission, refl, ior, spec, shin, spec_color);
            Rng rng(S.rng[s]);
            stf(S.alb, s, albedo);
            float4* SH = S.sh + (size_t)s * kMaxSh;
            uint8_t* OC = S.occl + (size_t)s * kMaxSh;
            int ns = 0;
            const f3 T = path ? ldf(S.T, s) : f3(1, 1, 1);
            const f3 sn = rec.normal;
            const f3 sh_o = rec.p + rec.normal * st.shadow_bias;
            for (int li = 0; li < D.n_lights; ++li) {
                const LightRec L = load_light(D.lights, li);
                const f3 lpos = ld3(L.pos), I = ld3(L.I);
                const int nsamp = path ? 1 : ((L.radius > 0.0f) ? st.soft_samples : 1);
                for (int q = 0; q < nsamp; ++q) {
                    f3 L_pos = lpos;
                    f3 Ld;
                    float dist = 1.0f;
                    if (L.kind == PCL_LIGHT_POINT) {
                        if (L.radius > 0.0f) L_pos = jitter_light(lpos, L.radius, rec.p, rng);
                        Ld = L_pos - rec.p; dist = Ld.length(); Ld = Ld / dist;
                    } else {
                        Ld = -lpos; dist = 1e9f;
                    }
                    const float ndotl = clampf(sn.dot(Ld), 0.0f, 1.0f);
                    f3 c;
                    if (path) {
                        c = T * albedo * I * ndotl;                 // :914-915
                    } else {
                        c = albedo * I * ndotl;                     // :558-566
                        if (spec > 0.0f && ndotl > 0.0f) {
                            f3 V = -r.d.normalized();
                            f3 H = (Ld + V).normalized();
                            float ndoth = clampf(sn.dot(H), 0.0f, 1.0f);
                            float sp = powf_cr(ndoth, shin);
                            c += spec_color * I * (sp * spec);
                        }
                    }
                    SH[ns] = make_float4(c.x, c.y, c.z, __int_as_float(li | (nsamp << 8)));
                    OC[ns] = 0;
                    if (!L.emissive) {
                        C.q.sh_who[o] = make_int2(s, ns);
                        C.q.sh_ray[2 * o] = make_float4(sh_o.x, sh_o.y, sh_o.z, st.shadow_bias);
                        C.q.sh_ray[2 * o + 1] = make_float4(Ld.x, Ld.y, Ld.z, dist - st.shadow_bias);
                        ++o;
                    }
                    ns++;
                }
            }
            S.nsh[s] = ns;
            if (path) {
                // radiance += T * emission comes before the NEE terms (:861)
                f3 R = ldf(S.R, s);
                R += T * emission;
                stf(S.R, s, R);
                const int depth = S.depth[s];
                outcome = V_P_BOUNCE;
                f3 Tn = T;
                if (st.bounces - depth > 3) {   // Russian roulette, :922-926
                    float p = fmaxf(Tn.x, fmaxf(Tn.y, Tn.z));
                    if (rng.unit() > p) outcome = V_P_END;
                    else Tn = Tn / p;
                }
                if (outcome == V_P_BOUNCE) {
                    f3 nd;
                    f3 NT = Tn * albedo;
                    if (kind == PCL_MAT_METAL && refl > 0.0f) {
                        f3 rf = r.d.normalized().reflect(sn).normalized();
                        f3 up = (fabsf(rf.y) < 0.9f) ? f3(0, 1, 0) : f3(1, 0, 0);
                        f3 t1 = rf.cross(up).normalized();
                        f3 t2 = rf.cross(t1).normalized();
                        float r1 = rng.unit() * 0.1f;
                        float r2 = rng.unit() * 6.28f;
                        nd = (rf + t1 * (r1 * cosf_cr(r2)) + t2 * (r1 * sinf_cr(r2))).normalized();
                        NT = Tn * albedo * refl;
                    } else if (kind == PCL_MAT_DIELECTRIC) {
                        Frame* F = S.frames + (size_t)s * kMaxFrames;
                        int nf = 0;
                        Ray rr = push_glass(F, nf, r, rec, albedo, ior, depth, st.shadow_bias);
                        S.nf[s] = nf;
                        stf(S.gT, s, Tn);
                        stf(S.gE, s, emission);
                        stf(S.ray_o, s, rr.o);
                        stf(S.ray_d, s, rr.d);
                        outcome = V_P_GLASS;
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
                    if (outcome == V_P_BOUNCE) {
                        stf(S.ray_o, s, rec.p + rec.normal * st.shadow_bias);
                        stf(S.ray_d, s, nd);
                        stf(S.NT, s, NT);
                    }
                }
            } else {
                stf(S.pend, s, emission);
                outcome = V_W_LEAF;
                if (kind == PCL_MAT_METAL && refl > 0.0f) {
                    Frame* F = S.frames + (size_t)s * kMaxFrames;
                    const int nf = S.nf[s];
                    Frame& f = F[nf];
                    f.type = FR_METAL; f.albedo = albedo; f.w0 = refl; f.depth = S.depth[s];
                    S.nf[s] = nf + 1;
                    stf(S.ray_o, s, rec.p + rec.normal * st.shadow_bias);
                    stf(S.ray_d, s, r.d.normalized().reflect(rec.normal).normalized());
                    outcome = V_W_CHILD;
                } else if (kind == PCL_MAT_DIELECTRIC) {
                    Frame* F = S.frames + (size_t)s * kMaxFrames;
                    int nf = S.nf[s];
                    Ray rr = push_glass(F, nf, r, rec, albedo, ior, S.depth[s], st.shadow_bias);
                    S.nf[s] = nf;
                    stf(S.ray_o, s, rr.o);
                    stf(S.ray_d, s, rr.d);
                    outcome = V_W_CHILD;
                    S.outcome_aux[s] = 1;   // glass: the local color is discarded
                }
            }
            S.rng[s] = rng.s;
        }
        S.outcome[s] = outcome;
    }
}

// Folds one vertex's light samples exactly as the reference's loops do.
__device__ f3 fold_whitted_local(const Slots& S, int s, const f3& emission, const f3& albedo, const f3& amb) {
    const float4* SH = S.sh + (size_t)s * kMaxSh;
    const uint8_t* OC = S.occl + (size_t)s * kMaxSh;
    const int ns = S.nsh[s];
    f3 local = emission;
    int q = 0;
    while (q < ns) {
        const int meta = __float_as_int(SH[q].w);
        const int li = meta & 0xFF, nsamp = (meta >> 8) & 0xFF;
        f3 contrib(0, 0, 0);
        for (int k = 0; k < nsamp; ++k, ++q) {
            if (OC[q]) continue;
            contrib += f3(SH[q].x, SH[q].y, SH[q].z);
        }
        local += contrib / (float)nsamp;
        (void)li;
    }
    local += albedo * amb;
    return local;
}

__global__ __launch_bounds__(kWfBlock, 4) void k_resolve(Ctx C) {
    const int cur = C.q.ctr[Q_CUR];
    const int n = (int)C.q.ctr[Q_COUNT0 + cur];
    const int32_t* Q = C.q.ray_q[cur];
    const Slots& S = C.sl;
    uint32_t rays = 0;
    int seg_t = 0, seg_hi = 0;
    while (true) {
        const int base = seg_fetch(&C.q.ctr[Q_HEAD_RESOLVE], n, seg_t, seg_hi);
        if (base < 0) break;
        const int k = base + lane_id();
        int a = ADV_IDLE;
        int s = -1;
        if (k < seg_hi) {
            s = Q[k];
            const ItemRec& it = C.items[S.item[s]];
            const DevScene& D = C.scenes[it.scene];
            const f3 amb = ld3(D.amb);
            const int outcome = S.outcome[s];
            int status;
            f3 result(0, 0, 0);
            switch (outcome) {
            case V_W_MISS: status = 3; result = ld3(D.bg); break;
            case V_W_LEAF: status = 3; result = fold_whitted_local(S, s, ldf(S.pend, s), ldf(S.alb, s), amb); break;
            case V_W_CHILD: {
                if (S.outcome_aux[s]) {
                    S.outcome_aux[s] = 0;          // glass: local discarded (its draws already made)
                } else {
                    Frame* F = S.frames + (size_t)s * kMaxFrames;
                    F[S.nf[s] - 1].acc = fold_whitted_local(S, s, ldf(S.pend, s), ldf(S.alb, s), amb);
                }
                S.depth[s] = S.depth[s] - 1;
                status = ADV_QUERY;
                break;
            }
            case V_P_MISS: {
                f3 R = ldf(S.R, s);
                R += ldf(S.T, s) * ld3(D.bg);
                stf(S.R, s, R);
                status = finish_sample(C, s, R, rays);
                break;
            }
            default: {   // path vertex with a hit: NEE terms in light order, then ambient
                f3 R = ldf(S.R, s);
                const float4* SH = S.sh + (size_t)s * kMaxSh;
                const uint8_t* OC = S.occl + (size_t)s * kMaxSh;
                const int ns = S.nsh[s];
                for (int q = 0; q < ns; ++q)
                    if (!OC[q]) R += f3(SH[q].x, SH[q].y, SH[q].z);
                R += ldf(S.T, s) * ldf(S.alb, s) * amb;
                stf(S.R, s, R);
                if (outcome == V_P_END) {
                    status = finish_sample(C, s, R, rays);
                } else if (outcome == V_P_GLASS) {
                    S.mode[s] = M_WHITTED;
                    S.glass_root[s] = 1;
                    S.depth[s] = S.depth[s] - 1;   // trace(refl, depth - 1)
                    status = ADV_QUERY;
                } else {
                    stf(S.T, s, ldf(S.NT, s));
                    S.depth[s] = S.depth[s] - 1;
                    status = ADV_QUERY;
                }
                break;
            }
            }
            a = advance(C, s, status, result, rays);
        }
        a = refill(C, s, a, rays);
        const int pos = wave_append(a == ADV_QUERY, &C.q.ctr[Q_COUNT0 + (cur ^ 1)]);
        if (pos >= 0) C.q.ray_q[cur ^ 1][pos] = s;
    }
    wave_add_u64(rays, &C.q.stats[ST_RAYS]);
}

// Starts the job: every slot takes a pixel and its first query.
__global__ __launch_bounds__(kWfBlock, 4) void k_start(Ctx C, int n_slots) {
    uint32_t rays = 0;
    const int cur = C.q.ctr[Q_CUR];
    for (int base = blockIdx.x * kWfBlock + (threadIdx.x & ~(kWave - 1)); base < n_slots;
         base += gridDim.x * kWfBlock) {
        const int s = base + lane_id();
        int a = ADV_IDLE;
        if (s < n_slots) {
            C.sl.nf[s] = 0;
            C.sl.glass_root[s] = 0;
            C.sl.outcome_aux[s] = 0;
            if ((int64_t)s < C.q.total_pixels) {   // slot s starts on job pixel s
                if (assign_pixel(C, s, (int64_t)s)) {
                    start_sample(C, s, rays);
                    a = advance(C, s, ADV_QUERY, f3(0, 0, 0), rays);
                } else {
                    a = ADV_NEED_PIXEL;
                }
            } else {
                C.sl.item[s] = -1;
            }
        }
        a = refill(C, s, a, rays);
        const int pos = wave_append(a == ADV_QUERY, &C.q.ctr[Q_COUNT0 + cur]);
        if (pos >= 0) C.q.ray_q[cur][pos] = s;
    }
    wave_add_u64(rays, &C.q.stats[ST_RAYS]);
}

// End of an iteration: account the queries, flip the ray queue, clear heads.
__global__ void k_swap(Queues q) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t cur = q.ctr[Q_CUR];
    q.stats[ST_CLOSEST] += q.ctr[Q_COUNT0 + cur];
    q.stats[ST_ANY] += q.ctr[Q_SHADOW];
    q.stats[ST_ITERS] += q.ctr[Q_COUNT0 + cur] ? 1 : 0;
    q.ctr[Q_COUNT0 + cur] = 0;
    q.ctr[Q_CUR] = cur ^ 1;
    q.ctr[Q_SHADOW] = 0;
    for (int i = Q_HEAD_CLOSEST; i < Q_NCTR; ++i) q.ctr[i] = 0;
}

// Clears the job counters (a kernel, not a memset, so it may sit in a graph).
__global__ void k_reset(Queues q, uint32_t first_free_pixel) {
    if (threadIdx.x < Q_NCTR) q.ctr[threadIdx.x] = threadIdx.x == Q_PIXEL ? first_free_pixel : 0u;
    if (threadIdx.x < ST_N) q.stats[threadIdx.x] = 0;
}

}  // namespace wf
}  // namespace pcl

// ---------------------------------------------------------------------------
// Host engine

namespace pcl {
namespace wf {

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

#define WF_TRY(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) { err = std::string(#call) + ": " + hipGetErrorString(e_); \
                                return e_ == hipErrorOutOfMemory ? PCL_ERR_OOM : PCL_ERR_HIP; } \
    } while (0)

Engine::~Engine() { release(); }

void Engine::release() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
    for (void* p : bufs) (void)hipFree(p);
    bufs.clear();
    cap_slots = 0;
    cap_items = 0;
    cap_shadow = 0;
}

int Engine::alloc(size_t bytes, void** p, std::string& err) {
    WF_TRY(hipMalloc(p, bytes < 256 ? 256 : bytes));
    bufs.push_back(*p);
    return PCL_OK;
}

int Engine::ensure(int slots, int items, std::string& err) {
    if (slots <= cap_slots && items <= cap_items) return PCL_OK;
    release();
    const size_t S = (size_t)slots;
    int rc = PCL_OK;
    void* p = nullptr;
#define A(field, type, count)                                        \
    if ((rc = alloc(sizeof(type) * (count), &p, err)) != PCL_OK) return rc; \
    field = static_cast<type*>(p);
    A(sl.item, int32_t, S) A(sl.pixel, int32_t, S) A(sl.rng, uint32_t, S) A(sl.sample, int32_t, S)
    A(sl.acc, float, 3 * S) A(sl.mode, int32_t, S) A(sl.depth, int32_t, S) A(sl.nf, int32_t, S)
    A(sl.ray_o, float, 3 * S) A(sl.ray_d, float, 3 * S) A(sl.hit, int32_t, S) A(sl.tacc, float, S)
    A(sl.T, float, 3 * S) A(sl.NT, float, 3 * S) A(sl.R, float, 3 * S) A(sl.pend, float, 3 * S)
    A(sl.alb, float, 3 * S) A(sl.gT, float, 3 * S) A(sl.gE, float, 3 * S) A(sl.outcome, int32_t, S)
    A(sl.outcome_aux, int32_t, S) A(sl.glass_root, int32_t, S) A(sl.nsh, int32_t, S)
    A(sl.sh, float4, S * kMaxSh) A(sl.occl, uint8_t, S * kMaxSh)
    A(sl.frames, Frame, S * kMaxFrames)
    A(q.ray_q[0], int32_t, S) A(q.ray_q[1], int32_t, S)
    A(q.sh_who, int2, S * kMaxSh) A(q.sh_ray, float4, 2 * S * kMaxSh)
    A(q.ctr, uint32_t, 64) A(q.stats, unsigned long long, 16)
    A(d_items, ItemRec, items) A(d_scenes, DevScene, items)
#undef A
    cap_slots = slots;
    cap_items = items;
    cap_shadow = slots * kMaxSh;
    return PCL_OK;
}

template <int K>
static void launch_iter(const Ctx& c, int grid_trav, int grid_other, hipStream_t st) {
    hipLaunchKernelGGL(k_closest<K>, dim3(grid_trav), dim3(kWfBlock), 0, st, c);
    hipLaunchKernelGGL(k_shade, dim3(grid_other), dim3(kWfBlock), 0, st, c);
    hipLaunchKernelGGL(k_anyhit<K>, dim3(grid_trav), dim3(kWfBlock), 0, st, c);
    hipLaunchKernelGGL(k_resolve, dim3(grid_other), dim3(kWfBlock), 0, st, c);
    hipLaunchKernelGGL(k_swap, dim3(1), dim3(64), 0, st, c.q);
}

int Engine::run(hipStream_t stream, const std::vector<ItemRec>& items, const std::vector<DevScene>& scenes,
                const Config& cfg_in, int cus, RunStats& out, std::string& err) {
    Config cfg = cfg_in;
    int64_t total = 0;
    for (const auto& it : items) total += padded_pixels(it.st.width, it.st.height);
    if (cfg.slots <= 0) cfg.slots = (int)std::min<int64_t>(total, 2 << 20);
    cfg.slots = (int)std::max<int64_t>(1, std::min<int64_t>(cfg.slots, total));
    if (cfg.blocks_per_cu <= 0) cfg.blocks_per_cu = 8;
    if (cfg.stack_k != 8 && cfg.stack_k != 16 && cfg.stack_k != 32) cfg.stack_k = 16;
    if (cfg.iters_per_graph <= 0) cfg.iters_per_graph = 4;
    if (total <= 0) { out = RunStats{}; return PCL_OK; }
    if (total >= (int64_t)1 << 31) { err = "job larger than 2^31 pixels"; return PCL_ERR_LIMIT; }
    int rc = ensure(cfg.slots, (int)items.size(), err);
    if (rc) return rc;
    std::vector<ItemRec> its = items;
    for (size_t i = 0; i < its.size(); ++i) its[i].scene = (int32_t)i;   // scene table is per item
    WF_TRY(hipMemcpyAsync(d_items, its.data(), its.size() * sizeof(ItemRec), hipMemcpyHostToDevice, stream));
    WF_TRY(hipMemcpyAsync(d_scenes, scenes.data(), scenes.size() * sizeof(DevScene), hipMemcpyHostToDevice, stream));
    Ctx c;
    c.scenes = d_scenes;
    c.items = d_items;
    c.n_items = (int32_t)its.size();
    c.sl = sl;
    c.q = q;
    c.q.total_pixels = total;
    const int grid_trav = cus * cfg.blocks_per_cu;
    const int grid_other = cus * 4;
    const bool graph_ok = (graph_exec && graph_key.slots == cfg.slots && graph_key.blocks_per_cu == cfg.blocks_per_cu &&
                           graph_key.stack_k == cfg.stack_k && graph_key.iters_per_graph == cfg.iters_per_graph &&
                           graph_items == (int)its.size() && graph_total == total);
    if (!graph_ok) {
        if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
        hipGraph_t g = nullptr;
        WF_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < cfg.iters_per_graph; ++i) {
            if (cfg.stack_k == 8) launch_iter<8>(c, grid_trav, grid_other, stream);
            else if (cfg.stack_k == 32) launch_iter<32>(c, grid_trav, grid_other, stream);
            else launch_iter<16>(c, grid_trav, grid_other, stream);
        }
        WF_TRY(hipStreamEndCapture(stream, &g));
        hipError_t e = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { graph_exec = nullptr; err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e); return PCL_ERR_HIP; }
        graph_key = cfg;
        graph_items = (int)its.size();
        graph_total = total;
    }
    hipEvent_t e0, e1;
    WF_TRY(hipEventCreate(&e0));
    WF_TRY(hipEventCreate(&e1));
    WF_TRY(hipEventRecord(e0, stream));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, stream, c.q, (uint32_t)std::min<int64_t>(total, cfg.slots));
    hipLaunchKernelGGL(k_start, dim3((cfg.slots + kWfBlock - 1) / kWfBlock), dim3(kWfBlock), 0, stream, c, cfg.slots);
    WF_TRY(hipGetLastError());
    uint32_t h_ctr[Q_NCTR];
    int launches = 0;
    const double t_begin = now_ms();
    while (true) {
        WF_TRY(hipGraphLaunch(graph_exec, stream));
        ++launches;
        WF_TRY(hipMemcpyAsync(h_ctr, q.ctr, sizeof(h_ctr), hipMemcpyDeviceToHost, stream));
        WF_TRY(hipStreamSynchronize(stream));
        if (h_ctr[Q_COUNT0 + h_ctr[Q_CUR]] == 0) break;
        // every iteration retires at least one query of a finite program; a
        // job still running after this long is a bug, reported, not waited on
        if (now_ms() - t_begin > 300000.0) { err = "wavefront did not drain in 300 s"; return PCL_ERR_LIMIT; }
    }
    WF_TRY(hipEventRecord(e1, stream));
    unsigned long long h_st[ST_N];
    WF_TRY(hipMemcpyAsync(h_st, q.stats, sizeof(h_st), hipMemcpyDeviceToHost, stream));
    WF_TRY(hipStreamSynchronize(stream));
    float ms = 0.0f;
    WF_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out.rays = h_st[ST_RAYS];
    out.closest = h_st[ST_CLOSEST];
    out.anyhit = h_st[ST_ANY];
    out.iterations = h_st[ST_ITERS];
    out.graph_launches = launches;
    out.ms = ms;
    out.cfg = cfg;
    return PCL_OK;
}

}  // namespace wf
}  // namespace pcl
