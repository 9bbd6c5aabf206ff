// wavefront.hip - batched, persistent wavefront tracer.
//
// Renders a job of many (scene, camera, settings) items in one pass over a
// fixed pool of pixel slots.  Each slot owns one pixel at a time and runs the
// reference's per-pixel program (src/renderer.hpp:309-390: n*n samples in
// order on the pixel's xorshift stream; Whitted trace() trees or
// trace_path() chains) as a state machine that stops at every BVH query.
// One iteration of the machine is four kernels:
//
//   closest+shade : persistent; pops the ray queue, Scene::hit, then shades
//              the hit: hit record, material, every light sample's
//              contribution and shadow ray, bounce / child ray.  Every RNG
//              draw of the vertex happens here, in the reference's order
//              (light jitter, Russian roulette, bounce).  A vertex with few
//              shadow rays traces them here too; the others go to the
//              shadow queue                             [LDS ring stack]
//   anyhit   : persistent; pops the shadow queue, Scene::occluded -> one
//              flag per sample                          [LDS ring stack]
//   resolve  : folds the unoccluded contributions in the reference's order,
//              unwinds Whitted frames, finishes samples and pixels, pulls the
//              next pixel of the job, and pushes the slot's next query
//   swap     : flips the queues and clears the work heads
//
// Queues are compacted with wave ballots and prefix sums (one atomic per
// wave) and split into stripes with their own counters (see wavefront.hpp),
// so no counter takes more than a small share of the atomics.  Kernel grids
// are fixed (persistent waves fetch 64-item chunks, home stripe first), so an
// iteration is captured once as a hipGraph and replayed until the ray queue
// drains.  Near the end of a job the slots stop at sample boundaries and one
// per-pixel launch finishes the remaining samples (the drain, see
// Engine::run).  Slot count, persistent blocks per CU, the LDS stack depth,
// iterations per graph and the inline-shadow limit are the autotuned knobs.
//
// Floating-point results are the megakernel's (and so the reference's): every
// color is accumulated with the same operations in the same order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "glass_sort.hpp"
#include "render.hpp"
#include "wavefront.hpp"

namespace pcl {
namespace wf {

constexpr int kWave = 64;
constexpr int kWfBlock = 128;
constexpr int kDrainLanes = 4;   // automatic drain: this many times the traversal grids' lanes

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
    V_P_GLASS = 7,     // path glass: Whitted subtree from the pushed frame
    V_DONE = 8         // pixel finished and written by the closest-hit kernel (FOLD)
};

// ---- wave-level queue helpers --------------------------------------------

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ int wave_gid() { return (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); }
// A slot's stripe.
__device__ __forceinline__ int stripe_of(int s) { return (s >> 6) & (kStripes - 1); }

// A counter other waves are changing right now: read it at device scope.
__device__ __forceinline__ uint32_t ld_live(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

// Striped work fetch.  Stripe g holds len(g) items behind the head counter
// ctr[sidx(head, g)].  A wave takes `want` items from its home stripe while
// that lasts; afterwards it reads every stripe's head (one per lane) and takes
// from the first stripe with work in a wave-specific rotation.  A stripe
// whose atomic comes back past its end is dead for this wave, so the scan
// loop ends after at most kStripes misses whatever the reads return.  Returns
// false once every stripe is exhausted; otherwise items [base, base + got) of
// stripe g.
struct Fetcher {
    unsigned long long dead;   // stripes known to be exhausted
    int home;                  // home stripe, or -1 once it is exhausted
    int rot;                   // scan rotation
    PCL_D explicit Fetcher(int h) : dead(0), home(h), rot((h + 1 + (wave_gid() >> 6) * 7) & (kStripes - 1)) {}

    template <class Len>
    PCL_D bool next(uint32_t* ctr, int head, uint32_t want, Len len, int& g, uint32_t& base, uint32_t& got) {
        if (dead == ~0ull) return false;
        if (home >= 0) {
            const uint32_t L = len(home);
            uint32_t b = 0;
            if (lane_id() == 0) b = atomicAdd(&ctr[sidx(head, home)], want);
            b = __shfl(b, 0, kWave);
            if (b < L) { g = home; base = b; got = min(want, L - b); return true; }
            dead |= 1ull << home;
            home = -1;
        }
        while (true) {
            const int mine = lane_id();
            const bool has = ld_live(&ctr[sidx(head, mine)]) < len(mine);
            const unsigned long long m = __ballot(has) & ~dead;
            if (!m) { dead = ~0ull; return false; }   // heads only grow: for good
            const unsigned long long r = rot ? ((m >> rot) | (m << (kWave - rot))) : m;
            const int pick = (rot + __ffsll((long long)r) - 1) & (kStripes - 1);
            const uint32_t L = len(pick);
            uint32_t b = 0;
            if (lane_id() == 0) b = atomicAdd(&ctr[sidx(head, pick)], want);
            b = __shfl(b, 0, kWave);
            if (b < L) { g = pick; base = b; got = min(want, L - b); return true; }
            dead |= 1ull << pick;
        }
    }
};

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
    int32_t inline_max;   // shadow rays per vertex traced in the closest-hit kernel (Config::inline_shadows)
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

__device__ __forceinline__ float4 f4(const f3& v, float w = 0.0f) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ f3 xyz(const float4& v) { return f3(v.x, v.y, v.z); }

// The slot state a kernel works on, held in registers: loaded once, stored
// once (no read-after-write round trips through global memory).
// ray_set: o / d were given a new ray here (a new sample, a refraction
// child); otherwise the slot's ray is the one the shading kernel wrote, which
// the lane neither loads nor stores.
struct Lane {
    int item, pixel, sample;
    uint32_t rng;
    int mode, glass_root, outcome_aux, outcome, depth, nf, nsh;
    bool ray_set;
    f3 o, d, acc, T, R;
};

__device__ __forceinline__ int pack_flags(const Lane& L) {
    return L.mode | (L.glass_root << 4) | (L.outcome_aux << 5) | (L.outcome << 8);
}

__device__ __forceinline__ void load_lane(const Slots& S, int s, Lane& L) {
    const int4 m = S.meta[s], c = S.ctl[s];
    L.item = m.x; L.pixel = m.y; L.sample = m.z; L.rng = (uint32_t)m.w;
    L.mode = c.x & 15; L.glass_root = (c.x >> 4) & 1; L.outcome_aux = (c.x >> 5) & 1; L.outcome = c.x >> 8;
    L.depth = c.y; L.nf = c.z; L.nsh = c.w;
    L.ray_set = false;
    if (L.outcome == V_DONE) {   // the slot only wants a new pixel
        L.acc = L.T = L.R = f3(0, 0, 0);
        return;
    }
    L.acc = xyz(S.acc[s]); L.T = xyz(S.T[s]); L.R = xyz(S.R[s]);
}

__device__ __forceinline__ void store_lane(const Slots& S, int s, const Lane& L) {
    S.meta[s] = make_int4(L.item, L.pixel, L.sample, (int)L.rng);
    S.ctl[s] = make_int4(pack_flags(L), L.depth, L.nf, L.nsh);
    if (L.ray_set) {
        S.ray_o[s] = f4(L.o);
        S.ray_d[s] = f4(L.d);
    }
    S.acc[s] = f4(L.acc);
    // Whitted slots never read T, and R only below a path's glass vertex
    if (L.mode == M_PATH) S.T[s] = f4(L.T);
    if (L.mode == M_PATH || L.glass_root) S.R[s] = f4(L.R);
}

// Camera ray of the lane's current sample (render_tile_region :361-367).
// Inlined at each call site: as a call it cost the resolve kernel 5.5 % of
// the path-traced Cornell frame (every wave has lanes starting a sample).
PCL_D void start_sample(const Ctx& C, Lane& L, uint32_t& rays) {
    const ItemRec& it = C.items[L.item];
    const RSettings& st = it.st;
    const int W = st.width, H = st.height, n = st.samples;
    const int i = L.pixel % W, j = L.pixel / W;
    const int sy = L.sample / n, sx = L.sample % n;
    Rng rng(L.rng);
    float u = (i + (sx + 0.5f) / n) / (float)W;
    float v = 1.0f - (j + (sy + 0.5f) / n) / (float)H;
    u += rng.sym() * (0.5f / (W * n));
    v += rng.sym() * (0.5f / (H * n));
    Ray r = camera_ray(it.cam, u, v, rng);
    L.rng = rng.s;
    L.o = r.o;
    L.d = r.d;
    L.ray_set = true;
    L.nf = 0;
    L.glass_root = 0;
    if (st.pathtrace) {
        rays++;                      // one per trace_path() call
        L.mode = M_PATH;
        L.depth = st.bounces;
        L.T = f3(1, 1, 1);
        L.R = f3(0, 0, 0);
    } else {
        L.mode = M_WHITTED;
        L.depth = st.max_depth;
    }
}

// Slot program results.  ADV_HANDOFF: the slot stopped at a sample boundary
// in drain mode (state kept: next sample index, rng, acc) for the finish kernel.
enum : int { ADV_IDLE = 0, ADV_QUERY = 1, ADV_NEED_PIXEL = 2, ADV_HANDOFF = 4 };

// Job pixels are numbered in 64-pixel row segments (row-major) so that a
// wave's 64 consecutive slots start on one coherent run of pixels, as in the
// per-pixel megakernel.  Rows are padded to whole segments; a padding pixel
// returns false.  Sets up render_tile_region's per-pixel state (:313).
__device__ bool assign_pixel(const Ctx& C, Lane& L, int64_t p) {
    const int it = find_item(C.items, C.n_items, p);
    const int r = (int)(p - C.items[it].pix_begin);
    const int W = C.items[it].st.width, H = C.items[it].st.height;
    const int segs_x = (W + 63) >> 6;
    const int seg = r >> 6, in = r & 63;
    const int i = (seg % segs_x) * 64 + in, j = seg / segs_x;
    if (i >= W || j >= H) return false;
    L.item = it;
    L.pixel = j * W + i;
    L.rng = (uint32_t)(i * 97861u + j * 1973u + 1u);
    L.sample = 0;
    L.acc = f3(0, 0, 0);
    return true;
}

// acc += sample color; starts the next sample (ADV_QUERY pending) or writes
// the pixel (ADV_NEED_PIXEL).  In drain mode a pixel with samples left stops
// before the next sample's first draw (ADV_HANDOFF).
__device__ int finish_sample(const Ctx& C, Lane& L, const f3& color, uint32_t& rays, bool drain) {
    L.acc += color;
    const ItemRec& it = C.items[L.item];
    const int n = it.st.samples;
    if (L.sample + 1 < n * n) {
        L.sample += 1;
        if (drain) return ADV_HANDOFF;
        start_sample(C, L, rays);
        return ADV_QUERY;
    }
    const f3 px = L.acc / (float)(n * n);
    float* o = it.out + (size_t)L.pixel * 3;
    o[0] = px.x; o[1] = px.y; o[2] = px.z;
    return ADV_NEED_PIXEL;
}

// Folds a Whitted result into the frames (the recursion's return path).
// ADV_QUERY: a refraction child (ray/depth set) or a new sample is pending;
// ADV_NEED_PIXEL: the pixel is done.
__device__ int unwind(const Ctx& C, int s, Lane& L, f3 result, uint32_t& rays, bool drain) {
    Frame* F = C.sl.frames + (size_t)s * kMaxFrames;
    while (L.nf > 0) {
        Frame& f = F[L.nf - 1];
        if (f.type == FR_METAL) {
            result = f.acc * (1.0f - f.w0) + result * f.albedo * f.w0;
            --L.nf;
        } else if (f.type == FR_GLASS_TIR) {
            result = result * f.albedo;
            --L.nf;
        } else if (f.type == FR_GLASS_REFL) {
            f.acc = result;
            f.type = FR_GLASS_REFR;
            L.o = f.ro;
            L.d = f.rd;
            L.ray_set = true;
            L.depth = f.depth - 1;
            return ADV_QUERY;
        } else {
            f3 g = f.acc * f.w0 + result * f.w1;
            result = g * f.albedo;
            --L.nf;
        }
    }
    if (L.glass_root) {       // trace_path's glass: R += T*glass - T*emission
        const f3 T = xyz(C.sl.gT[s]), e = xyz(C.sl.gE[s]);
        L.R += T * result - T * e;
        result = L.R;
    }
    return finish_sample(C, L, result, rays, drain);
}

// Drives the lane until it needs a closest-hit query (ADV_QUERY, ray set) or
// a new pixel (ADV_NEED_PIXEL).  Whitted nodes with depth <= 0 return black
// at once.  status: ADV_QUERY = the lane's ray is a fresh query; 3 = fold
// `result` into the frames.
__device__ int advance(const Ctx& C, int s, Lane& L, int status, f3 result, uint32_t& rays, bool drain) {
    for (int guard = 0; guard < 1 << 20; ++guard) {
        if (status == ADV_NEED_PIXEL || status == ADV_IDLE || status == ADV_HANDOFF) return status;
        if (status == 3) { status = unwind(C, s, L, result, rays, drain); continue; }
        if (L.mode == M_WHITTED) {
            rays++;                                  // trace() entry (:404)
            if (L.depth <= 0) { status = 3; result = f3(0, 0, 0); continue; }
            return ADV_QUERY;
        }
        if (L.depth > 0) return ADV_QUERY;           // path: while (depth > 0)
        status = finish_sample(C, L, L.R, rays, drain);
    }
    return ADV_IDLE;
}

// Hands new pixels to every lane whose slot finished one: one fetch per wave
// for the whole wave's requests.  The part's 64-pixel segments ("tiles") are
// dealt to the stripes round-robin (local tile t to stripe t % kStripes), and
// the parts of a job take interleaved tiles, so while the stripes and the
// parts advance together the whole GPU works on one front of the image, whose
// BVH working set the caches share.  Loops because a new pixel can finish
// without a query (zero depth / zero bounces) or be a tile's padding.
__device__ int refill(const Ctx& C, int s, Lane& L, int a, Fetcher& pf, uint32_t& rays, bool drain, bool pool_left) {
    if (!pool_left) {   // every job pixel is taken: no fetch
        if (a == ADV_NEED_PIXEL) {
            L.item = -1;
            a = ADV_IDLE;
        }
        return a;
    }
    const uint32_t tiles = (uint32_t)(C.q.total_pixels >> 6);
    auto len = [&](int g) -> uint32_t {
        return (tiles / kStripes + ((uint32_t)g < tiles % kStripes ? 1u : 0u)) * 64u;
    };
    while (true) {
        const unsigned long long need = __ballot(a == ADV_NEED_PIXEL);
        if (!need) return a;
        int g = 0;
        uint32_t base = 0, got = 0;
        const bool ok = pf.next(C.q.ctr, S_PIXEL, (uint32_t)__popcll(need), len, g, base, got);
        if (a == ADV_NEED_PIXEL) {
            if (!ok) {
                L.item = -1;
                a = ADV_IDLE;
            } else {
                const uint32_t r = (uint32_t)__popcll(need & ((1ull << lane_id()) - 1ull));
                const uint32_t k = base + r;   // the stripe's k-th pixel
                const int64_t tile = ((int64_t)(k >> 6) * kStripes + g) * C.q.parts + C.q.part;
                if (r < got && assign_pixel(C, L, tile * 64 + (k & 63))) {
                    start_sample(C, L, rays);
                    a = advance(C, s, L, ADV_QUERY, f3(0, 0, 0), rays, drain);
                }
            }
        }
    }
}

template <int K, bool ONE>
__global__ __launch_bounds__(kWfBlock) void k_anyhit(Ctx C) {
    if (C.q.ctr[C_PENDING] == 0) return;   // empty iteration
    __shared__ unsigned long long lds[K * kWfBlock];
    LaneStack<K, kStackK + kStackSpill - K> st;
    st.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
    st.set_stride(kWfBlock);
    auto len = [&](int g) -> uint32_t { return C.q.ctr[sidx(S_SHADOW, g)]; };
    const Slots& S = C.sl;
    TravCount<false> cnt;
    Fetcher f(wave_gid() & (kStripes - 1));
    int g;
    uint32_t base, got;
    while (f.next(C.q.ctr, S_HEAD_ANY, kWave, len, g, base, got)) {
        if ((uint32_t)lane_id() < got) {
            const size_t k = (size_t)g * C.q.shcap + base + lane_id();
            const int2 who = C.q.sh_who[k];
            const float4 a = C.q.sh_ray[2 * k], b = C.q.sh_ray[2 * k + 1];
            const DevScene& D = ONE ? C.scenes[0] : C.scenes[C.items[S.meta[who.x].x].scene];
            Ray r;
            r.o = f3(a.x, a.y, a.z);
            r.d = f3(b.x, b.y, b.z);
            float tb, ta;
            const bool occ = traverse<true>(D, r, a.w, b.w, st, tb, ta, cnt) >= 0;
            S.occl[(size_t)who.y * S.sh_stride + who.x] = occ ? 1 : 0;
        }
    }
}

// Shading of one vertex (see the file comment).  Light samples are stored in
// the slot's sample array in the reference's loop order; the shadow rays go
// straight into the shadow queue at an offset allocated per wave up front
// (the number of tested samples depends only on the lights and the settings).
// One wave's chunk of vertices of stripe g (s = -1: lane without one).  Every
// lane of the wave calls it: the shadow-queue offsets are allocated per wave.
// ro/rd: the query ray with the closest-hit result in .w (t_max at
// acceptance, hit slot).  FOLD: a Whitted vertex that ends its pixel -- no
// child ray, no frame waiting, the pixel's last sample, every shadow ray
// traced here -- is folded here with resolve's operations in resolve's order
// (fold_whitted_local, finish_sample) and the pixel written; its light
// samples and material never go through the slot arrays, and k_resolve only
// hands the slot a new pixel (V_DONE).  FRESH: the slot's first query of the
// job (k_start_closest), whose sample sum is zero and whose meta record is
// not in memory yet.  Returns the vertex outcome (V_NONE without a vertex).
template <bool ONE, bool FOLD, bool FRESH, class ST>
__device__ __forceinline__ int shade_chunk(const Ctx& C, int g, int s, int4 meta, int4 ctl, float4 ro, float4 rd,
                                           ST& stk, uint32_t& n_inline) {
        const Slots& S = C.sl;
        const ItemRec* itp = ONE ? &C.items[0] : (s >= 0) ? &C.items[meta.x] : nullptr;
        const DevScene* Dp = ONE ? &C.scenes[0] : itp ? &C.scenes[itp->scene] : nullptr;
        const bool path = (ctl.x & 15) == M_PATH;
        const int slot = (s >= 0) ? __float_as_int(rd.w) : -1;
        uint32_t n_test = 0;
        if (slot >= 0) {
            for (int li = 0; li < Dp->n_lights; ++li) {
                const LightRec L = load_light(Dp->lights, li);
                if (!L.emissive) n_test += path ? 1 : ((L.radius > 0.0f) ? itp->st.soft_samples : 1);
            }
        }
        // a vertex with at most inline_max shadow rays traces them here, right
        // after its closest hit; the others go through the shadow queue
        const bool inl = (int)n_test <= C.inline_max;
        size_t o = (size_t)g * C.q.shcap + wave_append_n(inl ? 0u : n_test, &C.q.ctr[sidx(S_SHADOW, g)]);
        if (s < 0) return V_NONE;
        const ItemRec& it = *itp;
        const DevScene& D = *Dp;
        const RSettings& st = it.st;
        Ray r;
        r.o = xyz(ro);
        r.d = xyz(rd);
        int outcome;
        int flags = ctl.x & 0xF0;        // keeps glass_root / outcome_aux
        int nf = ctl.z, ns = 0;
        const bool last = FOLD && !path && nf == 0 && !(ctl.x & 0x10) && meta.z + 1 == st.samples * st.samples;
        f3 result(0, 0, 0);   // a FOLDed vertex's value
        if (slot < 0) {
            outcome = path ? V_P_MISS : V_W_MISS;
            if (last) {
                result = ld3(D.bg);
                outcome = V_DONE;
            }
            if (FRESH) S.meta[s] = meta;
        } else {
            HitRec rec;
            PrimView P = load_prim(D.prims, slot);
            float t;
            hit_prim<true>(P, r, 1e-4f, ro.w, t, &rec);
            int kind; f3 albedo, emission, spec_color; float refl, ior, spec, shin;
            load_mat(D.mats, rec.mat, kind, albedo, emission, refl, ior, spec, shin, spec_color);
            Rng rng((uint32_t)meta.w);
            const bool term = last && inl && !(kind == PCL_MAT_METAL && refl > 0.0f) && kind != PCL_MAT_DIELECTRIC;
            if (!term) S.alb[s] = f4(albedo);
            result = emission;
            float4* SH = S.sh + s;   // sample q at SH[q * S.sh_stride]
            uint8_t* OC = S.occl + s;
            const f3 T = path ? xyz(S.T[s]) : f3(1, 1, 1);
            const f3 sn = rec.normal;
            const f3 sh_o = rec.p + rec.normal * st.shadow_bias;
            for (int li = 0; li < D.n_lights; ++li) {
                const LightRec L = load_light(D.lights, li);
                const f3 lpos = ld3(L.pos), I = ld3(L.I);
                const int nsamp = path ? 1 : ((L.radius > 0.0f) ? st.soft_samples : 1);
                f3 contrib(0, 0, 0);
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
                    int skip = 0;   // trace_path_mis adds nothing for ndotl <= 0
                    if (path && st.pathtrace == 2) {
                        if (!(ndotl <= 0.0f)) c = T * mis_light_term(L, albedo, I, Ld, dist, ndotl);
                        else skip = 1 << 16;
                    } else if (path) {
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
                    if (!term) SH[ns * S.sh_stride] = make_float4(c.x, c.y, c.z, __int_as_float(li | (nsamp << 8) | skip));
                    bool occ = false;
                    if (!L.emissive && inl) {
                        Ray sr;
                        sr.o = sh_o;
                        sr.d = Ld;
                        float tb, ta;
                        TravCount<false> tc;
                        occ = traverse<true>(D, sr, st.shadow_bias, dist - st.shadow_bias, stk, tb, ta, tc) >= 0;
                        ++n_inline;
                    }
                    if (!term) OC[ns * S.sh_stride] = occ;
                    else if (!occ) contrib += c;
                    if (!L.emissive && !inl) {
                        C.q.sh_who[o] = make_int2(s, ns);
                        C.q.sh_ray[2 * o] = make_float4(sh_o.x, sh_o.y, sh_o.z, st.shadow_bias);
                        C.q.sh_ray[2 * o + 1] = make_float4(Ld.x, Ld.y, Ld.z, dist - st.shadow_bias);
                        ++o;
                    }
                    ns++;
                }
                if (term) result += contrib / (float)nsamp;
            }
            if (path) {
                // radiance += T * emission comes before the NEE terms (:861)
                S.R[s] = f4(xyz(S.R[s]) + T * emission);
                const int depth = ctl.y;
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
                        nf = 0;
                        Ray rr = push_glass(F, nf, r, rec, albedo, ior, depth, st.shadow_bias);
                        S.gT[s] = f4(Tn);
                        S.gE[s] = f4(emission);
                        S.ray_o[s] = f4(rr.o);
                        S.ray_d[s] = f4(rr.d);
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
                        S.ray_o[s] = f4(rec.p + rec.normal * st.shadow_bias);
                        S.ray_d[s] = f4(nd);
                        S.NT[s] = f4(NT);
                    }
                }
            } else if (term) {
                result += albedo * ld3(D.amb);
                outcome = V_DONE;
            } else {
                S.pend[s] = f4(emission);
                outcome = V_W_LEAF;
                if (kind == PCL_MAT_METAL && refl > 0.0f) {
                    Frame* F = S.frames + (size_t)s * kMaxFrames;
                    Frame& f = F[nf];
                    f.type = FR_METAL; f.albedo = albedo; f.w0 = refl; f.depth = ctl.y;
                    nf += 1;
                    S.ray_o[s] = f4(rec.p + rec.normal * st.shadow_bias);
                    S.ray_d[s] = f4(r.d.normalized().reflect(rec.normal).normalized());
                    outcome = V_W_CHILD;
                } else if (kind == PCL_MAT_DIELECTRIC) {
                    Frame* F = S.frames + (size_t)s * kMaxFrames;
                    Ray rr = push_glass(F, nf, r, rec, albedo, ior, ctl.y, st.shadow_bias);
                    S.ray_o[s] = f4(rr.o);
                    S.ray_d[s] = f4(rr.d);
                    outcome = V_W_CHILD;
                    flags |= 1 << 5;   // glass: the local color is discarded
                }
            }
            meta.w = (int)rng.s;
            S.meta[s] = meta;
        }
        if (outcome == V_DONE) {   // finish_sample of the pixel's last sample
            const f3 acc = (FRESH ? f3(0, 0, 0) : xyz(S.acc[s])) + result;
            const f3 px = acc / (float)(st.samples * st.samples);
            float* o = it.out + (size_t)meta.y * 3;
            o[0] = px.x; o[1] = px.y; o[2] = px.z;
        }
        S.ctl[s] = make_int4((ctl.x & 15) | flags | (outcome << 8), ctl.y, nf, ns);
        return outcome;
}

// Closest hit and shading in one pass: the lane shades its vertex right
// after its traversal, with the ray and the hit still in registers, so the
// hit never goes through memory and the slot state is read once.  ONE: the
// job has a single item (one scene, camera and settings), so its records
// are read with uniform (scalar) loads instead of per lane through the
// slot's item index.  FOLD: see shade_chunk (jobs with inline shadow rays).
template <int K, bool ONE, bool FOLD>
__global__ __launch_bounds__(kWfBlock, 4) void k_closest_shade(Ctx C) {
    if (C.q.ctr[C_PENDING] == 0) return;   // empty iteration
    __shared__ unsigned long long lds[K * kWfBlock];
    LaneStack<K, kStackK + kStackSpill - K> st;
    st.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
    st.set_stride(kWfBlock);
    const int cur = C.q.ctr[C_CUR];
    auto len = [&](int g) -> uint32_t { return C.q.ctr[sidx(S_COUNT0 + cur, g)]; };
    const Slots& S = C.sl;
    TravCount<false> cnt;
    uint32_t n_inline = 0;
    const int home = wave_gid() & (kStripes - 1);
    Fetcher f(home);
    int g;
    uint32_t base, got;
    while (f.next(C.q.ctr, S_HEAD_CLOSEST, kWave, len, g, base, got)) {
        const int s = ((uint32_t)lane_id() < got) ? C.q.ray_q[cur][(size_t)g * C.q.qcap + base + lane_id()] : -1;
        int4 meta = make_int4(0, 0, 0, 0), ctl = make_int4(0, 0, 0, 0);
        float4 ro = make_float4(0, 0, 0, 0), rd = make_float4(0, 0, 0, 0);
        if (s >= 0) {
            meta = S.meta[s]; ctl = S.ctl[s]; ro = S.ray_o[s]; rd = S.ray_d[s];
            const DevScene& D = ONE ? C.scenes[0] : C.scenes[C.items[meta.x].scene];
            Ray r;
            r.o = xyz(ro);
            r.d = xyz(rd);
            float tb, ta;
            const int slot = traverse<false>(D, r, 1e-4f, 1e9f, st, tb, ta, cnt);
            ro.w = ta;
            rd.w = __int_as_float(slot);
        }
        shade_chunk<ONE, FOLD, false>(C, g, s, meta, ctl, ro, rd, st, n_inline);
    }
    wave_add_u64(n_inline, &C.q.stats[(1 + home) * kStatLine + ST_ANY]);
}
// Folds one vertex's light samples exactly as the reference's loops do.
__device__ f3 fold_whitted_local(const Slots& S, int s, int ns, const f3& emission, const f3& albedo, const f3& amb) {
    const float4* SH = S.sh + s;
    const uint8_t* OC = S.occl + s;
    const size_t str = (size_t)S.sh_stride;
    f3 local = emission;
    int q = 0;
    while (q < ns) {
        const int nsamp = (__float_as_int(SH[q * str].w) >> 8) & 0xFF;
        f3 contrib(0, 0, 0);
        for (int k = 0; k < nsamp; ++k, ++q) {
            if (OC[q * str]) continue;
            const float4 v = SH[q * str];
            contrib += f3(v.x, v.y, v.z);
        }
        local += contrib / (float)nsamp;
    }
    local += albedo * amb;
    return local;
}

template <bool ONE>
__global__ __launch_bounds__(kWfBlock, 4) void k_resolve(Ctx C) {
    const int cur = C.q.ctr[C_CUR];
    auto len = [&](int g) -> uint32_t { return C.q.ctr[sidx(S_COUNT0 + cur, g)]; };
    const Slots& S = C.sl;
    if (C.q.ctr[C_PENDING] == 0) return;   // empty iteration
    const bool drain = C.q.ctr[C_DRAIN] != 0;
    const bool pool_left = C.q.ctr[C_POOL] != 0;   // as of the last swap; it only shrinks
    uint32_t rays = 0;
    const int home = wave_gid() & (kStripes - 1);
    Fetcher f(home), pf(home);
    int g;
    uint32_t base, got;
    while (f.next(C.q.ctr, S_HEAD_RESOLVE, kWave, len, g, base, got)) {
        int a = ADV_IDLE;
        int s = -1;
        Lane L;
        if ((uint32_t)lane_id() < got) {
            s = C.q.ray_q[cur][(size_t)g * C.q.qcap + base + lane_id()];
            load_lane(S, s, L);
            const DevScene& D = ONE ? C.scenes[0] : C.scenes[C.items[L.item].scene];
            const f3 amb = ld3(D.amb);
            int status;
            f3 result(0, 0, 0);
            switch (L.outcome) {
            case V_DONE: status = ADV_NEED_PIXEL; break;   // written by the closest-hit kernel
            case V_W_MISS: status = 3; result = ld3(D.bg); break;
            case V_W_LEAF: status = 3; result = fold_whitted_local(S, s, L.nsh, xyz(S.pend[s]), xyz(S.alb[s]), amb); break;
            case V_W_CHILD: {
                if (L.outcome_aux) {
                    L.outcome_aux = 0;            // glass: local discarded (its draws already made)
                } else {
                    Frame* F = S.frames + (size_t)s * kMaxFrames;
                    F[L.nf - 1].acc = fold_whitted_local(S, s, L.nsh, xyz(S.pend[s]), xyz(S.alb[s]), amb);
                }
                L.depth -= 1;
                status = ADV_QUERY;
                break;
            }
            case V_P_MISS: {
                L.R += L.T * ld3(D.bg);
                status = finish_sample(C, L, L.R, rays, drain);
                break;
            }
            default: {   // path vertex with a hit: NEE terms in light order, then ambient
                const float4* SH = S.sh + s;
                const uint8_t* OC = S.occl + s;
                const size_t str = (size_t)S.sh_stride;
                for (int q = 0; q < L.nsh; ++q)
                    if (!OC[q * str]) {
                        const float4 v = SH[q * str];
                        if (!(__float_as_int(v.w) & (1 << 16))) L.R += f3(v.x, v.y, v.z);
                    }
                L.R += L.T * xyz(S.alb[s]) * amb;
                if (L.outcome == V_P_END) {
                    status = finish_sample(C, L, L.R, rays, drain);
                } else if (L.outcome == V_P_GLASS) {
                    L.mode = M_WHITTED;
                    L.glass_root = 1;
                    L.depth -= 1;                 // trace(refl, depth - 1)
                    status = ADV_QUERY;
                } else {
                    L.T = xyz(S.NT[s]);
                    L.depth -= 1;
                    status = ADV_QUERY;
                }
                break;
            }
            }
            L.outcome = V_NONE;
            a = advance(C, s, L, status, result, rays, drain);
        }
        a = refill(C, s, L, a, pf, rays, drain, pool_left);
        if (s >= 0) {
            if (a == ADV_QUERY) {
                store_lane(S, s, L);
            } else {   // idle or handed off: the finish list reads meta and acc only
                S.meta[s] = make_int4(L.item, L.pixel, L.sample, (int)L.rng);
                if (a == ADV_HANDOFF) S.acc[s] = f4(L.acc);
            }
        }
        const int pos = wave_append(a == ADV_QUERY, &C.q.ctr[sidx(S_COUNT0 + (cur ^ 1), g)]);
        if (pos >= 0) C.q.ray_q[cur ^ 1][(size_t)g * C.q.qcap + pos] = s;
    }
    wave_add_u64(rays, &C.q.stats[(1 + home) * kStatLine + ST_RAYS]);
}

// Starts the job and runs the closest-hit half of its first iteration: every
// slot takes a pixel, makes its first ray (as k_resolve starts a pixel) and,
// with the ray still in registers, traces and shades it as k_closest_shade
// does; k_anyhit, k_resolve and k_swap then finish the iteration.  The slot
// state is stored only where the shading or k_resolve reads it: T / R of a
// path before the shading, which writes the meta record, the ray and the
// control word itself; the (zero) sample sum after it, unless the shading
// finished the pixel.  Slot wave b (slots 64b .. 64b + 63) belongs to stripe
// b % kStripes.
// One lane per slot, the LDS stack of the default depth.
template <bool ONE, bool FOLD>
__global__ __launch_bounds__(kWfBlock, 4) void k_start_closest(Ctx C, int n_slots) {
    __shared__ unsigned long long lds[kStackK * kWfBlock];
    Stack st;
    st.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
    st.set_stride(kWfBlock);
    uint32_t rays = 0, n_inline = 0;
    const int cur = C.q.ctr[C_CUR];
    const int s0 = wave_gid() * kWave;
    const int s = s0 + lane_id();
    const int g = stripe_of(s0);
    Fetcher pf(g);
    Lane L;
    L.item = -1; L.pixel = 0; L.sample = 0; L.rng = 0;
    L.mode = 0; L.glass_root = 0; L.outcome_aux = 0; L.outcome = V_NONE; L.depth = 0; L.nf = 0; L.nsh = 0;
    L.o = L.d = L.acc = L.T = L.R = f3(0, 0, 0);
    L.ray_set = false;
    int a = (s < n_slots) ? ADV_NEED_PIXEL : ADV_IDLE;
    a = refill(C, s, L, a, pf, rays, false, true);
    const Slots& S = C.sl;
    if (s < n_slots) {
        if (L.mode == M_PATH) {
            S.T[s] = f4(L.T);
            S.R[s] = f4(L.R);
        }
        if (a != ADV_QUERY) {   // no query (an idle slot): the record as it is
            S.meta[s] = make_int4(L.item, L.pixel, L.sample, (int)L.rng);
            S.acc[s] = f4(L.acc);
            S.ctl[s] = make_int4(pack_flags(L), L.depth, L.nf, L.nsh);
        }
    }
    const int pos = wave_append(a == ADV_QUERY, &C.q.ctr[sidx(S_COUNT0 + cur, g)]);
    if (pos >= 0) C.q.ray_q[cur][(size_t)g * C.q.qcap + pos] = s;
    wave_add_u64(rays, &C.q.stats[(1 + g) * kStatLine + ST_RAYS]);
    // the first closest-hit query (k_closest_shade's body for this slot)
    int4 meta = make_int4(0, 0, 0, 0), ctl = make_int4(0, 0, 0, 0);
    float4 ro = make_float4(0, 0, 0, 0), rd = make_float4(0, 0, 0, 0);
    if (a == ADV_QUERY) {
        meta = make_int4(L.item, L.pixel, L.sample, (int)L.rng);
        ctl = make_int4(pack_flags(L), L.depth, L.nf, L.nsh);
        const DevScene& D = ONE ? C.scenes[0] : C.scenes[C.items[L.item].scene];
        Ray r;
        r.o = L.o;
        r.d = L.d;
        float tb, ta;
        TravCount<false> cnt;
        const int slot = traverse<false>(D, r, 1e-4f, 1e9f, st, tb, ta, cnt);
        ro = f4(L.o, ta);
        rd = f4(L.d, __int_as_float(slot));
    }
    const int outcome = shade_chunk<ONE, FOLD, true>(C, g, a == ADV_QUERY ? s : -1, meta, ctl, ro, rd, st, n_inline);
    if (a == ADV_QUERY && outcome != V_DONE) S.acc[s] = f4(L.acc);
    wave_add_u64(n_inline, &C.q.stats[(1 + g) * kStatLine + ST_ANY]);
}

// End of an iteration, one lane per stripe: account the queries, flip the
// ray queue, clear the stripe's heads, publish the next queue's length and
// the number of job pixels no slot has taken yet.
__global__ void k_swap(Queues q) {
    const int g = threadIdx.x;
    const uint32_t cur = q.ctr[C_CUR];
    uint32_t n = q.ctr[sidx(S_COUNT0 + cur, g)];
    uint32_t nsh = q.ctr[sidx(S_SHADOW, g)];
    uint32_t nn = q.ctr[sidx(S_COUNT0 + (cur ^ 1), g)];
    const uint32_t tiles = (uint32_t)(q.total_pixels >> 6);
    const uint32_t plen = (tiles / kStripes + ((uint32_t)g < tiles % kStripes ? 1u : 0u)) * 64u;
    const uint32_t phead = q.ctr[sidx(S_PIXEL, g)];
    uint32_t pool = phead < plen ? plen - phead : 0u;
    q.ctr[sidx(S_COUNT0 + cur, g)] = 0;
    q.ctr[sidx(S_SHADOW, g)] = 0;
    for (int h = S_HEAD_CLOSEST; h <= S_HEAD_RESOLVE; ++h) q.ctr[sidx(h, g)] = 0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        n += __shfl_xor(n, m, kWave);
        nsh += __shfl_xor(nsh, m, kWave);
        nn += __shfl_xor(nn, m, kWave);
        pool += __shfl_xor(pool, m, kWave);
    }
    if (g == 0) {
        q.stats[ST_CLOSEST] += n;
        q.stats[ST_ANY] += nsh;
        q.stats[ST_ITERS] += n ? 1 : 0;
        q.ctr[C_CUR] = cur ^ 1;
        q.ctr[C_PENDING] = nn;
        q.ctr[C_POOL] = pool;
    }
}

__global__ void k_set_word(uint32_t* p, uint32_t v) { *p = v; }

// After a drained run every slot still holding a pixel stopped at a sample
// boundary (ADV_HANDOFF); the others hold none (item -1).  Lists the former's
// records, in slot order within each wave.
__global__ __launch_bounds__(kWfBlock) void k_finish_list(Ctx C, int n_slots) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    int4 m = make_int4(-1, 0, 0, 0);
    if (s < n_slots) m = C.sl.meta[s];
    const int pos = wave_append(m.x >= 0, &C.q.ctr[C_NFINISH]);
    if (pos >= 0) {
        C.q.fin_meta[pos] = m;
        C.q.fin_acc[pos] = C.sl.acc[s];
    }
}

// Deferred glass vertices of the finish launch: the task array (GlassSink),
// sample k of listed pixel p (entry from + p) at samp[(k * n_px + p) * 3],
// and per listed pixel the sum of its samples before the first deferred one
// and that sample's index.
struct FinishDefer {
    GlassSink sink;
    float* samp;
    float* prefix;
    int32_t* first;
    int64_t n_px;
};

// Finishes the listed pixels [from .. to) one per lane with the per-pixel
// program (render_samples, the megakernel's loop from the record's next
// sample on).  The tail of a job -- few pixels with long sample chains, such
// as pixels that see glass -- then runs without a barrier per query.  With
// DEFER (path-traced jobs) a path that hits glass leaves the vertex's Whitted
// subtree to k_finish_glass, as the megakernel's deferred program does, and a
// pixel that deferred a sample is left to k_finish_sum; DEFER 2 compiles the
// subtree walk that skips the generator past Whitted's draws (an item with
// soft lights), 1 leaves it out.
// Register budget as the megakernel's for the same mode.
template <int PT, int DEFER>
__global__ __launch_bounds__(kWfBlock, PT == 1 ? 4 : 3) void k_finish(Ctx C, const int4* meta, const float4* accs,
                                                                    int from, int to, FinishDefer fd) {
    __shared__ unsigned long long lds[kStackK * kWfBlock];
    const int e = from + blockIdx.x * blockDim.x + threadIdx.x;
    LaneCount<false> c;
    if (e < to) {
        const int4 m = meta[e];
        const ItemRec& it = C.items[m.x];
        const DevScene& D = C.scenes[it.scene];
        alignas(16) unsigned char fbuf[sizeof(Frame) * kMaxFrames];
        Frame* F = reinterpret_cast<Frame*>(fbuf);
        Stack stk;
        stk.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
        stk.set_stride(kWfBlock);
        const int W = it.st.width;
        Rng rng((uint32_t)m.w);
        f3 px;
        bool done = true;
        if constexpr (DEFER > 0) {
            const int64_t p = e - from;
            GlassSink sink = fd.sink;
            sink.draws = it.draws;
            sink.pixel = (int32_t)p;
            sink.item = m.x;
            int first;
            done = render_samples_defer<DEFER == 2, true>(D, it.cam, it.st, m.y % W, m.y / W, m.z, rng, xyz(accs[e]),
                                                          F, stk, c, sink, fd.samp + p * 3, (size_t)fd.n_px * 3, px,
                                                          first);
            fd.first[p] = first;
            if (!done) {
                float* q = fd.prefix + p * 3;
                q[0] = px.x; q[1] = px.y; q[2] = px.z;
            }
        } else {
            px = render_samples<PT>(D, it.cam, it.st, m.y % W, m.y / W, m.z, rng, xyz(accs[e]), F, stk, c);
        }
        if (done) {
            float* o = it.out + (size_t)m.y * 3;
            o[0] = px.x; o[1] = px.y; o[2] = px.z;
        }
    }
    unsigned long long* line = &C.q.stats[(1 + (blockIdx.x & (kStripes - 1))) * kStatLine];
    wave_add_u64(c.rays, line + ST_RAYS);
    wave_add_u64(c.closest, line + ST_CLOSEST);
    wave_add_u64(c.anyhit, line + ST_ANY);
}

// The finish launch's deferred glass vertices: one lane per task (grid-stride
// over 64-task chunks), glass_task's subtree in the task's item's scene, the
// sample's value into its slot.
__global__ __launch_bounds__(kWfBlock, 4) void k_finish_glass(Ctx C, const float4* __restrict__ tasks,
                                                             const uint32_t* count, uint32_t cap, float* samp,
                                                             int64_t n_px, const uint32_t* __restrict__ perm) {
    __shared__ unsigned long long lds[kStackK * kWfBlock];
    const uint32_t n = min(*count, cap);
    LaneCount<false> c;
    alignas(16) unsigned char fbuf[sizeof(Frame) * kMaxFrames];
    Frame* F = reinterpret_cast<Frame*>(fbuf);
    Stack stk;
    stk.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
    stk.set_stride(kWfBlock);
    const uint32_t wave = (blockIdx.x * kWfBlock + threadIdx.x) >> 6, waves = (gridDim.x * kWfBlock) >> 6;
    for (uint32_t base = wave * 64u; base < n; base += waves * 64u) {
        const uint32_t i = base + (threadIdx.x & 63);
        if (i >= n) continue;
        const uint32_t t = perm ? perm[i] : i;   // the order of glass_sort.hpp
        const float4* q = tasks + (size_t)t * 8;
        const ItemRec& it = C.items[__float_as_int(q[7].w)];
        const f3 v = glass_task(C.scenes[it.scene], it.st, q, F, stk, c);
        const int64_t p = __float_as_int(q[4].w), k = __float_as_int(q[5].w);
        float* o = samp + (k * n_px + p) * 3;
        o[0] = v.x; o[1] = v.y; o[2] = v.z;
    }
    unsigned long long* line = &C.q.stats[(1 + (blockIdx.x & (kStripes - 1))) * kStatLine];
    wave_add_u64(c.rays, line + ST_RAYS);
    wave_add_u64(c.closest, line + ST_CLOSEST);
    wave_add_u64(c.anyhit, line + ST_ANY);
}

// Listed pixels that deferred a sample: the prefix plus samples first..n^2-1
// in order, divided by n^2 -- render_samples' sum, same additions, same order.
__global__ __launch_bounds__(256) void k_finish_sum(Ctx C, const int4* __restrict__ meta, int from,
                                                   const float* __restrict__ samp, const float* __restrict__ prefix,
                                                   const int32_t* __restrict__ first, int64_t n_px) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_px) return;
    const int k0 = first[p];
    const int4 m = meta[from + p];
    const ItemRec& it = C.items[m.x];
    const int nn = it.st.samples * it.st.samples;
    if (k0 >= nn) return;
    f3 acc(prefix[p * 3], prefix[p * 3 + 1], prefix[p * 3 + 2]);
    for (int k = k0; k < nn; ++k) {
        const float* s = samp + ((int64_t)k * n_px + p) * 3;
        acc += f3(s[0], s[1], s[2]);
    }
    const f3 px = acc / (float)nn;
    float* o = it.out + (size_t)m.y * 3;
    o[0] = px.x; o[1] = px.y; o[2] = px.z;
}

// Clears the job counters (a kernel, not a memset, so it may sit in a graph).
// C_PENDING starts non-zero so the first iteration runs; C_POOL starts at the
// job's pixel count (the swap kernel keeps both up to date afterwards).
__global__ void k_reset(Queues q) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kCtrWords; i += gridDim.x * blockDim.x)
        q.ctr[i] = i == C_PENDING ? 1u : i == C_POOL ? (uint32_t)q.total_pixels : 0u;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kStatWords; i += gridDim.x * blockDim.x) q.stats[i] = 0;
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

Engine::~Engine() {
    release();
    if (h_pending) (void)hipHostFree(h_pending);
    for (hipEvent_t& e : ev_pending)
        if (e) (void)hipEventDestroy(e);
}

void Engine::release() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
    for (void* p : bufs) (void)hipFree(p);
    bufs.clear();
    cap_slots = 0;
    cap_items = 0;
    for (void* p : {(void*)d_gtasks, (void*)d_gcount, (void*)d_samp, (void*)d_prefix, (void*)d_first,
                    (void*)d_gkeys, (void*)d_gperm, (void*)d_gbins})
        if (p) (void)hipFree(p);
    d_gtasks = nullptr; d_gcount = nullptr; d_samp = nullptr; d_prefix = nullptr; d_first = nullptr;
    d_gkeys = d_gperm = d_gbins = nullptr;
    gtask_cap = samp_cap = px_cap = gkey_cap = 0;
}

// The finish launch's deferred-glass buffers: 12 B per listed pixel and
// sample, a task array sized for a quarter of those samples (the listed
// pixels are the job's longest, often glass-heavy: 30 glass vertices per
// listed pixel on the 256-spp Cornell + mesh frame, against a capacity of 64;
// tasks beyond it are evaluated in place), 16 B per listed pixel; together
// under half of the device's free memory.  The task counter is 32-bit, so
// n_px * nn stays below 2^32.
bool Engine::defer_buffers(int64_t n_px, int64_t nn) {
    if (n_px <= 0 || n_px * nn >= ((int64_t)1 << 32)) return false;
    const size_t samp = (size_t)(n_px * nn * 3);
    const size_t tasks = defer_tasks > 0 ? (size_t)defer_tasks
                                         : std::min<size_t>((size_t)(n_px * nn / 4 + 4096), (size_t)1 << 27);
    size_t need = 0;
    if (samp > samp_cap) need += samp * sizeof(float);
    if (tasks > gtask_cap) need += tasks * (8 * sizeof(float4) + 2 * sizeof(uint32_t));
    if ((size_t)n_px > px_cap) need += (size_t)n_px * 16;
    if (need) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || need > free_b / 2) return false;
    }
    if (!d_gcount && hipMalloc((void**)&d_gcount, 256) != hipSuccess) { d_gcount = nullptr; return false; }
    auto grow = [](auto** p, size_t& cap, size_t want, size_t elem) -> bool {
        if (want <= cap) return true;
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        cap = 0;
        if (hipMalloc((void**)p, want * elem) != hipSuccess) { *p = nullptr; return false; }
        cap = want;
        return true;
    };
    size_t pxc = px_cap, pxc2 = px_cap;
    if (!grow(&d_samp, samp_cap, samp, sizeof(float)) || !grow(&d_gtasks, gtask_cap, tasks, 8 * sizeof(float4)) ||
        !grow(&d_prefix, pxc, (size_t)n_px, 3 * sizeof(float)) || !grow(&d_first, pxc2, (size_t)n_px, sizeof(int32_t))) {
        px_cap = std::min(pxc, pxc2);
        return false;
    }
    px_cap = std::min(pxc, pxc2);
    size_t kc = gkey_cap, kc2 = gkey_cap, bc = d_gbins ? 512 : 0;
    const bool ok = grow(&d_gkeys, kc, tasks, sizeof(uint32_t)) && grow(&d_gperm, kc2, tasks, sizeof(uint32_t)) &&
                    grow(&d_gbins, bc, 512, sizeof(uint32_t));
    gkey_cap = std::min(kc, kc2);
    return ok;
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
    A(sl.meta, int4, S) A(sl.ctl, int4, S) A(sl.ray_o, float4, S) A(sl.ray_d, float4, S)
    A(sl.acc, float4, S) A(sl.T, float4, S) A(sl.NT, float4, S) A(sl.R, float4, S) A(sl.pend, float4, S)
    A(sl.alb, float4, S) A(sl.gT, float4, S) A(sl.gE, float4, S)
    A(sl.sh, float4, S * kMaxSh) A(sl.occl, uint8_t, S * kMaxSh)
    sl.sh_stride = (int64_t)S;
    A(sl.frames, Frame, S * kMaxFrames)
    const size_t qcap = ((S + 63) / 64 + kStripes - 1) / kStripes * 64;   // a stripe's slots
    const size_t Q = qcap * kStripes, SH = Q * kMaxSh;
    A(q.ray_q[0], int32_t, Q) A(q.ray_q[1], int32_t, Q)
    A(q.fin_meta, int4, S) A(q.fin_acc, float4, S)
    A(q.sh_who, int2, SH) A(q.sh_ray, float4, 2 * SH)
    A(q.ctr, uint32_t, kCtrWords) A(q.stats, unsigned long long, kStatWords)
    q.qcap = (int32_t)qcap;
    q.shcap = (int32_t)(qcap * kMaxSh);
    A(d_items, ItemRec, items) A(d_scenes, DevScene, items)
#undef A
    cap_slots = slots;
    cap_items = items;
    return PCL_OK;
}

template <int K, bool ONE, bool FOLD>
static void launch_iter(const Ctx& c, int grid_trav, int grid_other, hipStream_t st) {
    hipLaunchKernelGGL((k_closest_shade<K, ONE, FOLD>), dim3(grid_trav), dim3(kWfBlock), 0, st, c);
    hipLaunchKernelGGL((k_anyhit<K, ONE>), dim3(grid_trav), dim3(kWfBlock), 0, st, c);
    hipLaunchKernelGGL(k_resolve<ONE>, dim3(grid_other), dim3(kWfBlock), 0, st, c);
    hipLaunchKernelGGL(k_swap, dim3(1), dim3(kStripes), 0, st, c.q);
}

// FOLD (Engine::run decides): the closest-hit kernel finishes terminal
// one-sample Whitted pixels itself.
template <int K>
static void launch_iter(const Ctx& c, bool one, bool fold, int grid_trav, int grid_other, hipStream_t st) {
    if (one) {
        if (fold) launch_iter<K, true, true>(c, grid_trav, grid_other, st);
        else launch_iter<K, true, false>(c, grid_trav, grid_other, st);
    } else {
        if (fold) launch_iter<K, false, true>(c, grid_trav, grid_other, st);
        else launch_iter<K, false, false>(c, grid_trav, grid_other, st);
    }
}

int Engine::run(hipStream_t stream, const std::vector<ItemRec>& items, const std::vector<DevScene>& scenes,
                const Config& cfg_in, int cus, RunStats& out, std::string& err, int part, int parts,
                hipEvent_t done) {
    Config cfg = cfg_in;
    int64_t job = 0;
    for (const auto& it : items) job += padded_pixels(it.st.width, it.st.height);
    if (job >= (int64_t)1 << 31) { err = "job larger than 2^31 pixels"; return PCL_ERR_LIMIT; }
    const int64_t total = part_pixels(job, part, parts);
    if (total <= 0) {
        out = RunStats{};
        if (done) WF_TRY(hipEventRecord(done, stream));
        return PCL_OK;
    }
    if (cfg.slots <= 0) cfg.slots = (int)std::min<int64_t>(total, 2 << 20);
    cfg.slots = (int)std::max<int64_t>(1, std::min<int64_t>(cfg.slots, total));
    if (cfg.blocks_per_cu <= 0) cfg.blocks_per_cu = 8;
    if (cfg.stack_k != 8 && cfg.stack_k != 16 && cfg.stack_k != 32) cfg.stack_k = 16;
    if (cfg.iters_per_graph <= 0) cfg.iters_per_graph = 4;
    // auto: drain once fewer slots are active than four times the persistent
    // traversal grids' lanes (blocks_per_cu blocks of kWfBlock lanes per CU):
    // with its glass vertices deferred the finish launch runs the tail
    // better than iterations with that few slots (Cornell + mesh, 256 spp:
    // 2.39 s, against 2.45 s at once the grids' lanes, 2.48 s at six times;
    // the 64-scene job the same at one to four times)
    if (cfg.drain_below == 0) cfg.drain_below = kDrainLanes * cus * cfg.blocks_per_cu * kWfBlock;
    if (cfg.inline_shadows == 0) cfg.inline_shadows = 1;
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
    c.inline_max = cfg.inline_shadows;
    c.sl = sl;
    c.q = q;
    c.q.total_pixels = total;
    c.q.part = part;
    c.q.parts = parts;
    const bool one = its.size() == 1;   // part of the graph key through graph_items
    // Finishing pixels in the closest-hit kernel pays where most pixels end
    // at their first vertex: one-sample Whitted jobs (the 1-spp stress frame
    // 2 % faster, 10 % less HBM traffic); with more samples only a pixel's
    // last sample could, and the larger kernel cost the 16-spp 64-scene job
    // 1 %.  Needs inline shadow rays.
    bool fold = cfg.inline_shadows > 0;
    for (const auto& it : its) fold &= !it.st.pathtrace && it.st.samples == 1;
    const int grid_trav = cus * cfg.blocks_per_cu;
    const int grid_other = cus * 8;   // 4 waves per SIMD, the register limit of shade / resolve
    const bool graph_ok = (graph_exec && graph_key.slots == cfg.slots && graph_key.blocks_per_cu == cfg.blocks_per_cu &&
                           graph_key.stack_k == cfg.stack_k && graph_key.iters_per_graph == cfg.iters_per_graph &&
                           graph_key.inline_shadows == cfg.inline_shadows &&
                           graph_items == (int)its.size() && graph_total == total && graph_part == part &&
                           graph_parts == parts && graph_fold == fold);
    if (!graph_ok) {
        if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
        hipGraph_t g = nullptr;
        WF_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < cfg.iters_per_graph; ++i) {
            if (cfg.stack_k == 8) launch_iter<8>(c, one, fold, grid_trav, grid_other, stream);
            else if (cfg.stack_k == 32) launch_iter<32>(c, one, fold, grid_trav, grid_other, stream);
            else launch_iter<16>(c, one, fold, grid_trav, grid_other, stream);
        }
        WF_TRY(hipStreamEndCapture(stream, &g));
        hipError_t e = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) { graph_exec = nullptr; err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e); return PCL_ERR_HIP; }
        graph_key = cfg;
        graph_items = (int)its.size();
        graph_total = total;
        graph_part = part;
        graph_parts = parts;
        graph_fold = fold;
    }
    if (!h_pending)
        WF_TRY(hipHostMalloc((void**)&h_pending, 2 * kStatusWords * sizeof(uint32_t), hipHostMallocDefault));
    for (hipEvent_t& e : ev_pending)
        if (!e) WF_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    bool any_path = false, any_whitted = false;
    for (const auto& it : its) (it.st.pathtrace ? any_path : any_whitted) = true;
    // finish kernel over records [from .. to) on stream st
    bool any_glass = false, any_draws = false, draws_ok = true;
    int64_t nn_max = 1;
    for (const auto& it : its) {
        any_glass |= it.glass != 0;
        any_draws |= it.draws > 0;
        draws_ok &= it.draws <= (1 << 15);   // a subtree's draws stay below 2^31
        nn_max = std::max<int64_t>(nn_max, (int64_t)it.st.samples * it.st.samples);
    }
    // Deferred glass vertices in the finish launch (as the megakernel's): for
    // path-traced jobs with a dielectric, when the buffers fit.
    bool deferred = false;
    auto finish = [&](const int4* m, const float4* a, uint32_t from, uint32_t to, hipStream_t st) -> hipError_t {
        if (to <= from) return hipSuccess;
        const dim3 grid((to - from + kWfBlock - 1) / kWfBlock);
        FinishDefer fd{};
        const int64_t n_px = (int64_t)(to - from);
        deferred = any_path && !any_whitted && any_glass && draws_ok && defer_glass != 0 &&
                   defer_buffers(n_px, nn_max);
        if (deferred) {
            hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, st, d_gcount, 0u);
            fd.sink.tasks = d_gtasks;
            fd.sink.count = d_gcount;
            const size_t tcap = std::min(gtask_cap, gkey_cap);   // either may have grown alone before a failure
            fd.sink.cap = (uint32_t)std::min<size_t>(tcap, defer_tasks > 0 ? (size_t)defer_tasks : tcap);
            fd.sink.keys = d_gkeys;
            fd.samp = d_samp;
            fd.prefix = d_prefix;
            fd.first = d_first;
            fd.n_px = n_px;
            if (any_draws) hipLaunchKernelGGL((k_finish<1, 2>), grid, dim3(kWfBlock), 0, st, c, m, a, (int)from, (int)to, fd);
            else hipLaunchKernelGGL((k_finish<1, 1>), grid, dim3(kWfBlock), 0, st, c, m, a, (int)from, (int)to, fd);
            glass_sort(d_gkeys, d_gcount, fd.sink.cap, d_gbins, d_gperm, st);
            int nb = 0;
            hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_finish_glass, kWfBlock, 0);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_finish_glass, dim3(cus * std::max(nb, 1)), dim3(kWfBlock), 0, st, c, d_gtasks,
                               d_gcount, fd.sink.cap, d_samp, n_px, d_gperm);
            hipLaunchKernelGGL(k_finish_sum, dim3((unsigned)((n_px + 255) / 256)), dim3(256), 0, st, c, m, (int)from,
                               d_samp, d_prefix, d_first, n_px);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            return hipMemcpyAsync(&h_pending[1], d_gcount, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        }
        if (any_path && !any_whitted) hipLaunchKernelGGL((k_finish<1, 0>), grid, dim3(kWfBlock), 0, st, c, m, a, (int)from, (int)to, fd);
        else if (!any_path) hipLaunchKernelGGL((k_finish<0, 0>), grid, dim3(kWfBlock), 0, st, c, m, a, (int)from, (int)to, fd);
        else hipLaunchKernelGGL((k_finish<-1, 0>), grid, dim3(kWfBlock), 0, st, c, m, a, (int)from, (int)to, fd);
        return hipGetLastError();
    };
    // timing events, destroyed on every return path
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() { for (hipEvent_t x : e) if (x) (void)hipEventDestroy(x); }
    } ev;
    WF_TRY(hipEventCreate(&ev.e[0]));
    WF_TRY(hipEventCreate(&ev.e[1]));
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    WF_TRY(hipEventRecord(e0, stream));
    hipLaunchKernelGGL(k_reset, dim3(16), dim3(256), 0, stream, c.q);
    {   // the first iteration: start + closest hit in one kernel, then the rest of the iteration
        const dim3 sg((cfg.slots + kWfBlock - 1) / kWfBlock), b(kWfBlock);
        if (one) {
            if (fold) hipLaunchKernelGGL((k_start_closest<true, true>), sg, b, 0, stream, c, cfg.slots);
            else hipLaunchKernelGGL((k_start_closest<true, false>), sg, b, 0, stream, c, cfg.slots);
            hipLaunchKernelGGL((k_anyhit<kStackK, true>), dim3(grid_trav), b, 0, stream, c);
            hipLaunchKernelGGL(k_resolve<true>, dim3(grid_other), b, 0, stream, c);
        } else {
            if (fold) hipLaunchKernelGGL((k_start_closest<false, true>), sg, b, 0, stream, c, cfg.slots);
            else hipLaunchKernelGGL((k_start_closest<false, false>), sg, b, 0, stream, c, cfg.slots);
            hipLaunchKernelGGL((k_anyhit<kStackK, false>), dim3(grid_trav), b, 0, stream, c);
            hipLaunchKernelGGL(k_resolve<false>, dim3(grid_other), b, 0, stream, c);
        }
        hipLaunchKernelGGL(k_swap, dim3(1), dim3(kStripes), 0, stream, c.q);
    }
    WF_TRY(hipGetLastError());
    // Two graph launches in flight: while the host reads the status words
    // after launch k, launch k + 1 already runs, so the GPU never idles for
    // the round trip.  The launch after the last busy one finds every queue
    // empty and only resets the counters.
    int launches = 0;
    auto issue = [&]() -> hipError_t {
        hipError_t e = hipGraphLaunch(graph_exec, stream);
        const int k = launches++ & 1;
        if (e == hipSuccess)
            e = hipMemcpyAsync(&h_pending[kStatusWords * k], q.ctr, kStatusWords * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipEventRecord(ev_pending[k], stream);
        return e;
    };
    const double t_begin = now_ms();
    bool drained = false;
    WF_TRY(issue());
    WF_TRY(issue());
    for (int k = 0;; ++k) {
        WF_TRY(hipEventSynchronize(ev_pending[k & 1]));
        const uint32_t* hp = &h_pending[kStatusWords * (k & 1)];
        if (hp[C_PENDING] == 0) break;
        if (!drained && cfg.drain_below > 0 && hp[C_POOL] == 0 && hp[C_PENDING] < (uint32_t)cfg.drain_below) {
            // takes effect from the launch after the one in flight
            hipLaunchKernelGGL(k_set_word, dim3(1), dim3(1), 0, stream, q.ctr + C_DRAIN, 1u);
            WF_TRY(hipGetLastError());
            drained = true;
        }
        WF_TRY(issue());   // reuses slot k & 1, already read
        // every iteration retires at least one query of a finite program; a
        // job still running after this long is a bug and reported as one,
        // after the device finishes the (at most two) graph launches in
        // flight -- each a fixed number of bounded iterations -- so that no
        // kernel still uses the buffers when the caller gets the error
        if (now_ms() - t_begin > 300000.0) {
            (void)hipDeviceSynchronize();
            err = "wavefront did not drain in 300 s";
            return PCL_ERR_LIMIT;
        }
    }
    uint32_t n_finish = 0;
    if (drained) {
        hipLaunchKernelGGL(k_finish_list, dim3((cfg.slots + kWfBlock - 1) / kWfBlock), dim3(kWfBlock), 0, stream, c,
                           cfg.slots);
        WF_TRY(hipGetLastError());
        WF_TRY(hipMemcpyAsync(&h_pending[0], q.ctr + C_NFINISH, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        WF_TRY(hipStreamSynchronize(stream));
        n_finish = h_pending[0];
        if (n_finish > (uint32_t)cfg.slots) {
            (void)hipDeviceSynchronize();
            err = "finish list longer than the slot pool";
            return PCL_ERR_LIMIT;
        }
        WF_TRY(finish(q.fin_meta, q.fin_acc, 0, n_finish, stream));
    }
    WF_TRY(hipEventRecord(e1, stream));
    if (done) WF_TRY(hipEventRecord(done, stream));
    std::vector<unsigned long long> h_st(kStatWords);
    WF_TRY(hipMemcpyAsync(h_st.data(), q.stats, kStatWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    WF_TRY(hipStreamSynchronize(stream));
    for (int g = 0; g < kStripes; ++g)
        for (int w : {ST_RAYS, ST_CLOSEST, ST_ANY}) h_st[w] += h_st[(1 + g) * kStatLine + w];
    float ms = 0.0f;
    WF_TRY(hipEventElapsedTime(&ms, e0, e1));
    out.rays = h_st[ST_RAYS];
    out.closest = h_st[ST_CLOSEST];
    out.anyhit = h_st[ST_ANY];
    out.iterations = h_st[ST_ITERS];
    out.graph_launches = launches;
    out.finished = n_finish;
    out.deferred = deferred ? std::min<uint64_t>(h_pending[1], std::min<size_t>(gtask_cap, defer_tasks > 0 ? (size_t)defer_tasks : gtask_cap)) : 0;
    out.ms = ms;
    out.cfg = cfg;
    return PCL_OK;
}

Tracer::~Tracer() {
    for (hipStream_t& s : side)
        if (s) (void)hipStreamDestroy(s);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    for (hipEvent_t& e : ev_done)
        if (e) (void)hipEventDestroy(e);
}

void Tracer::release() {
    for (Engine& e : eng) e.release();
}

size_t Tracer::cap_slots() const {
    size_t n = 0;
    for (const Engine& e : eng) n += (size_t)e.cap_slots;
    return n;
}

int Tracer::run(int device, hipStream_t stream, const std::vector<ItemRec>& items, const std::vector<DevScene>& scenes,
                const Config& cfg_in, int cus, RunStats& out, std::string& err) {
    for (Engine& e : eng) {
        e.defer_glass = defer_glass;
        e.defer_tasks = defer_tasks;
    }
    Config cfg = cfg_in;
    int64_t total = 0;
    for (const auto& it : items) total += padded_pixels(it.st.width, it.st.height);
    if (total <= 0) { out = RunStats{}; out.cfg = cfg; return PCL_OK; }
    // job-level knobs resolved here, then split among the parts
    if (cfg.slots <= 0) cfg.slots = (int)std::min<int64_t>(total, 2 << 20);
    cfg.slots = (int)std::max<int64_t>(1, std::min<int64_t>(cfg.slots, total));
    if (cfg.blocks_per_cu <= 0) cfg.blocks_per_cu = 8;
    if (cfg.drain_below == 0) cfg.drain_below = kDrainLanes * cus * cfg.blocks_per_cu * kWfBlock;
    if (cfg.streams <= 0) {
        // Concurrent parts pay where iterations repeat: a one-sample job runs
        // a handful of iterations and is faster as one engine (headline frame
        // 1.18 ms vs 1.37 ms in two parts); path-traced jobs with >= 16 spp
        // run best in three (Cornell + mesh 256 spp: 3.17 / 2.87 / 2.73 s in
        // one / two / three parts), other jobs in two (64 Whitted scenes at 16
        // spp: 155.5 / 151 / 157 ms).  Four parts are slower than one with
        // the four hardware queues HIP gives a process by default.
        bool multi = false, long_paths = false;
        for (const auto& it : items) {
            multi |= it.st.samples > 1;
            long_paths |= it.st.pathtrace && it.st.samples >= 4;
        }
        cfg.streams = long_paths ? 3 : multi ? 2 : 1;
    }
    cfg.streams = std::min(cfg.streams, kMaxParts);
    const int parts = (int)std::max<int64_t>(1, std::min<int64_t>(cfg.streams, total >> 6));
    cfg.streams = parts;
    if (parts == 1) {
        const int rc = eng[0].run(stream, items, scenes, cfg, cus, out, err);
        out.cfg.streams = 1;
        return rc;
    }
    Config pc = cfg;
    pc.slots = std::max(1, (cfg.slots + parts - 1) / parts);
    if (cfg.drain_below > 0) pc.drain_below = std::max(1, cfg.drain_below / parts);
    for (int i = 0; i < parts; ++i) {
        if (i > 0 && !side[i]) WF_TRY(hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking));
        if (!ev_done[i]) WF_TRY(hipEventCreate(&ev_done[i]));
    }
    if (!ev_fork) WF_TRY(hipEventCreate(&ev_fork));
    // the parts start after everything the caller queued on the context's stream
    WF_TRY(hipEventRecord(ev_fork, stream));
    for (int i = 1; i < parts; ++i) WF_TRY(hipStreamWaitEvent(side[i], ev_fork, 0));
    std::vector<RunStats> rs(parts);
    std::vector<std::string> errs(parts);
    std::vector<int> rcs(parts, PCL_OK);
    // no exception may leave while a part's thread still runs (nor cross the C ABI)
    auto run_part = [&](int i) {
        try {
            rcs[i] = eng[i].run(i ? side[i] : stream, items, scenes, pc, cus, rs[i], errs[i], i, parts, ev_done[i]);
        } catch (const std::bad_alloc&) {
            rcs[i] = PCL_ERR_OOM;
            errs[i] = "host allocation failed";
        } catch (...) {
            rcs[i] = PCL_ERR_HIP;
            errs[i] = "host exception";
        }
    };
    std::vector<std::thread> th;
    std::vector<int> here;   // parts whose thread could not be started: run on this thread
    for (int i = 1; i < parts; ++i) {
        try {
            th.emplace_back([&, i]() {
                if (hipSetDevice(device) != hipSuccess) { rcs[i] = PCL_ERR_HIP; errs[i] = "hipSetDevice"; return; }
                run_part(i);
            });
        } catch (...) {
            here.push_back(i);
        }
    }
    run_part(0);
    for (int i : here) run_part(i);
    for (std::thread& t : th) t.join();
    for (int i = 0; i < parts; ++i)
        if (rcs[i] != PCL_OK) {
            // a part that failed may have left work queued: let every stream finish
            for (int k = 0; k < parts; ++k) (void)hipStreamSynchronize(k ? side[k] : stream);
            err = "part " + std::to_string(i) + ": " + errs[i];
            return rcs[i];
        }
    out = RunStats{};
    float ms = 0.0f;
    for (int i = 0; i < parts; ++i) {
        out.rays += rs[i].rays;
        out.closest += rs[i].closest;
        out.anyhit += rs[i].anyhit;
        out.iterations = std::max(out.iterations, rs[i].iterations);
        out.finished += rs[i].finished;
        out.deferred += rs[i].deferred;
        out.graph_launches += rs[i].graph_launches;
        float m = 0.0f;
        WF_TRY(hipEventElapsedTime(&m, ev_fork, ev_done[i]));
        ms = std::max(ms, m);
    }
    out.ms = ms;
    out.cfg = rs[0].cfg;
    out.cfg.slots = cfg.slots;
    out.cfg.drain_below = cfg.drain_below;
    out.cfg.streams = parts;
    return PCL_OK;
}

}  // namespace wf
}  // namespace pcl
