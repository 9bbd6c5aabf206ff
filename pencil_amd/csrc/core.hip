// core.hip - gfx950 kernels and the C ABI of include/pencil_hip.h.
//
// Kernels
//   k_trace_rays<ANY, COUNT> : one lane per ray, Scene::hit / Scene::occluded
//                              over the reference-topology BVH.
//   k_render                 : one lane per pixel (128-pixel row segment per block),
//                              n*n samples in order, Whitted trace() or
//                              trace_path() with NEE.
//   k_render_p               : the same per-pixel program on a persistent grid:
//                              each wave pulls 8x8 pixel tiles from
//                              striped counters (large images; see
//                              kPersistMinChunksPerWave).
// Both keep the traversal stack's top 16 entries per lane in LDS (16 x 8 B x
// 128 lanes = 16 KB per block), spilling deeper entries to a private array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "adjoint_launch.hpp"
#include "glass_sort.hpp"
#include "tonemap_launch.hpp"
#include "host_build.hpp"
#include "lbvh.hpp"
#include "pencil_hip.h"
#include "render.hpp"
#include "wavefront.hpp"

using namespace pcl;

namespace {

constexpr int kBlock = 128;
constexpr int kCtrSlots = 16;   // device counter words (>= CTR_N)

enum { CTR_RAYS = 0, CTR_CLOSEST, CTR_ANY, CTR_NODES, CTR_T0, CTR_N = CTR_T0 + 6 };

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Wave-level sums, one atomic per counter per wave.
template <bool COUNT>
__device__ __forceinline__ void flush_counts(const LaneCount<COUNT>& c, unsigned long long* ctr) {
    constexpr int N = COUNT ? CTR_N : CTR_NODES;
    uint32_t v[CTR_N] = {c.rays, c.closest, c.anyhit, c.trav.nodes, c.trav.t0, c.trav.t1,
                         c.trav.t2, c.trav.t3, c.trav.t4, c.trav.t5};
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (v[k]) atomicAdd(ctr + k, (unsigned long long)v[k]);
    }
}

template <bool ANY, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_trace_rays(DevScene S, const pcl_ray* __restrict__ rays, int n,
                                                       pcl_hit* __restrict__ hits, unsigned long long* ctr) {
    __shared__ unsigned long long lds[kStackK * kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    LaneCount<COUNT> c;
    if (i < n) {
        const float4* q = reinterpret_cast<const float4*>(rays + i);
        const float4 a = q[0], b = q[1];
        Ray r;
        r.o = f3(a.x, a.y, a.z);
        r.d = f3(a.w, b.x, b.y);
        Stack st;
        st.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
        st.set_stride(kBlock);
        float tb, ta;
        // an empty interval misses, as the reference's box test decides; a
        // NaN bound is taken as empty (the slab test's max/min would drop it)
        const int slot = (b.z <= b.w) ? traverse<ANY>(S, r, b.z, b.w, st, tb, ta, c.trav) : -1;
        pcl_hit h;
        if (ANY) {
            h.t = slot >= 0 ? 1.0f : -1.0f;
            h.prim_id = slot >= 0 ? 1 : -1;
        } else {
            h.t = slot >= 0 ? tb : -1.0f;
            h.prim_id = slot >= 0 ? gptr(S.prims)[slot].id : -1;
        }
        hits[i] = h;
    }
    if (COUNT) flush_counts(c, ctr);
}

#ifndef PCL_SHALLOW_WAVES
#define PCL_SHALLOW_WAVES 3
#endif
// Waves per SIMD the register budget allows: Whitted 3, path tracing 4, the
// frame-free max_depth <= 1 program PCL_SHALLOW_WAVES (with 8x8 tiles per
// wave, the 1080p 1-spp stress frame 0.838 ms wall at 4 waves, 0.819 at 3,
// 0.814 at 2, 0.876 at 5).
template <int PT>
struct RenderWaves { static constexpr int v = PT == 1 ? 4 : PT == 2 ? PCL_SHALLOW_WAVES : 3; };

// Buffers of the deferred glass vertices (render_pixel_defer, k_glass, k_sum).
struct DeferBufs {
    GlassSink sink;    // tasks, count, cap
    float* samp;       // sample k of pixel p at (k * n_px + p) * 3
    float* prefix;     // per pixel: sum of the samples before the first deferred one
    int32_t* first;    // per pixel: first deferred sample, n^2 when none
    int64_t n_px;
};

// The per-pixel body shared by both grids: render_pixel, or with DEFER its
// deferred-glass form (path tracing only), which leaves the pixel to k_sum
// when it deferred a sample.  DEFER: 0 = in place, 1 = deferred (Whitted
// draws nothing: hard lights), 2 = deferred with the subtree walk (soft
// lights).
template <int PT, int DEFER, class LC>
__device__ __forceinline__ void shade_pixel(const DevScene& S, const CamRec& C, const RSettings& st, int i, int j,
                                            Frame* F, Stack& stk, LC& c, float* out, DeferBufs& db) {
    const int64_t p = (int64_t)j * st.width + i;
    f3 px;
    if (DEFER) {
        int first;
        db.sink.pixel = (int32_t)p;
        if (!render_pixel_defer<DEFER == 2>(S, C, st, i, j, F, stk, c, db.sink, db.samp + p * 3,
                                            (size_t)db.n_px * 3, px, first)) {
            float* q = db.prefix + p * 3;
            q[0] = px.x; q[1] = px.y; q[2] = px.z;
            db.first[p] = first;
            return;
        }
        db.first[p] = first;
    } else {
        px = render_pixel<PT>(S, C, st, i, j, F, stk, c);
    }
    float* o = out + p * 3;
    o[0] = px.x; o[1] = px.y; o[2] = px.z;
}

// One instantiation per tracing mode (PT: Whitted / path), each with the
// register budget measured fastest for it: Whitted at most 168 VGPRs, 3 waves
// per SIMD (faster than 183 / 2 waves; 128 / 4 waves is no faster); path
// tracing 128 VGPRs, 4 waves (the 1M-triangle Cornell frame 1.19x faster
// than at 168 / 3 waves, and faster than 96 / 5 waves).
template <bool COUNT, int PT, int DEFER>
__global__ __launch_bounds__(kBlock, RenderWaves<PT>::v) void k_render(DevScene S, CamRec C, RSettings st, float* __restrict__ out,
                                                   unsigned long long* ctr, DeferBufs db) {
    __shared__ unsigned long long lds[kStackK * kBlock];
    // A block is 128 consecutive pixels of one row, blocks in row-major
    // order: measured 7% faster on the 1M-primitive frame and 16% on the
    // 64-scene batch than 8x8-pixel tiles per wave.  (Giving each of the 8
    // XCDs a contiguous band of the image, for L2 locality, measured 43%
    // slower: the bands' costs differ and the busiest XCD sets the time.)
    const int segs_x = (st.width + 127) / 128;
    const int i = (blockIdx.x % segs_x) * 128 + threadIdx.x;
    const int j = blockIdx.x / segs_x;
    LaneCount<COUNT> c;
    if (i < st.width && j < st.height) {
        // raw storage: a Frame array would run f3's zeroing constructors on
        // all 16 frames of every lane (~770 B of scratch writes per pixel)
        alignas(16) unsigned char fbuf[sizeof(Frame) * kMaxFrames];
        Frame* F = reinterpret_cast<Frame*>(fbuf);
        Stack stk;
        stk.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
        stk.set_stride(kBlock);
        shade_pixel<PT, DEFER>(S, C, st, i, j, F, stk, c, out, db);
    }
    flush_counts(c, ctr);
}

// Scratch memory of pcl_adjoint_device_noalloc, one copy per device (held
// until the process ends).  Layout: 16 counter words at offset 0, the
// reduced gradient + loss (PCL_ADJ_MAX_MATERIALS * 6 + 1 doubles) at
// kScrOut, the per-tile partials from kScrPartial on.  Host code never uses
// this name as an address: pcl_init looks up each device's copy with
// hipGetSymbolAddress and keeps it in the context.
constexpr size_t kScratchBytes = (size_t)PCL_NOALLOC_SCRATCH_BYTES;
constexpr size_t kScrOut = 256;
constexpr size_t kScrPartial = 4096;
static_assert(kScrOut + (PCL_ADJ_MAX_MATERIALS * 6 + 1) * sizeof(double) <= kScrPartial, "scratch layout");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Zeroes two word ranges in one launch (the render's counters and chunk heads).
__global__ void k_zero2_u64(unsigned long long* a, int na, unsigned long long* b, int nb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na) a[i] = 0ull;
    if (i < nb) b[i] = 0ull;
}

// ---- persistent render: waves pull 8x8-pixel chunks -----------------------
//
// Chunk c is the 8x8-pixel tile c of the image (row-major tiles).
// Chunks are dealt to kRStripes stripes round-robin (stripe g holds chunks g,
// g + 64, ...), each stripe with its own head counter on its own 128-byte
// line: device-scope atomics on one line serialize, so one counter for the
// whole frame would cost ~11 ns x 32,400 chunks.  A wave drains its home
// stripe first, then takes chunks from any stripe with work left.
constexpr int kRStripes = 64;
constexpr int kRLine = 32;        // uint32 words per 128-byte line
constexpr int kRHeadWords = kRStripes * kRLine;
// Measured on MI355X (kernel times, persistent vs one block per 128 pixels):
// 1080p 1 spp Whitted depth 1 (10.5 chunks per resident wave) 0.90 vs 1.29 ms;
// 1080p 16 spp path tracing (7.9) 218 vs 224 ms; 1024^2 16 spp Whitted
// depth 6 (5.3) 12.1 vs 11.9 ms; 512^2 64 spp path tracing (1.0) 35.1 vs
// 33.9 ms.  Short per-pixel programs gain most (block dispatch dominates).
constexpr int kPersistMinChunksPerWave = 6;

struct ChunkFetch {
    unsigned long long dead = 0;  // stripes known to be exhausted
    int home;                     // home stripe, or -1 once it is exhausted
    int rot;                      // scan rotation
    __device__ explicit ChunkFetch(int wave) : home(wave & (kRStripes - 1)), rot((wave * 7 + 1) & (kRStripes - 1)) {}
    __device__ static uint32_t len(int g, int n) { return n > g ? (uint32_t)((n - g + kRStripes - 1) / kRStripes) : 0u; }
    // Wave-uniform: the next chunk for this wave, or false when every stripe is
    // exhausted (heads only grow, so each stripe misses at most once).
    __device__ bool next(uint32_t* heads, int n, int& chunk) {
        const int lane = threadIdx.x & 63;
        if (home >= 0) {
            uint32_t k = 0;
            if (lane == 0) k = atomicAdd(&heads[home * kRLine], 1u);
            k = __shfl(k, 0, 64);
            if (k < len(home, n)) { chunk = (int)k * kRStripes + home; return true; }
            dead |= 1ull << home;
            home = -1;
        }
        while (true) {
            const bool has = __hip_atomic_load(&heads[lane * kRLine], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                             len(lane, n);
            const unsigned long long m = __ballot(has) & ~dead;
            if (!m) { dead = ~0ull; return false; }
            const unsigned long long r = rot ? ((m >> rot) | (m << (64 - rot))) : m;
            const int pick = (rot + __ffsll((long long)r) - 1) & (kRStripes - 1);
            uint32_t k = 0;
            if (lane == 0) k = atomicAdd(&heads[pick * kRLine], 1u);
            k = __shfl(k, 0, 64);
            if (k < len(pick, n)) { chunk = (int)k * kRStripes + pick; return true; }
            dead |= 1ull << pick;
        }
    }
};

// k_render with a persistent grid: same per-pixel program, an 8x8 pixel tile
// per wave (rays of a square footprint share more BVH nodes than a 64-pixel
// row's: the 1080p 1-spp stress frame 0.86 -> 0.82 ms); a wave takes its next
// tile as soon as its own lanes are done, independent of the other wave of
// its block.
template <bool COUNT, int PT, int DEFER>
__global__ __launch_bounds__(kBlock, RenderWaves<PT>::v) void k_render_p(DevScene S, CamRec C, RSettings st,
                                                                     float* __restrict__ out, unsigned long long* ctr,
                                                                     uint32_t* heads, int n_chunks, DeferBufs db) {
    __shared__ unsigned long long lds[kStackK * kBlock];
    const int tiles_x = (st.width + 7) / 8;
    const int lane = threadIdx.x & 63;
    LaneCount<COUNT> c;
    alignas(16) unsigned char fbuf[sizeof(Frame) * kMaxFrames];   // raw storage: no zeroing constructors
    Frame* F = reinterpret_cast<Frame*>(fbuf);
    Stack stk;
    stk.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
    stk.set_stride(kBlock);
    ChunkFetch f((int)((blockIdx.x * kBlock + threadIdx.x) >> 6));
    int chunk;
    while (f.next(heads, n_chunks, chunk)) {
        const int ty = chunk / tiles_x;
        const int i = (chunk - ty * tiles_x) * 8 + (lane & 7), j = ty * 8 + (lane >> 3);
        if (i < st.width && j < st.height) shade_pixel<PT, DEFER>(S, C, st, i, j, F, stk, c, out, db);
    }
    flush_counts(c, ctr);
}

// The deferred glass vertices: one lane per task (grid-stride over 64-task
// chunks), glass_task's subtree and sample value into the sample's slot.
// 4 waves per SIMD (128 VGPRs): the soft-light Cornell + mesh frame 3 %
// faster than at 3 (156 VGPRs); at 5 (96 VGPRs) 4 % slower.
template <bool COUNT>
__global__ __launch_bounds__(kBlock, 4) void k_glass(DevScene S, RSettings st, const float4* __restrict__ tasks,
                                                     const uint32_t* count, uint32_t cap, float* samp, int64_t n_px,
                                                     unsigned long long* ctr, const uint32_t* __restrict__ perm) {
    __shared__ unsigned long long lds[kStackK * kBlock];
    const uint32_t n = min(*count, cap);
    LaneCount<COUNT> c;
    alignas(16) unsigned char fbuf[sizeof(Frame) * kMaxFrames];
    Frame* F = reinterpret_cast<Frame*>(fbuf);
    Stack stk;
    stk.lds = (PCL_LDS unsigned long long*)(lds + threadIdx.x);
    stk.set_stride(kBlock);
    const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) >> 6, waves = (gridDim.x * kBlock) >> 6;
    for (uint32_t base = wave * 64u; base < n; base += waves * 64u) {
        const uint32_t i = base + (threadIdx.x & 63);
        if (i >= n) continue;
        const uint32_t t = perm ? perm[i] : i;
        const float4* q = tasks + (size_t)t * 8;
        const f3 v = glass_task(S, st, q, F, stk, c);
        const int64_t p = __float_as_int(q[4].w), k = __float_as_int(q[5].w);
        float* o = samp + (k * n_px + p) * 3;
        o[0] = v.x; o[1] = v.y; o[2] = v.z;
    }
    flush_counts(c, ctr);
}

// Pixels that deferred a sample: the prefix plus samples first..n^2-1 in
// order, divided by n^2 -- render_pixel's sum, same additions, same order.
__global__ __launch_bounds__(256) void k_sum(float* __restrict__ out, const float* __restrict__ samp,
                                             const float* __restrict__ prefix, const int32_t* __restrict__ first,
                                             int64_t n_px, int nn) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_px) return;
    const int k0 = first[p];
    if (k0 >= nn) return;
    f3 acc(prefix[p * 3], prefix[p * 3 + 1], prefix[p * 3 + 2]);
    for (int k = k0; k < nn; ++k) {
        const float* s = samp + ((int64_t)k * n_px + p) * 3;
        acc += f3(s[0], s[1], s[2]);
    }
    const f3 px = acc / (float)nn;
    out[p * 3] = px.x; out[p * 3 + 1] = px.y; out[p * 3 + 2] = px.z;
}

struct SceneSlot {
    bool used = false;
    DevScene dev{};
    std::vector<void*> bufs;
    std::vector<int32_t> leaf_order;
    std::vector<LightRec> lights;   // host copy, for per-item sample budgets
    int32_t n_mats = 0;
    bool has_glass = false;         // some material is a dielectric
    pcl_scene_info info{};
};

}  // namespace

struct pcl_ctx {
    int device = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned long long* d_ctr = nullptr;
    // pinned host copy of a render's counters (kCtrSlots words) and its
    // deferred-task count: a copy into pageable memory goes through a staging
    // buffer on every frame
    unsigned long long* h_ctr = nullptr;
    std::string err;
    std::vector<SceneSlot> scenes;
    pcl::wf::Tracer* wf = nullptr;
    pcl::wf::Config wf_default;      // set by pcl_wf_autotune
    // adjoint work buffers, grown on demand
    double* d_adj_partial = nullptr;
    size_t adj_partial_cap = 0;
    double* d_adj_out = nullptr;     // kAdjMaxMats*6 + 1 doubles
    float* d_img = nullptr;          // forward image / target staging
    float* d_tgt = nullptr;
    size_t img_cap = 0;
    float* d_thr = nullptr;          // tone-map threshold table (256 floats)
    uint8_t* d_rgb8 = nullptr;
    size_t rgb8_cap = 0;
    unsigned char* d_scratch = nullptr;   // this device's copy of `scratch`
    uint32_t* d_heads = nullptr;          // chunk heads of the persistent render (kRHeadWords)
    // render grid: -1 automatic (persistent when the image has at least
    // kPersistMinChunksPerWave 8x8 tiles per resident wave), 0 one block
    // per 128 pixels, 1 persistent.  Environment PCL_RENDER_KERNEL overrides.
    int render_kernel = -1;
    int render_bpc[3] = {0, 0, 0};        // persistent blocks per CU (Whitted, path, depth <= 1); 0 = occupancy
    // deferred glass vertices of path-traced renders (render_pixel_defer),
    // grown on demand: tasks, per-sample values, per-pixel prefix / first
    float4* d_gtasks = nullptr;
    size_t gtask_cap = 0;
    uint32_t* d_gkeys = nullptr;          // per task: its subtree's shaded nodes (walked renders)
    uint32_t* d_gperm = nullptr;          // tasks in the glass launch's order
    size_t gkey_cap = 0;
    uint32_t* d_gbins = nullptr;          // 256 counts + 256 offsets of the task sort
    uint32_t* d_gcount = nullptr;         // one 128-byte line
    float* d_samp = nullptr;
    size_t samp_cap = 0;                  // floats
    float* d_prefix = nullptr;
    int32_t* d_first = nullptr;
    size_t px_cap = 0;
    // -1 automatic (defer when the scene allows it), 0 never, 1 whenever
    // allowed.  Environment PCL_DEFER_GLASS overrides.
    int defer_glass = -1;
    int64_t defer_tasks = 0;              // task array size override (PCL_DEFER_TASKS); 0 = automatic
};

namespace {

int fail(pcl_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIP_TRY(ctx, call)                                                                          \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail((ctx), e_ == hipErrorOutOfMemory ? PCL_ERR_OOM : PCL_ERR_HIP,               \
                        std::string(#call) + ": " + hipGetErrorString(e_));                         \
    } while (0)

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

SceneSlot* get_scene(pcl_ctx* ctx, int32_t s) {
    if (!ctx || s < 0 || s >= (int32_t)ctx->scenes.size() || !ctx->scenes[s].used) return nullptr;
    return &ctx->scenes[s];
}

template <class T>
int upload(pcl_ctx* ctx, SceneSlot& sl, const std::vector<T>& v, const T** out) {
    *out = nullptr;
    if (v.empty()) return PCL_OK;
    void* p = nullptr;
    HIP_TRY(ctx, hipMalloc(&p, v.size() * sizeof(T)));
    sl.bufs.push_back(p);
    HIP_TRY(ctx, hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    *out = static_cast<const T*>(p);
    sl.info.device_bytes += v.size() * sizeof(T);
    return PCL_OK;
}

void release(SceneSlot& sl) {
    for (void* p : sl.bufs) (void)hipFree(p);
    sl.bufs.clear();
    sl.used = false;
}

int check_settings(pcl_ctx* ctx, const pcl_settings* s) {
    if (!s || s->width <= 0 || s->height <= 0 || s->samples <= 0)
        return fail(ctx, PCL_ERR_INVALID, "settings: width, height and samples must be positive");
    if ((int64_t)s->width * s->height > (int64_t)1 << 28)
        return fail(ctx, PCL_ERR_LIMIT, "settings: image larger than 2^28 pixels");
    if (!s->pathtrace && s->max_depth > kMaxFrames)
        return fail(ctx, PCL_ERR_LIMIT, "settings: max_depth above " + std::to_string(kMaxFrames));
    if (s->pathtrace && s->pathtrace_bounces > kMaxFrames)
        return fail(ctx, PCL_ERR_LIMIT, "settings: pathtrace_bounces above " + std::to_string(kMaxFrames));
    if (s->soft_shadow_samples < 0) return fail(ctx, PCL_ERR_INVALID, "settings: soft_shadow_samples < 0");
    if (!std::isfinite(s->shadow_bias)) return fail(ctx, PCL_ERR_INVALID, "settings: shadow_bias must be finite");
    if (s->pathtrace < 0 || s->pathtrace > 2)
        return fail(ctx, PCL_ERR_INVALID, "settings: pathtrace must be 0 (Whitted), 1 (path) or 2 (path + MIS)");
    return PCL_OK;
}

RSettings to_dev(const pcl_settings* s) {
    RSettings r;
    r.width = s->width; r.height = s->height; r.samples = s->samples; r.max_depth = s->max_depth;
    r.shadow_bias = s->shadow_bias; r.soft_samples = s->soft_shadow_samples;
    r.pathtrace = s->pathtrace; r.bounces = s->pathtrace_bounces;
    return r;
}

void fill_stats(pcl_stats* st, const unsigned long long* h, float kms, double t0) {
    if (!st) return;
    st->rays_traced = h[CTR_RAYS];
    st->closest_queries = h[CTR_CLOSEST];
    st->anyhit_queries = h[CTR_ANY];
    st->kernel_ms = kms;
    st->total_ms = now_ms() - t0;
    st->deferred_glass = 0;
}

}  // namespace

// Host allocations (and host threads) can throw; no exception crosses the C
// ABI: an entry point's function-try-block turns one into an error code.
#define PCL_CATCH(ctx)                                                                                   \
    catch (const std::bad_alloc&) { return fail((ctx), PCL_ERR_OOM, "host allocation failed"); }          \
    catch (const std::exception& e_) { return fail((ctx), PCL_ERR_HIP, std::string("host: ") + e_.what()); } \
    catch (...) { return fail((ctx), PCL_ERR_HIP, "host exception"); }

extern "C" {

int pcl_abi_version(void) { return PCL_ABI_VERSION; }

int pcl_init(int device, pcl_ctx** out) try {
    if (!out) return PCL_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PCL_ERR_HIP;
    if (device < 0 || device >= n) return PCL_ERR_INVALID;
    auto ctx = std::make_unique<pcl_ctx>();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) return PCL_ERR_HIP;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return PCL_ERR_HIP;
    if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) return PCL_ERR_HIP;
    if (hipMalloc(&ctx->d_ctr, kCtrSlots * sizeof(unsigned long long)) != hipSuccess) return PCL_ERR_OOM;
    {
        float thr[256];
        pcl::tonemap_thresholds(thr);
        if (hipMalloc(&ctx->d_thr, sizeof(thr)) != hipSuccess) return PCL_ERR_OOM;
        if (hipMemcpy(ctx->d_thr, thr, sizeof(thr), hipMemcpyHostToDevice) != hipSuccess) return PCL_ERR_HIP;
    }
    if (hipMalloc(&ctx->d_heads, kRHeadWords * sizeof(uint32_t)) != hipSuccess) return PCL_ERR_OOM;
    if (const char* e = std::getenv("PCL_RENDER_KERNEL")) ctx->render_kernel = std::atoi(e);
    if (const char* e = std::getenv("PCL_DEFER_GLASS")) ctx->defer_glass = std::atoi(e);
    if (const char* e = std::getenv("PCL_DEFER_TASKS")) ctx->defer_tasks = std::max<int64_t>(1, std::atoll(e));
    if (hipMalloc(&ctx->d_gcount, 128) != hipSuccess) return PCL_ERR_OOM;
    if (const char* e = std::getenv("PCL_RENDER_BPC"))
        ctx->render_bpc[0] = ctx->render_bpc[1] = ctx->render_bpc[2] = std::atoi(e);
    // the device address of this device's scratch array (device is current)
    if (hipGetSymbolAddress((void**)&ctx->d_scratch, HIP_SYMBOL(scratch)) != hipSuccess || !ctx->d_scratch)
        return PCL_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->cus = prop.multiProcessorCount;
    ctx->wf = new pcl::wf::Tracer();
    ctx->wf->defer_glass = ctx->defer_glass;
    ctx->wf->defer_tasks = ctx->defer_tasks;
    *out = ctx.release();
    return PCL_OK;
} PCL_CATCH(nullptr)

void pcl_destroy(pcl_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& s : ctx->scenes) if (s.used) release(s);
    delete ctx->wf;
    if (ctx->d_ctr) (void)hipFree(ctx->d_ctr);
    if (ctx->h_ctr) (void)hipHostFree(ctx->h_ctr);
    if (ctx->d_adj_partial) (void)hipFree(ctx->d_adj_partial);
    if (ctx->d_adj_out) (void)hipFree(ctx->d_adj_out);
    if (ctx->d_img) (void)hipFree(ctx->d_img);
    if (ctx->d_tgt) (void)hipFree(ctx->d_tgt);
    if (ctx->d_thr) (void)hipFree(ctx->d_thr);
    if (ctx->d_rgb8) (void)hipFree(ctx->d_rgb8);
    if (ctx->d_heads) (void)hipFree(ctx->d_heads);
    for (void* p : {(void*)ctx->d_gkeys, (void*)ctx->d_gperm, (void*)ctx->d_gbins})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)ctx->d_gtasks, (void*)ctx->d_gcount, (void*)ctx->d_samp, (void*)ctx->d_prefix,
                    (void*)ctx->d_first})
        if (p) (void)hipFree(p);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* pcl_last_error(const pcl_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// use_bvh == 2: the records go up in insertion order with their boxes, the
// device builds the tree (lbvh.hip) and the leaf order, root box and depth
// come back.  Work arrays are freed before returning.
static int build_on_device(pcl_ctx* ctx, SceneSlot& sl, HostLayout& L) {
    const int n = (int)L.prims.size();
    DevScene& D = sl.dev;
    auto keep = [&](size_t bytes, void** p) -> int {
        HIP_TRY(ctx, hipMalloc(p, bytes));
        sl.bufs.push_back(*p);
        sl.info.device_bytes += bytes;
        return PCL_OK;
    };
    void *prims = nullptr, *wide = nullptr, *slot_of = nullptr;
    int rc;
    if ((rc = keep(sizeof(PrimRec) * n, &prims)) || (rc = keep(sizeof(int32_t) * n, &slot_of))) return rc;
    std::vector<void*> tmp;
    auto scratch = [&](size_t bytes, void** p) -> int {
        HIP_TRY(ctx, hipMalloc(p, bytes < 256 ? 256 : bytes));
        tmp.push_back(*p);
        return PCL_OK;
    };
    auto free_tmp = [&]() { for (void* p : tmp) (void)hipFree(p); };
    void *in = nullptr, *boxes = nullptr, *order = nullptr, *ka = nullptr, *kb = nullptr, *pi = nullptr,
         *pl = nullptr, *nd = nullptr, *sorted = nullptr, *out16 = nullptr, *work = nullptr, *nodes = nullptr,
         *flag = nullptr, *widx = nullptr;
    const size_t work_bytes = pcl::lbvh::temp_bytes(n);
    if ((rc = scratch(sizeof(NodeRec) * (n - 1), &nodes)) || (rc = scratch(sizeof(int32_t) * n, &flag)) ||
        (rc = scratch(sizeof(int32_t) * n, &widx)) ||
        (rc = scratch(sizeof(PrimRec) * n, &in)) || (rc = scratch(sizeof(float) * 6 * n, &boxes)) ||
        (rc = scratch(sizeof(int32_t) * n, &order)) || (rc = scratch(sizeof(uint64_t) * n, &ka)) ||
        (rc = scratch(sizeof(uint64_t) * n, &kb)) || (rc = scratch(sizeof(int32_t) * n, &pi)) ||
        (rc = scratch(sizeof(int32_t) * n, &pl)) || (rc = scratch(sizeof(int32_t) * n, &nd)) ||
        (rc = scratch(sizeof(float) * 6 * n, &sorted)) || (rc = scratch(sizeof(float) * 16, &out16)) ||
        (rc = scratch(work_bytes, &work))) {
        free_tmp();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(in, L.prims.data(), sizeof(PrimRec) * n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(boxes, L.boxes.data(), sizeof(float) * 6 * n, hipMemcpyHostToDevice, ctx->stream);
    int depth = 0;
    if (e == hipSuccess)
        e = pcl::lbvh::build(ctx->stream, (const PrimRec*)in, (const float*)boxes, n, (PrimRec*)prims,
                             (NodeRec*)nodes, (int32_t*)order, (int32_t*)slot_of, work, work_bytes, (uint64_t*)ka,
                             (uint64_t*)kb, (int32_t*)pi, (int32_t*)pl, (int32_t*)nd, (float*)sorted,
                             (float*)out16, &depth);
    float h6[6] = {0};
    if (e == hipSuccess) e = hipMemcpyAsync(h6, out16, sizeof(h6), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(L.leaf_order.data(), order, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream);
    int n_wide = 0;
    if (e == hipSuccess)
        e = pcl::lbvh::wide_count(ctx->stream, (const int32_t*)nd, n, (int32_t*)flag, (int32_t*)widx, work,
                                  work_bytes, &n_wide);
    if (e == hipSuccess && (n_wide <= 0 || n_wide > n - 1)) e = hipErrorUnknown;
    if (e == hipSuccess && (rc = keep(sizeof(Node4) * n_wide, &wide))) { free_tmp(); return rc; }
    if (e == hipSuccess)
        e = pcl::lbvh::wide_emit(ctx->stream, (const NodeRec*)nodes, n, (const int32_t*)nd, (const int32_t*)widx,
                                 (Node4*)wide);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    free_tmp();
    if (e != hipSuccess) return fail(ctx, PCL_ERR_HIP, std::string("device BVH build: ") + hipGetErrorString(e));
    for (int k = 0; k < 6; ++k) L.root_box[k] = h6[k];
    L.depth = depth;
    L.root = 0;
    L.wide_root = 0;
    // even-depth nodes on a path: at most depth / 2 rounded up, each pushing <= 3
    L.stack_need = 3 * ((depth + 1) / 2);
    L.n_wide = n_wide;
    D.prims = (const PrimRec*)prims;
    D.nodes = (const Node4*)wide;
    D.slot_of = (const int32_t*)slot_of;
    return PCL_OK;
}

int pcl_scene_upload(pcl_ctx* ctx, const pcl_scene_desc* desc, int32_t* scene_out) try {
    if (!ctx || !desc || !scene_out) return fail(ctx, PCL_ERR_INVALID, "null argument");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const double t0 = now_ms();
    HostLayout L;
    std::string err;
    int rc = build_layout(desc, L, err);
    if (rc != PCL_OK) return fail(ctx, rc, err);
    const double t1 = now_ms();
    int32_t idx = -1;
    for (size_t i = 0; i < ctx->scenes.size(); ++i)
        if (!ctx->scenes[i].used) { idx = (int32_t)i; break; }
    if (idx < 0) { idx = (int32_t)ctx->scenes.size(); ctx->scenes.emplace_back(); }
    SceneSlot& sl = ctx->scenes[idx];
    sl = SceneSlot{};
    sl.used = true;
    DevScene& D = sl.dev;
    int r = PCL_OK;
    const bool device_build = desc->use_bvh == 2 && desc->n_prims > 1;
    if (device_build) {
        if ((r = build_on_device(ctx, sl, L)) || (r = upload(ctx, sl, L.mats, &D.mats)) ||
            (r = upload(ctx, sl, L.lights, &D.lights))) {
            release(sl);
            return r;
        }
    } else if ((r = upload(ctx, sl, L.wide, &D.nodes)) || (r = upload(ctx, sl, L.prims, &D.prims)) ||
               (r = upload(ctx, sl, L.mats, &D.mats)) || (r = upload(ctx, sl, L.lights, &D.lights)) ||
               (r = upload(ctx, sl, L.slot_of, &D.slot_of))) {
        release(sl);
        return r;
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { release(sl); return fail(ctx, PCL_ERR_HIP, hipGetErrorString(e)); }
    // the traversal addresses nodes with 32-bit byte offsets (intersect.hpp)
    const int64_t n_wide = device_build ? L.n_wide : (int64_t)L.wide.size();
    if (n_wide * (int64_t)sizeof(Node4) > ((int64_t)1 << 32)) {
        release(sl);
        return fail(ctx, PCL_ERR_LIMIT, std::to_string(n_wide) + " 4-wide BVH nodes: more than the 4 GiB the "
                                        "traversal addresses");
    }
    if (L.stack_need > kStackK + kStackSpill) {
        release(sl);
        return fail(ctx, PCL_ERR_LIMIT, "BVH traversal needs " + std::to_string(L.stack_need) +
                                            " stack entries, more than the traversal stack (" +
                                            std::to_string(kStackK + kStackSpill) + ")");
    }
    D.n_prims = desc->n_prims;
    D.n_lights = desc->n_lights;
    D.root = L.wide_root;
    D.use_bvh = desc->use_bvh ? 1 : 0;   // 1 and 2 both leave a BVH on the device
    for (int k = 0; k < 6; ++k) D.root_box[k] = L.root_box[k];
    for (int k = 0; k < 3; ++k) { D.bg[k] = desc->background[k]; D.amb[k] = desc->ambient[k]; }
    sl.leaf_order = std::move(L.leaf_order);
    sl.lights = L.lights;
    sl.n_mats = desc->n_materials;
    for (int32_t m = 0; m < desc->n_materials; ++m) sl.has_glass |= desc->materials[m].kind == PCL_MAT_DIELECTRIC;
    sl.info.n_prims = desc->n_prims;
    sl.info.n_nodes = device_build ? desc->n_prims - 1 : (int32_t)L.nodes.size();
    sl.info.depth = L.depth;
    sl.info.root = L.root;
    sl.info.n_wide_nodes = device_build ? L.n_wide : (int32_t)L.wide.size();
    sl.info.stack_need = L.stack_need;
    sl.info.build_ms = t1 - t0;
    sl.info.upload_ms = now_ms() - t1;
    *scene_out = idx;
    return PCL_OK;
} PCL_CATCH(ctx)

int pcl_scene_free(pcl_ctx* ctx, int32_t scene) {
    SceneSlot* sl = get_scene(ctx, scene);
    if (!sl) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    release(*sl);
    return PCL_OK;
}

int pcl_scene_info_get(pcl_ctx* ctx, int32_t scene, pcl_scene_info* info) {
    SceneSlot* sl = get_scene(ctx, scene);
    if (!sl || !info) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    *info = sl->info;
    return PCL_OK;
}

int pcl_scene_leaf_order(pcl_ctx* ctx, int32_t scene, int32_t* out, int32_t n) try {
    SceneSlot* sl = get_scene(ctx, scene);
    if (!sl || !out) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    if (n < (int32_t)sl->leaf_order.size()) return fail(ctx, PCL_ERR_INVALID, "output too small");
    if (!sl->leaf_order.empty()) std::memcpy(out, sl->leaf_order.data(), sl->leaf_order.size() * sizeof(int32_t));
    return PCL_OK;
} PCL_CATCH(ctx)

int pcl_bvh_build_host(const pcl_scene_desc* desc, pcl_scene_info* info, int32_t* leaf_order, int32_t n) try {
    if (!desc) return PCL_ERR_INVALID;
    const double t0 = now_ms();
    HostLayout L;
    std::string err;
    int rc = build_layout(desc, L, err);
    if (rc != PCL_OK) return rc;
    if (info) {
        *info = pcl_scene_info{};
        info->n_prims = desc->n_prims;
        info->n_nodes = (int32_t)L.nodes.size();
        info->depth = L.depth;
        info->root = L.root;
        info->n_wide_nodes = (int32_t)L.wide.size();
        info->stack_need = L.stack_need;
        info->device_bytes = L.wide.size() * sizeof(Node4) + L.prims.size() * sizeof(PrimRec) +
                             L.mats.size() * sizeof(MatRec) + L.lights.size() * sizeof(LightRec) +
                             L.slot_of.size() * sizeof(int32_t);
        info->build_ms = now_ms() - t0;
    }
    if (leaf_order) {
        if (n < (int32_t)L.leaf_order.size()) return PCL_ERR_INVALID;
        if (!L.leaf_order.empty()) std::memcpy(leaf_order, L.leaf_order.data(), L.leaf_order.size() * sizeof(int32_t));
    }
    return PCL_OK;
} PCL_CATCH(nullptr)

static int launch_trace(pcl_ctx* ctx, SceneSlot* sl, const pcl_ray* d_rays, int32_t n, int32_t any_hit,
                        pcl_hit* d_hits, bool count) {
    const int grid = (n + kBlock - 1) / kBlock;
    if (any_hit) {
        if (count) hipLaunchKernelGGL((k_trace_rays<true, true>), dim3(grid), dim3(kBlock), 0, ctx->stream, sl->dev, d_rays, n, d_hits, ctx->d_ctr);
        else hipLaunchKernelGGL((k_trace_rays<true, false>), dim3(grid), dim3(kBlock), 0, ctx->stream, sl->dev, d_rays, n, d_hits, ctx->d_ctr);
    } else {
        if (count) hipLaunchKernelGGL((k_trace_rays<false, true>), dim3(grid), dim3(kBlock), 0, ctx->stream, sl->dev, d_rays, n, d_hits, ctx->d_ctr);
        else hipLaunchKernelGGL((k_trace_rays<false, false>), dim3(grid), dim3(kBlock), 0, ctx->stream, sl->dev, d_rays, n, d_hits, ctx->d_ctr);
    }
    HIP_TRY(ctx, hipGetLastError());
    return PCL_OK;
}

int pcl_trace_rays_device(pcl_ctx* ctx, int32_t scene, const pcl_ray* d_rays, int32_t n, int32_t any_hit,
                          pcl_hit* d_hits, pcl_stats* stats) try {
    SceneSlot* sl = get_scene(ctx, scene);
    if (!sl) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    if (n < 0 || (n > 0 && (!d_rays || !d_hits))) return fail(ctx, PCL_ERR_INVALID, "bad ray buffers");
    const double t0 = now_ms();
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    unsigned long long h[kCtrSlots] = {0};
    float kms = 0.0f;
    if (n > 0) {
        HIP_TRY(ctx, hipMemsetAsync(ctx->d_ctr, 0, sizeof(h), ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        int rc = launch_trace(ctx, sl, d_rays, n, any_hit, d_hits, false);
        if (rc) return rc;
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
        HIP_TRY(ctx, hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    }
    if (any_hit) h[CTR_ANY] = (unsigned long long)n; else h[CTR_CLOSEST] = (unsigned long long)n;
    fill_stats(stats, h, kms, t0);
    return PCL_OK;
} PCL_CATCH(ctx)

int pcl_trace_rays(pcl_ctx* ctx, int32_t scene, const pcl_ray* rays, int32_t n, int32_t any_hit, pcl_hit* hits,
                   pcl_stats* stats) try {
    SceneSlot* sl = get_scene(ctx, scene);
    if (!sl) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    if (n < 0 || (n > 0 && (!rays || !hits))) return fail(ctx, PCL_ERR_INVALID, "bad ray buffers");
    if (n == 0) return pcl_trace_rays_device(ctx, scene, nullptr, 0, any_hit, nullptr, stats);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    void *dr = nullptr, *dh = nullptr;
    HIP_TRY(ctx, hipMalloc(&dr, (size_t)n * sizeof(pcl_ray)));
    if (hipMalloc(&dh, (size_t)n * sizeof(pcl_hit)) != hipSuccess) { (void)hipFree(dr); return fail(ctx, PCL_ERR_OOM, "hits"); }
    int rc = PCL_OK;
    if (hipMemcpy(dr, rays, (size_t)n * sizeof(pcl_ray), hipMemcpyHostToDevice) != hipSuccess) rc = fail(ctx, PCL_ERR_HIP, "copy rays");
    if (!rc) rc = pcl_trace_rays_device(ctx, scene, (const pcl_ray*)dr, n, any_hit, (pcl_hit*)dh, stats);
    if (!rc && hipMemcpy(hits, dh, (size_t)n * sizeof(pcl_hit), hipMemcpyDeviceToHost) != hipSuccess) rc = fail(ctx, PCL_ERR_HIP, "copy hits");
    (void)hipFree(dr);
    (void)hipFree(dh);
    return rc;
} PCL_CATCH(ctx)

// Work counters for a ray batch: internal nodes entered and primitive tests,
// summed over the batch (a separate instrumented launch).
static void fill_counts(pcl_work_counts* out, const unsigned long long* h) {
    if (!out) return;
    out->node_visits = h[CTR_NODES];
    for (int k = 0; k < 6; ++k) out->prim_tests[k] = h[CTR_T0 + k];
}

int pcl_trace_rays_counts(pcl_ctx* ctx, int32_t scene, const pcl_ray* d_rays, int32_t n, int32_t any_hit,
                          pcl_hit* d_hits, pcl_work_counts* counts) try {
    SceneSlot* sl = get_scene(ctx, scene);
    if (!sl) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    if (n <= 0 || !d_rays || !d_hits) return fail(ctx, PCL_ERR_INVALID, "bad ray buffers");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    unsigned long long h[kCtrSlots] = {0};
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_ctr, 0, sizeof(h), ctx->stream));
    int rc = launch_trace(ctx, sl, d_rays, n, any_hit, d_hits, true);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(h, ctx->d_ctr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    fill_counts(counts, h);
    return PCL_OK;
} PCL_CATCH(ctx)

constexpr int kMaxWalkDraws = 1 << 15;   // per shaded node, for deferral with the walk

// Generator draws of one shaded Whitted node (whitted_local): a unit_disc
// (2 draws) per sample of every point light with a radius.  Passed to the
// deferred programs (GlassSink::draws, ItemRec::draws).
static int whitted_draws_per_node(const SceneSlot& sl, const RSettings& st) {
    int64_t d = 0;
    for (const LightRec& L : sl.lights)
        if (L.kind == PCL_LIGHT_POINT && L.radius > 0.0f) d += 2 * (int64_t)st.soft_samples;
    return (int)std::min<int64_t>(d, INT32_MAX);
}

// Deferred glass vertices (render_pixel_defer): path tracing, a scene with a
// dielectric, and buffers that fit: the per-sample values of the frame (12 B
// per pixel and sample) and a task array sized for an eighth of the samples
// (128 B each; tasks beyond it are evaluated in place), together under half
// of the device's free memory.  Returns false (render in place) otherwise.
static bool defer_buffers(pcl_ctx* ctx, const SceneSlot& sl, const RSettings& st) {
    if (!st.pathtrace || !sl.has_glass || ctx->defer_glass == 0) return false;
    // a subtree's draws (draws per node x at most 2^16 nodes) stay below 2^31
    if (whitted_draws_per_node(sl, st) > kMaxWalkDraws) return false;
    const int64_t n_px = (int64_t)st.width * st.height, nn = (int64_t)st.samples * st.samples;
    if (n_px * nn >= ((int64_t)1 << 32)) return false;   // the task counter is 32-bit
    const size_t samp = (size_t)(n_px * nn * 3);
    const size_t tasks = ctx->defer_tasks > 0 ? (size_t)ctx->defer_tasks
                                              : (size_t)std::min<int64_t>(n_px * nn / 8 + 4096, (int64_t)1 << 27);
    size_t need = 0;
    if (samp > ctx->samp_cap) need += samp * sizeof(float);
    if (tasks > ctx->gtask_cap) need += tasks * (8 * sizeof(float4) + 2 * sizeof(uint32_t));
    if ((size_t)n_px > ctx->px_cap) need += (size_t)n_px * 16;
    if (need) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || need > free_b / 2) return false;
    }
    auto grow = [](auto** p, size_t& cap, size_t want, size_t elem) -> bool {
        if (want <= cap) return true;
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        cap = 0;
        if (hipMalloc((void**)p, want * elem) != hipSuccess) { *p = nullptr; return false; }
        cap = want;
        return true;
    };
    size_t pxc = ctx->px_cap;
    if (!grow(&ctx->d_samp, ctx->samp_cap, samp, sizeof(float)) ||
        !grow(&ctx->d_gtasks, ctx->gtask_cap, tasks, 8 * sizeof(float4)) ||
        !grow(&ctx->d_prefix, pxc, (size_t)n_px, 3 * sizeof(float)))
        return false;
    size_t pxc2 = ctx->px_cap;
    if (!grow(&ctx->d_first, pxc2, (size_t)n_px, sizeof(int32_t))) return false;
    ctx->px_cap = std::min(pxc, pxc2);
    size_t kc = ctx->gkey_cap, kc2 = ctx->gkey_cap, bc = ctx->d_gbins ? 512 : 0;
    const bool ok = grow(&ctx->d_gkeys, kc, tasks, sizeof(uint32_t)) && grow(&ctx->d_gperm, kc2, tasks, sizeof(uint32_t)) &&
                    grow(&ctx->d_gbins, bc, 512, sizeof(uint32_t));
    ctx->gkey_cap = std::min(kc, kc2);
    return ok;
}

static int render_impl(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                       float* d_out_rgb, pcl_stats* stats, pcl_work_counts* counts) {
    SceneSlot* sl = get_scene(ctx, scene);
    if (!sl) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    if (!cam || !d_out_rgb) return fail(ctx, PCL_ERR_INVALID, "null camera or output");
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    const double t0 = now_ms();
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const CamRec C = make_camera(*cam);
    const RSettings st = to_dev(settings);
    const int tiles = ((st.width + 127) / 128) * st.height;   // 128-pixel row segments
    if (!ctx->h_ctr) HIP_TRY(ctx, hipHostMalloc((void**)&ctx->h_ctr, (kCtrSlots + 1) * sizeof(unsigned long long),
                                                hipHostMallocDefault));
    unsigned long long* h = ctx->h_ctr;
    const int n_chunks = ((st.width + 7) / 8) * ((st.height + 7) / 8);   // 8x8 tiles
    // Whitted with max_depth <= 1 has no recursion: the frame-free program
    const int mode = st.pathtrace ? 1 : (st.max_depth <= 1 ? 2 : 0);
    const bool defer = mode == 1 && defer_buffers(ctx, *sl, st);
    DeferBufs db{};
    if (defer) {
        db.sink.tasks = ctx->d_gtasks;
        db.sink.count = ctx->d_gcount;
        // keys and tasks share the capacity (either array may have grown alone before a failed call)
        const size_t tcap = std::min(ctx->gtask_cap, ctx->gkey_cap);
        db.sink.cap = (uint32_t)std::min<size_t>(tcap, ctx->defer_tasks > 0 ? (size_t)ctx->defer_tasks : tcap);
        db.sink.draws = whitted_draws_per_node(*sl, st);
        db.sink.keys = ctx->d_gkeys;   // the glass launch's order
        db.samp = ctx->d_samp;
        db.prefix = ctx->d_prefix;
        db.first = ctx->d_first;
        db.n_px = (int64_t)st.width * st.height;
    }
    // deferred with the subtree walk only where Whitted draws (soft lights)
    const bool walk = defer && db.sink.draws > 0;
    auto pkern = mode == 1 ? (walk ? (counts ? k_render_p<true, 1, 2> : k_render_p<false, 1, 2>)
                              : defer ? (counts ? k_render_p<true, 1, 1> : k_render_p<false, 1, 1>)
                                      : (counts ? k_render_p<true, 1, 0> : k_render_p<false, 1, 0>))
               : mode == 2 ? (counts ? k_render_p<true, 2, 0> : k_render_p<false, 2, 0>)
                           : (counts ? k_render_p<true, 0, 0> : k_render_p<false, 0, 0>);
    int& bpc = ctx->render_bpc[mode];
    if (bpc <= 0) {
        int nb = 0;
        HIP_TRY(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pkern, kBlock, 0));
        bpc = nb > 0 ? nb : 1;
    }
    const int pgrid = std::min(ctx->cus * bpc, (n_chunks + 1) / 2);
    const bool persist = ctx->render_kernel == 1 ||
                         (ctx->render_kernel < 0 && n_chunks >= kPersistMinChunksPerWave * 2 * pgrid);
    // counters (and the chunk heads) cleared by a kernel: no memset node if a caller captures the stream
    hipLaunchKernelGGL(k_zero2_u64, dim3(kRHeadWords / 2 / 64), dim3(64), 0, ctx->stream, ctx->d_ctr, kCtrSlots,
                       reinterpret_cast<unsigned long long*>(ctx->d_heads), persist ? kRHeadWords / 2 : 0);
    HIP_TRY(ctx, hipGetLastError());
    if (defer) {   // the task count (its own line)
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(16), 0, ctx->stream, reinterpret_cast<unsigned long long*>(ctx->d_gcount),
                           16, 0ull);
        HIP_TRY(ctx, hipGetLastError());
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (persist) {
        hipLaunchKernelGGL(pkern, dim3(pgrid), dim3(kBlock), 0, ctx->stream, sl->dev, C, st, d_out_rgb, ctx->d_ctr,
                           ctx->d_heads, n_chunks, db);
    } else {
        auto kern = mode == 1 ? (walk ? (counts ? k_render<true, 1, 2> : k_render<false, 1, 2>)
                                 : defer ? (counts ? k_render<true, 1, 1> : k_render<false, 1, 1>)
                                         : (counts ? k_render<true, 1, 0> : k_render<false, 1, 0>))
                  : mode == 2 ? (counts ? k_render<true, 2, 0> : k_render<false, 2, 0>)
                              : (counts ? k_render<true, 0, 0> : k_render<false, 0, 0>);
        hipLaunchKernelGGL(kern, dim3(tiles), dim3(kBlock), 0, ctx->stream, sl->dev, C, st, d_out_rgb, ctx->d_ctr, db);
    }
    HIP_TRY(ctx, hipGetLastError());
    if (defer) {
        const uint32_t* perm = nullptr;
        if (db.sink.keys) {   // tasks ordered by subtree size (glass_sort.hpp)
            glass_sort(ctx->d_gkeys, ctx->d_gcount, db.sink.cap, ctx->d_gbins, ctx->d_gperm, ctx->stream);
            HIP_TRY(ctx, hipGetLastError());
            perm = ctx->d_gperm;
        }
        int nb = 0;
        HIP_TRY(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, counts ? k_glass<true> : k_glass<false>, kBlock, 0));
        const int ggrid = ctx->cus * std::max(nb, 1);
        hipLaunchKernelGGL(counts ? k_glass<true> : k_glass<false>, dim3(ggrid), dim3(kBlock), 0, ctx->stream, sl->dev,
                           st, ctx->d_gtasks, ctx->d_gcount, db.sink.cap, ctx->d_samp, db.n_px, ctx->d_ctr, perm);
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(k_sum, dim3((unsigned)((db.n_px + 255) / 256)), dim3(256), 0, ctx->stream, d_out_rgb,
                           ctx->d_samp, ctx->d_prefix, ctx->d_first, db.n_px, st.samples * st.samples);
        HIP_TRY(ctx, hipGetLastError());
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(h, ctx->d_ctr, kCtrSlots * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                ctx->stream));
    uint32_t* h_deferred = reinterpret_cast<uint32_t*>(h + kCtrSlots);
    *h_deferred = 0;
    if (defer)
        HIP_TRY(ctx, hipMemcpyAsync(h_deferred, ctx->d_gcount, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const uint32_t n_deferred = *h_deferred;
    float kms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
    fill_stats(stats, h, kms, t0);
    if (stats) stats->deferred_glass = std::min<uint64_t>(n_deferred, db.sink.cap);
    fill_counts(counts, h);
    return PCL_OK;
}

int pcl_render_device(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                      float* d_out_rgb, pcl_stats* stats) try {
    return render_impl(ctx, scene, cam, settings, d_out_rgb, stats, nullptr);
} PCL_CATCH(ctx)

int pcl_set_glass_deferral(pcl_ctx* ctx, int32_t enable, int64_t max_tasks) {
    if (!ctx) return PCL_ERR_INVALID;
    if (max_tasks < 0) return fail(ctx, PCL_ERR_INVALID, "max_tasks must be >= 0");
    ctx->defer_glass = enable ? 1 : 0;
    ctx->defer_tasks = max_tasks;
    ctx->wf->defer_glass = ctx->defer_glass;
    ctx->wf->defer_tasks = ctx->defer_tasks;
    return PCL_OK;
}

int pcl_render_counts(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                      float* d_out_rgb, pcl_stats* stats, pcl_work_counts* counts) try {
    pcl_work_counts tmp;
    return render_impl(ctx, scene, cam, settings, d_out_rgb, stats, counts ? counts : &tmp);
} PCL_CATCH(ctx)

static int img_buffers(pcl_ctx* ctx, size_t n);

int pcl_render(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings, float* out_rgb,
               pcl_stats* stats) try {
    if (!out_rgb) return fail(ctx, PCL_ERR_INVALID, "null output");
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (!get_scene(ctx, scene)) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)settings->width * settings->height * 3;
    // the image on the device: the context's staging buffer, grown on demand
    // and kept (no allocation per frame)
    if ((rc = img_buffers(ctx, n))) return rc;
    const double t0 = now_ms();
    rc = pcl_render_device(ctx, scene, cam, settings, ctx->d_img, stats);
    if (!rc && hipMemcpy(out_rgb, ctx->d_img, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(ctx, PCL_ERR_HIP, "copy image");
    if (!rc && stats) stats->total_ms = now_ms() - t0;
    return rc;
} PCL_CATCH(ctx)

static int batch_prepare(pcl_ctx* ctx, const pcl_batch_item* items, int32_t n, std::vector<pcl::wf::ItemRec>& its,
                         std::vector<DevScene>& scs) {
    if (!ctx || (n > 0 && !items) || n < 0) return fail(ctx, PCL_ERR_INVALID, "bad batch");
    its.clear();
    scs.clear();
    int64_t begin = 0;
    for (int32_t i = 0; i < n; ++i) {
        const pcl_batch_item& b = items[i];
        SceneSlot* sl = get_scene(ctx, b.scene);
        if (!sl) return fail(ctx, PCL_ERR_INVALID, "batch item " + std::to_string(i) + ": bad scene handle");
        if (!b.d_out_rgb) return fail(ctx, PCL_ERR_INVALID, "batch item " + std::to_string(i) + ": null output");
        int rc = check_settings(ctx, &b.settings);
        if (rc) return rc;
        if (b.settings.soft_shadow_samples < 1)
            return fail(ctx, PCL_ERR_INVALID, "wavefront needs soft_shadow_samples >= 1");
        int budget = 0;
        for (const auto& L : sl->lights)
            budget += b.settings.pathtrace ? 1 : ((L.radius > 0.0f) ? b.settings.soft_shadow_samples : 1);
        if (budget > pcl::wf::kMaxSh)
            return fail(ctx, PCL_ERR_LIMIT, "batch item " + std::to_string(i) + ": " + std::to_string(budget) +
                                                " light samples per shading point (max " +
                                                std::to_string(pcl::wf::kMaxSh) + ")");
        pcl::wf::ItemRec r{};
        r.cam = make_camera(b.camera);
        r.st = to_dev(&b.settings);
        r.out = b.d_out_rgb;
        r.scene = i;
        r.glass = sl->has_glass ? 1 : 0;
        r.draws = whitted_draws_per_node(*sl, r.st);
        r.pix_begin = begin;
        begin += pcl::wf::padded_pixels(b.settings.width, b.settings.height);
        its.push_back(r);
        scs.push_back(sl->dev);
    }
    return PCL_OK;
}

static pcl_wf_config to_abi(const pcl::wf::Config& c) {
    pcl_wf_config o;
    o.slots = c.slots; o.blocks_per_cu = c.blocks_per_cu; o.stack_k = c.stack_k; o.iters_per_graph = c.iters_per_graph;
    o.drain_below = c.drain_below;
    o.inline_shadows = c.inline_shadows;
    o.streams = c.streams;
    return o;
}

int pcl_render_batch(pcl_ctx* ctx, const pcl_batch_item* items, int32_t n, const pcl_wf_config* cfg,
                     pcl_batch_stats* stats) try {
    const double t0 = now_ms();
    std::vector<pcl::wf::ItemRec> its;
    std::vector<DevScene> scs;
    int rc = batch_prepare(ctx, items, n, its, scs);
    if (rc) return rc;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    pcl::wf::Config c = ctx->wf_default;
    if (cfg) {
        if (cfg->slots > 0) c.slots = cfg->slots;
        if (cfg->blocks_per_cu > 0) c.blocks_per_cu = cfg->blocks_per_cu;
        if (cfg->stack_k > 0) c.stack_k = cfg->stack_k;
        if (cfg->iters_per_graph > 0) c.iters_per_graph = cfg->iters_per_graph;
        if (cfg->drain_below != 0) c.drain_below = cfg->drain_below;
        if (cfg->inline_shadows != 0) c.inline_shadows = cfg->inline_shadows;
        if (cfg->streams < 0 || cfg->streams > pcl::wf::kMaxParts)
            return fail(ctx, PCL_ERR_INVALID, "streams must be 0..4");
        if (cfg->streams > 0) c.streams = cfg->streams;
    }
    pcl::wf::RunStats rs;
    std::string err;
    rc = ctx->wf->run(ctx->device, ctx->stream, its, scs, c, ctx->cus, rs, err);
    if (rc) return fail(ctx, rc, "wavefront: " + err);
    if (stats) {
        stats->rays_traced = rs.rays;
        stats->closest_queries = rs.closest;
        stats->anyhit_queries = rs.anyhit;
        stats->iterations = rs.iterations;
        stats->kernel_ms = rs.ms;
        stats->total_ms = now_ms() - t0;
        stats->used = to_abi(rs.cfg);
        stats->finished_pixels = rs.finished;
        stats->deferred_glass = rs.deferred;
    }
    return PCL_OK;
} PCL_CATCH(ctx)

int pcl_wf_autotune(pcl_ctx* ctx, const pcl_batch_item* items, int32_t n, pcl_wf_config* best, double* best_ms) try {
    std::vector<pcl::wf::ItemRec> its;
    std::vector<DevScene> scs;
    int rc = batch_prepare(ctx, items, n, its, scs);
    if (rc) return rc;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int64_t total = 0;
    for (auto& it : its) total += pcl::wf::padded_pixels(it.st.width, it.st.height);
    if (total <= 0) return fail(ctx, PCL_ERR_INVALID, "empty job");
    std::string err;
    auto time_cfg = [&](const pcl::wf::Config& c, float& ms) -> int {
        pcl::wf::RunStats rs;
        float bestv = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {   // first run builds the graph
            int r = ctx->wf->run(ctx->device, ctx->stream, its, scs, c, ctx->cus, rs, err);
            if (r) return r;
            if (rep > 0 && rs.ms < bestv) bestv = rs.ms;
        }
        ms = bestv;
        return PCL_OK;
    };
    pcl::wf::Config cur;
    cur.slots = (int)std::min<int64_t>(total, 1 << 20);
    cur.blocks_per_cu = 8;
    cur.stack_k = 16;
    cur.iters_per_graph = 4;
    float cur_ms = 0;
    if ((rc = time_cfg(cur, cur_ms))) return fail(ctx, rc, "autotune: " + err);
    std::vector<int> slot_opts;
    for (int64_t v : {(int64_t)1 << 18, (int64_t)1 << 19, (int64_t)1 << 20, (int64_t)1 << 21, (int64_t)1 << 22})
        if (v < total) slot_opts.push_back((int)v);
    slot_opts.push_back((int)std::min<int64_t>(total, 1 << 23));
    // The drain threshold and the concurrent engines are not tuned: the tail
    // they cut grows with the samples per pixel, so a short tuning job cannot
    // rank the candidates.  At 2x2 spp the tuner picked 48 waves' worth of
    // lanes per CU for the drain, 3% slower at 256 spp than the rule (the
    // traversal grids' lanes), and two engines, 6% slower at 256 spp than the
    // rule's three (Cornell + mesh: 2.65 s vs 2.50 s); both stay 0 = the
    // rule, applied to the job being rendered.  Shadow rays per vertex traced
    // inline: none, one, or up to 4 / 32.
    const std::vector<int> inline_opts{-1, 1, 4, 32};
    for (int knob = 0; knob < 5; ++knob) {
        std::vector<int> opts = knob == 0 ? slot_opts : knob == 1 ? std::vector<int>{4, 6, 10, 12, 16}
                              : knob == 2 ? std::vector<int>{8, 16, 32} : knob == 3 ? std::vector<int>{1, 4, 16}
                              : inline_opts;
        for (int v : opts) {
            pcl::wf::Config c = cur;
            (knob == 0 ? c.slots : knob == 1 ? c.blocks_per_cu : knob == 2 ? c.stack_k
                       : knob == 3 ? c.iters_per_graph : c.inline_shadows) = v;
            float ms;
            if ((rc = time_cfg(c, ms))) return fail(ctx, rc, "autotune: " + err);
            if (ms < cur_ms) { cur_ms = ms; cur = c; }
        }
    }
    ctx->wf_default = cur;
    // the candidates grew the engine's buffers to the largest slot count
    // tried (8M slots: ~25 GB); when that is far above the winner's, free
    // them and let the next job allocate what it needs
    if (ctx->wf->cap_slots() > 2 * (size_t)cur.slots) ctx->wf->release();
    if (best) *best = to_abi(cur);
    if (best_ms) *best_ms = cur_ms;
    return PCL_OK;
} PCL_CATCH(ctx)

// Grows the context's HDR staging buffers (forward image, target) to n floats.
static int img_buffers(pcl_ctx* ctx, size_t n) {
    if (n <= ctx->img_cap) return PCL_OK;
    if (ctx->d_img) (void)hipFree(ctx->d_img);
    if (ctx->d_tgt) (void)hipFree(ctx->d_tgt);
    ctx->d_img = ctx->d_tgt = nullptr;
    ctx->img_cap = 0;
    HIP_TRY(ctx, hipMalloc(&ctx->d_img, n * sizeof(float)));
    HIP_TRY(ctx, hipMalloc(&ctx->d_tgt, n * sizeof(float)));
    ctx->img_cap = n;
    return PCL_OK;
}

int pcl_tonemap_thresholds(float* t256) {
    if (!t256) return PCL_ERR_INVALID;
    pcl::tonemap_thresholds(t256);
    return PCL_OK;
}

int pcl_tonemap_rgb8_device(pcl_ctx* ctx, const float* d_hdr, int64_t n_px, float exposure, uint8_t* d_rgb8) {
    if (!ctx) return PCL_ERR_INVALID;
    if (n_px < 0 || (n_px > 0 && (!d_hdr || !d_rgb8))) return fail(ctx, PCL_ERR_INVALID, "bad tone-map buffers");
    if (((uintptr_t)d_hdr & 15) || ((uintptr_t)d_rgb8 & 3))
        return fail(ctx, PCL_ERR_INVALID, "tone map: HDR buffer must be 16-byte and output 4-byte aligned");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const float scale = std::pow(2.0f, exposure);   // tonemap_color's e, with the host's libm
    HIP_TRY(ctx, pcl::tm::launch(ctx->stream, d_hdr, n_px, scale, ctx->d_thr, d_rgb8));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PCL_OK;
}

int pcl_render_rgb8(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                    float exposure, uint8_t* out_rgb8, pcl_stats* stats) try {
    if (!out_rgb8) return fail(ctx, PCL_ERR_INVALID, "null output");
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (!get_scene(ctx, scene)) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    const double t0 = now_ms();
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t n_px = (size_t)settings->width * settings->height;
    if ((rc = img_buffers(ctx, n_px * 3))) return rc;
    if (n_px * 3 > ctx->rgb8_cap) {
        if (ctx->d_rgb8) (void)hipFree(ctx->d_rgb8);
        ctx->d_rgb8 = nullptr;
        ctx->rgb8_cap = 0;
        HIP_TRY(ctx, hipMalloc(&ctx->d_rgb8, n_px * 3 + 4));
        ctx->rgb8_cap = n_px * 3;
    }
    rc = render_impl(ctx, scene, cam, settings, ctx->d_img, stats, nullptr);
    if (rc) return rc;
    rc = pcl_tonemap_rgb8_device(ctx, ctx->d_img, (int64_t)n_px, exposure, ctx->d_rgb8);
    if (rc) return rc;
    HIP_TRY(ctx, hipMemcpy(out_rgb8, ctx->d_rgb8, n_px * 3, hipMemcpyDeviceToHost));
    if (stats) stats->total_ms = now_ms() - t0;
    return PCL_OK;
} PCL_CATCH(ctx)

int pcl_scene_set_materials(pcl_ctx* ctx, int32_t scene, const pcl_material* materials, int32_t n) try {
    SceneSlot* sl = get_scene(ctx, scene);
    if (!sl) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    if (!materials || n != sl->n_mats)
        return fail(ctx, PCL_ERR_INVALID, "set_materials: expected " + std::to_string(sl->n_mats) + " materials");
    std::vector<MatRec> recs;
    std::string err;
    const int rc = pack_materials(materials, n, recs, err);
    if (rc) return fail(ctx, rc, err);
    if (n == 0) return PCL_OK;
    sl->has_glass = false;
    for (int32_t m = 0; m < n; ++m) sl->has_glass |= materials[m].kind == PCL_MAT_DIELECTRIC;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(const_cast<MatRec*>(sl->dev.mats), recs.data(), recs.size() * sizeof(MatRec),
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return PCL_OK;
} PCL_CATCH(ctx)

// Grows the adjoint's partials buffer to this launch and makes sure the
// output buffer exists.
static int adj_buffers(pcl_ctx* ctx, const RSettings& st, int n_mat) {
    const size_t need = pcl::adj::partial_doubles(st, n_mat);
    if (need > ctx->adj_partial_cap) {
        if (ctx->d_adj_partial) (void)hipFree(ctx->d_adj_partial);
        ctx->d_adj_partial = nullptr;
        ctx->adj_partial_cap = 0;
        HIP_TRY(ctx, hipMalloc(&ctx->d_adj_partial, need * sizeof(double)));
        ctx->adj_partial_cap = need;
    }
    if (!ctx->d_adj_out) HIP_TRY(ctx, hipMalloc(&ctx->d_adj_out, (PCL_ADJ_MAX_MATERIALS * 6 + 1) * sizeof(double)));
    return PCL_OK;
}

static int adj_check(pcl_ctx* ctx, SceneSlot* sl, const pcl_camera* cam, const pcl_settings* settings,
                     double* grad) {
    if (!sl) return fail(ctx, PCL_ERR_INVALID, "bad scene handle");
    if (!cam || !grad) return fail(ctx, PCL_ERR_INVALID, "null camera or gradient output");
    int rc = check_settings(ctx, settings);
    if (rc) return rc;
    if (sl->n_mats > PCL_ADJ_MAX_MATERIALS)
        return fail(ctx, PCL_ERR_LIMIT, "adjoint: " + std::to_string(sl->n_mats) + " materials, at most " +
                                            std::to_string(PCL_ADJ_MAX_MATERIALS));
    return PCL_OK;
}

// Adjoint pass (which renders the image too) + reduction; copies n_mat*6
// (+ the loss) doubles back and fills stats from the pass's work counters.
static int adj_run(pcl_ctx* ctx, SceneSlot* sl, const pcl_camera* cam, const pcl_settings* settings,
                   const float* tgt, const float* adjw, float* img_out, double* grad, double* loss, pcl_stats* stats,
                   double t0) {
    const CamRec C = make_camera(*cam);
    const RSettings st = to_dev(settings);
    int rc = adj_buffers(ctx, st, sl->n_mats);
    if (rc) return rc;
    unsigned long long h[kCtrSlots] = {0};
    HIP_TRY(ctx, hipMemsetAsync(ctx->d_ctr, 0, sizeof(h), ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, pcl::adj::launch(ctx->stream, sl->dev, C, st, sl->n_mats, tgt, adjw, img_out, ctx->d_adj_partial,
                                  ctx->d_adj_out, ctx->d_ctr));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<double> hg((size_t)sl->n_mats * 6 + 1);
    HIP_TRY(ctx, hipMemcpyAsync(hg.data(), ctx->d_adj_out, hg.size() * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(h, ctx->d_ctr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    for (int p = 0; p < sl->n_mats * 6; ++p) grad[p] = hg[p];
    if (loss) *loss = hg[(size_t)sl->n_mats * 6];
    if (stats) {
        *stats = pcl_stats{};
        fill_stats(stats, h, ms, t0);
    }
    return PCL_OK;
}

int pcl_render_vjp_device(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                          const float* d_adj, double* grad, pcl_stats* stats) try {
    SceneSlot* sl = get_scene(ctx, scene);
    int rc = adj_check(ctx, sl, cam, settings, grad);
    if (rc) return rc;
    if (!d_adj) return fail(ctx, PCL_ERR_INVALID, "null image adjoint");
    const double t0 = now_ms();
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return adj_run(ctx, sl, cam, settings, nullptr, d_adj, nullptr, grad, nullptr, stats, t0);
} PCL_CATCH(ctx)

int pcl_adjoint_device(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                       const float* d_target, double* grad, double* loss, float* d_out_rgb, pcl_stats* stats) try {
    SceneSlot* sl = get_scene(ctx, scene);
    int rc = adj_check(ctx, sl, cam, settings, grad);
    if (rc) return rc;
    if (!d_target) return fail(ctx, PCL_ERR_INVALID, "null target");
    const double t0 = now_ms();
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    return adj_run(ctx, sl, cam, settings, d_target, nullptr, d_out_rgb, grad, loss, stats, t0);
} PCL_CATCH(ctx)

// pcl_adjoint_device with every work buffer in the device's scratch array:
// no allocation, the same kernels.  The counters are cleared with a kernel
// (a capturable node); partials and result are written in full by the pass.
int pcl_adjoint_device_noalloc(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                               const float* d_target, double* grad, double* loss, float* d_out_rgb,
                               pcl_stats* stats) try {
    SceneSlot* sl = get_scene(ctx, scene);
    int rc = adj_check(ctx, sl, cam, settings, grad);
    if (rc) return rc;
    if (!d_target) return fail(ctx, PCL_ERR_INVALID, "null target");
    if (!ctx->d_scratch) return fail(ctx, PCL_ERR_HIP, "no scratch memory on this device");
    const double t0 = now_ms();
    const RSettings st = to_dev(settings);
    const size_t need = pcl::adj::partial_doubles(st, sl->n_mats) * sizeof(double);
    if (need > kScratchBytes - kScrPartial)
        return fail(ctx, PCL_ERR_LIMIT, "adjoint without allocation: " + std::to_string(need) +
                                            " bytes of per-tile partials, the scratch holds " +
                                            std::to_string(kScratchBytes - kScrPartial) +
                                            " (8x8 tiles x (6 x materials + 1) doubles)");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const CamRec C = make_camera(*cam);
    unsigned long long* d_ctr = reinterpret_cast<unsigned long long*>(ctx->d_scratch);
    double* d_out = reinterpret_cast<double*>(ctx->d_scratch + kScrOut);
    double* d_part = reinterpret_cast<double*>(ctx->d_scratch + kScrPartial);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, ctx->stream, d_ctr, kCtrSlots, 0ull);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, pcl::adj::launch(ctx->stream, sl->dev, C, st, sl->n_mats, d_target, nullptr, d_out_rgb, d_part,
                                  d_out, d_ctr));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    double hg[PCL_ADJ_MAX_MATERIALS * 6 + 1];
    unsigned long long h[kCtrSlots];
    const size_t n_par = (size_t)sl->n_mats * 6;
    HIP_TRY(ctx, hipMemcpyAsync(hg, d_out, (n_par + 1) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(h, d_ctr, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    for (size_t p = 0; p < n_par; ++p) grad[p] = hg[p];
    if (loss) *loss = hg[n_par];
    if (stats) {
        *stats = pcl_stats{};
        fill_stats(stats, h, ms, t0);
    }
    return PCL_OK;
} PCL_CATCH(ctx)

int pcl_adjoint(pcl_ctx* ctx, int32_t scene, const pcl_camera* cam, const pcl_settings* settings,
                const float* target_rgb, double* grad, double* loss, float* out_rgb, pcl_stats* stats) try {
    SceneSlot* sl = get_scene(ctx, scene);
    int rc = adj_check(ctx, sl, cam, settings, grad);
    if (rc) return rc;
    if (!target_rgb) return fail(ctx, PCL_ERR_INVALID, "null target");
    const double t0 = now_ms();
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)settings->width * settings->height * 3;
    if ((rc = img_buffers(ctx, n))) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->d_tgt, target_rgb, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    rc = pcl_adjoint_device(ctx, scene, cam, settings, ctx->d_tgt, grad, loss, out_rgb ? ctx->d_img : nullptr, stats);
    if (rc) return rc;
    if (out_rgb) HIP_TRY(ctx, hipMemcpy(out_rgb, ctx->d_img, n * sizeof(float), hipMemcpyDeviceToHost));
    if (stats) stats->total_ms = now_ms() - t0;
    return PCL_OK;
} PCL_CATCH(ctx)

int pcl_device_alloc(pcl_ctx* ctx, size_t bytes, void** d_ptr) {
    if (!ctx || !d_ptr) return fail(ctx, PCL_ERR_INVALID, "null argument");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMalloc(d_ptr, bytes));
    return PCL_OK;
}
int pcl_device_free(pcl_ctx* ctx, void* d_ptr) {
    if (!ctx) return PCL_ERR_INVALID;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipFree(d_ptr));
    return PCL_OK;
}
int pcl_memcpy_h2d(pcl_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
    if (!ctx) return PCL_ERR_INVALID;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
    return PCL_OK;
}
int pcl_memcpy_d2h(pcl_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
    if (!ctx) return PCL_ERR_INVALID;
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
    return PCL_OK;
}

}  // extern "C"
