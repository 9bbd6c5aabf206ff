// wavefront.hpp - state of the batched wavefront tracer (see wavefront.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "render.hpp"

namespace pcl {
namespace wf {

constexpr int kMaxSh = 32;   // light samples per shading point (sum over lights)

// Job pixels per item: the image with each row padded to whole 64-pixel
// segments (segment order: row-major segments, pixels left to right).
inline int64_t padded_pixels(int w, int h) { return (int64_t)((w + 63) / 64 * 64) * (int64_t)h; }

// One job item: a scene, its camera frame, settings and device output.
struct ItemRec {
    CamRec cam;
    RSettings st;
    float* out;
    int32_t scene;
    int32_t glass;       // the scene has a dielectric material
    int64_t pix_begin;   // first job pixel index of this item
    int32_t draws;       // generator draws per shaded Whitted node (soft lights)
    int32_t pad;
};

// Pixel-slot state, structure of arrays of 16-byte records (one per slot),
// so a kernel loads and stores a slot's state with a few dwordx4 accesses.
struct Slots {
    int4* meta;          // item, pixel, sample index, rng state
    int4* ctl;           // mode | glass_root<<4 | glass-discard<<5 | outcome<<8, depth, frames, samples
    float4 *ray_o, *ray_d;   // current ray; after a closest query .w = t_max at acceptance / hit slot
    float4 *acc, *T, *NT, *R, *pend, *alb, *gT, *gE;
    // Light samples of the current vertex, sample-major ([kMaxSh][sh_stride]):
    // sample q of slot s at q * sh_stride + s, so the lanes of a wave working
    // on neighbouring slots touch neighbouring records.
    float4* sh;          // contribution + meta
    uint8_t* occl;       // occluded flags
    int64_t sh_stride;   // slots allocated
    Frame* frames;       // [slot][kMaxFrames]
};

// Work queues and counters are split into kStripes stripes.  Slot s belongs
// to stripe (s / 64) % kStripes for its whole life, and every queue holds a
// stripe's slots in that stripe's own segment, so each stripe's counters get
// 1/kStripes of the atomics.  Every counter sits in its own 128-byte line:
// device-scope atomics on one line serialize (about 11 ns each on MI355X),
// atomics on different lines proceed in parallel.
constexpr int kStripes = 64;   // equal to the wave size: a scan reads one stripe per lane
constexpr int kLine = 32;      // uint32 words per 128-byte line
// global counters, line 0: current queue, queries pending after the last
// iteration, drain flag (set by the host), job pixels not yet assigned after
// the last iteration, finish-list length
enum : int { C_CUR = 0, C_PENDING = 1, C_DRAIN = 2, C_POOL = 3, C_NFINISH = 4 };
constexpr int kStatusWords = 4;   // words 0..3 of line 0, copied to the host after every graph launch
enum : int {
    S_COUNT0 = 0, S_COUNT1 = 1, S_SHADOW = 2,   // queue lengths per stripe
    S_HEAD_CLOSEST = 3, S_HEAD_SHADE = 4, S_HEAD_ANY = 5, S_HEAD_RESOLVE = 6,   // work heads
    S_PIXEL = 7,   // next job pixel of the stripe's share of the job
    S_N = 8
};
PCL_HD int sidx(int type, int g) { return (1 + type * kStripes + g) * kLine; }
constexpr int kCtrWords = (1 + S_N * kStripes) * kLine;
// statistics: line 0 global (written by the swap kernel), then one line per stripe
enum : int { ST_RAYS = 0, ST_CLOSEST = 1, ST_ANY = 2, ST_ITERS = 3 };
constexpr int kStatLine = 16;   // u64 per line
constexpr int kStatWords = (1 + kStripes) * kStatLine;

struct Queues {
    int32_t* ray_q[2];        // slot ids awaiting a closest-hit query, [stripe][qcap]
    // Pixels the drain's finish kernel completes (at most one per slot), one
    // record each: (item, pixel, next sample, rng) and (acc, -).
    int4* fin_meta;
    float4* fin_acc;
    int2* sh_who;             // shadow queue: (slot, sample index), [stripe][shcap]
    float4* sh_ray;           // shadow queue: (o, t_min), (d, t_max) pairs
    uint32_t* ctr;            // counters, see sidx()
    unsigned long long* stats;
    int32_t qcap, shcap;      // per-stripe capacities
    int64_t total_pixels;     // this part's job pixels (a multiple of 64)
    // The job's 64-pixel tiles are dealt to `parts` engines that run
    // concurrently on their own streams: this engine's local tile k is job
    // tile k * parts + part.
    int32_t part, parts;
};

struct Config {
    int slots = 0;            // pixel slots in flight (queue depth); 0 = auto
    int blocks_per_cu = 0;    // persistent traversal blocks per CU; 0 = auto
    int stack_k = 0;          // LDS stack entries per lane (8, 16, 32); 0 = auto
    int iters_per_graph = 0;  // iterations captured per hipGraph; 0 = auto
    // Drain: once every job pixel is assigned and fewer than this many slots
    // are active, slots stop at their next sample boundary and one
    // per-pixel kernel finishes the remaining samples (no iteration
    // barriers in the tail).  0 = auto, < 0 = never.
    int drain_below = 0;
    // A vertex with at most this many shadow rays (light samples of
    // non-emissive lights) traces them in the closest-hit kernel right after
    // shading; vertices with more go through the shadow queue and the any-hit
    // kernel.  0 = auto (1), < 0 = never.
    int inline_shadows = 0;
    // Engines (each with its own stream, slots, queues and graph) the job's
    // tiles are dealt to, interleaved; they run concurrently, so one engine's
    // iteration tails and drain overlap another's iterations.  The slots are
    // split evenly among them.  0 = auto (1 to 3 by the job's settings), at
    // most kMaxParts.
    int streams = 0;
};

constexpr int kMaxParts = 4;

// Tiles and pixels of part `part` of `parts` in a job of `total` padded pixels.
inline int64_t part_pixels(int64_t total, int part, int parts) {
    const int64_t tiles = total >> 6;
    return tiles > part ? (tiles - part + parts - 1) / parts * 64 : 0;
}

struct RunStats {
    unsigned long long rays = 0, closest = 0, anyhit = 0, iterations = 0;
    unsigned long long finished = 0;   // pixels completed by the drain's finish kernel
    unsigned long long deferred = 0;   // glass vertices the finish kernel deferred
    int graph_launches = 0;
    float ms = 0.0f;
    Config cfg;
};

struct Engine {
    Slots sl{};
    Queues q{};
    ItemRec* d_items = nullptr;
    DevScene* d_scenes = nullptr;
    int cap_slots = 0, cap_items = 0;
    std::vector<void*> bufs;
    hipGraphExec_t graph_exec = nullptr;
    Config graph_key;
    int graph_items = 0;
    int64_t graph_total = 0;
    int graph_part = -1, graph_parts = 0;
    bool graph_fold = false;
    uint32_t* h_pending = nullptr;      // pinned, kStatusWords per graph launch in flight
    hipEvent_t ev_pending[2] = {nullptr, nullptr};
    // Deferred glass vertices of the finish launch (path-traced jobs): the
    // task array and its counter, the per-sample values of the listed pixels
    // and their prefix sums / first deferred samples; grown on demand and
    // kept until release().
    float4* d_gtasks = nullptr;
    uint32_t* d_gcount = nullptr;
    uint32_t* d_gkeys = nullptr;    // per task: the walk's count (soft lights), and the launch order
    uint32_t* d_gperm = nullptr;
    uint32_t* d_gbins = nullptr;    // 512 words of the sort
    size_t gkey_cap = 0;
    float* d_samp = nullptr;
    float* d_prefix = nullptr;
    int32_t* d_first = nullptr;
    size_t gtask_cap = 0, samp_cap = 0, px_cap = 0;
    int defer_glass = 1;   // 0: the finish launch evaluates glass vertices in place
    int64_t defer_tasks = 0;   // > 0: task array capacity limit (tests)

    ~Engine();
    void release();
    int alloc(size_t bytes, void** p, std::string& err);
    int ensure(int slots, int items, std::string& err);
    // Sizes the deferred-glass buffers for n_px listed pixels of at most nn
    // samples; false (finish in place) when they would not fit.
    bool defer_buffers(int64_t n_px, int64_t nn);
    // Runs part `part` of `parts` of the job on `stream`; `done`, when given,
    // is recorded on the stream after the part's last kernel.
    int run(hipStream_t stream, const std::vector<ItemRec>& items, const std::vector<DevScene>& scenes,
            const Config& cfg, int cus, RunStats& out, std::string& err, int part = 0, int parts = 1,
            hipEvent_t done = nullptr);
};

// The engines of a context and the streams of parts 1.. (part 0 runs on the
// context's stream).  run() deals the job to cfg.streams engines, drives
// parts 1.. from host threads of their own (each waits on its own stream)
// and part 0 from the calling thread, and returns when every part is done.
struct Tracer {
    Engine eng[kMaxParts];
    hipStream_t side[kMaxParts] = {};
    hipEvent_t ev_fork = nullptr, ev_done[kMaxParts] = {};
    int defer_glass = 1;       // handed to the engines (Engine::defer_glass)
    int64_t defer_tasks = 0;

    ~Tracer();
    void release();
    size_t cap_slots() const;
    int run(int device, hipStream_t stream, const std::vector<ItemRec>& items, const std::vector<DevScene>& scenes,
            const Config& cfg, int cus, RunStats& out, std::string& err);
};

}  // namespace wf
}  // namespace pcl
