// wavefront.hpp - state of the batched wavefront tracer (see wavefront.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "render.hpp"

namespace pcl {
namespace wf {

constexpr int kMaxSh = 32;   // light samples per shading point (sum over lights)

// Job pixels per item: the image padded to whole 8x8 tiles (tile order).
inline int64_t padded_pixels(int w, int h) { return (int64_t)((w + 7) / 8 * 8) * (int64_t)((h + 7) / 8 * 8); }

// One job item: a scene, its camera frame, settings and device output.
struct ItemRec {
    CamRec cam;
    RSettings st;
    float* out;
    int32_t scene;
    int32_t pad;
    int64_t pix_begin;   // first job pixel index of this item
};

// Pixel-slot state, structure of arrays (one entry per slot).
struct Slots {
    int32_t *item, *pixel, *sample, *mode, *depth, *nf, *hit, *outcome, *outcome_aux, *glass_root, *nsh;
    uint32_t* rng;
    float *acc, *ray_o, *ray_d, *tacc, *T, *NT, *R, *pend, *alb, *gT, *gE;
    float4* sh;          // [slot][kMaxSh] light-sample contribution + meta
    uint8_t* occl;      // [slot][kMaxSh] occluded flags
    Frame* frames;       // [slot][kMaxFrames]
};

constexpr int kXcds = 8;   // MI355X: 8 XCDs, each with its own L2
enum : int {
    Q_CUR = 0, Q_COUNT0 = 1, Q_COUNT1 = 2, Q_SHADOW = 3, Q_PIXEL = 4,
    // per-stage work heads, one per XCD segment of the queue
    Q_HEAD_CLOSEST = 16, Q_HEAD_SHADE = 24, Q_HEAD_ANY = 32, Q_HEAD_RESOLVE = 40, Q_NCTR = 48
};
enum : int { ST_RAYS = 0, ST_CLOSEST = 1, ST_ANY = 2, ST_ITERS = 3, ST_N = 8 };

struct Queues {
    int32_t* ray_q[2];        // slot ids awaiting a closest-hit query
    int2* sh_who;             // shadow queue: (slot, sample index)
    float4* sh_ray;           // shadow queue: (o, t_min), (d, t_max)
    uint32_t* ctr;            // Q_* counters
    unsigned long long* stats;
    int64_t total_pixels;
};

struct Config {
    int slots = 0;            // pixel slots in flight (queue depth); 0 = auto
    int blocks_per_cu = 0;    // persistent traversal blocks per CU; 0 = auto
    int stack_k = 0;          // LDS stack entries per lane (8, 16, 32); 0 = auto
    int iters_per_graph = 0;  // iterations captured per hipGraph; 0 = auto
};

struct RunStats {
    unsigned long long rays = 0, closest = 0, anyhit = 0, iterations = 0;
    int graph_launches = 0;
    float ms = 0.0f;
    Config cfg;
};

struct Engine {
    Slots sl{};
    Queues q{};
    ItemRec* d_items = nullptr;
    DevScene* d_scenes = nullptr;
    int cap_slots = 0, cap_items = 0, cap_shadow = 0;
    std::vector<void*> bufs;
    hipGraphExec_t graph_exec = nullptr;
    Config graph_key;
    int graph_items = 0;
    int64_t graph_total = 0;

    ~Engine();
    void release();
    int alloc(size_t bytes, void** p, std::string& err);
    int ensure(int slots, int items, std::string& err);
    int run(hipStream_t stream, const std::vector<ItemRec>& items, const std::vector<DevScene>& scenes,
            const Config& cfg, int cus, RunStats& out, std::string& err);
};

}  // namespace wf
}  // namespace pcl
