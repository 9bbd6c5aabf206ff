// ref_harness.cpp - fixture generator compiled against the reference headers.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target _ref/ref_harness)
// from this file plus the reference's own camera.cpp / stb_impl.cpp, with
// -I pointing at the reference sources where they lie.  It drives the
// reference's parser, BVH and Renderer and dumps what the golden PNGs cannot
// show: HDR pixels, primary-hit (t, primitive id), the BVH leaf order and split
// points, and the reference-convention ray count.
//
// Primitive identity: HitRecord carries no primitive id, so every primitive is
// wrapped in a delegating subclass that records its insertion index in a
// thread_local when its hit() succeeds.  out_rec is written only on success,
// so the last success is the record Scene::hit returns.  aabb() is forwarded
// unchanged, so the BVH built over the wrappers has the same topology.
//
// Usage:
//   ref_harness render <scene.pecl> <out.bin> [opts]   HDR float32 W*H*3 + u64 rays
//                       (--rows r0 r1: only rows [r0, r1) through the public
//                       render_tile_region, header W, r1-r0)
//   ref_harness hits   <scene.pecl> <W> <H> <out.bin>  pixel-center primary hits
//   ref_harness bvh    <scene.pecl> <out.bin>          leaf order + split points
//   ref_harness loss   <scene.pecl> <target.bin> <out.bin> [opts]
//                       L2 loss and central FD of it over albedo/emission
// opts: --width W --height H --samples n --depth d --pathtrace --mis --bounces b
//       --threads T --h step --no-fd --rows r0 r1
//       --repeat N   render N times after one load, printing each render's
//                    seconds (a CPU baseline: parse and BVH build untimed,
//                    as src/pencil.cpp:311-313 times render() only)
//       --no-tag     leave the primitives unwrapped (the reference's own
//                    objects; ids are then not recorded -- for timing)

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

// Scene and BVH keep their storage private; the harness needs read access.
#define private public
#include "pecl_parser.hpp"
#include "renderer.hpp"
#undef private

using namespace pencil;

static thread_local int g_last_id = -1;
// Primitive tests per kind (sphere, triangle, cube, plane, cylinder, cone):
// every Primitive::hit call the traversal makes, closest-hit and occlusion.
static std::atomic<unsigned long long> g_tests[6];

static int kind_of(const std::shared_ptr<Primitive>& p) {
    if (std::dynamic_pointer_cast<Sphere>(p)) return 0;
    if (std::dynamic_pointer_cast<Triangle>(p)) return 1;
    if (std::dynamic_pointer_cast<Cube>(p)) return 2;
    if (std::dynamic_pointer_cast<Plane>(p)) return 3;
    if (std::dynamic_pointer_cast<Cylinder>(p)) return 4;
    if (std::dynamic_pointer_cast<Cone>(p)) return 5;
    return -1;
}

struct Tagged : Primitive {
    std::shared_ptr<Primitive> inner;
    int id, kind;
    Tagged(std::shared_ptr<Primitive> p, int i) : inner(std::move(p)), id(i), kind(kind_of(inner)) {}
    bool hit(const Ray& r, float t_min, float t_max, HitRecord& out) const override {
        if (kind >= 0) g_tests[kind].fetch_add(1, std::memory_order_relaxed);
        if (inner->hit(r, t_min, t_max, out)) { g_last_id = id; return true; }
        return false;
    }
    AABB aabb() const override { return inner->aabb(); }
};

struct Opts {
    int width = 0, height = 0, samples = 0, depth = 0, threads = 0, bounces = 4;
    int row0 = -1, row1 = -1, repeat = 1;
    bool tag = true;
    bool pathtrace = false, fd = true, mis = false;
    float h = 1e-3f;
};

static Opts parse_opts(int argc, char** argv, int from) {
    Opts o;
    for (int i = from; i < argc; ++i) {
        std::string s = argv[i];
        auto nxt = [&]() { return (i + 1 < argc) ? argv[++i] : (char*)"0"; };
        if (s == "--width") o.width = atoi(nxt());
        else if (s == "--height") o.height = atoi(nxt());
        else if (s == "--samples") o.samples = atoi(nxt());
        else if (s == "--depth") o.depth = atoi(nxt());
        else if (s == "--threads") o.threads = atoi(nxt());
        else if (s == "--bounces") o.bounces = atoi(nxt());
        else if (s == "--pathtrace") o.pathtrace = true;
        else if (s == "--mis") o.mis = true;
        else if (s == "--h") o.h = (float)atof(nxt());
        else if (s == "--no-fd") o.fd = false;
        else if (s == "--rows") { o.row0 = atoi(nxt()); o.row1 = atoi(nxt()); }
        else if (s == "--repeat") o.repeat = std::max(1, atoi(nxt()));
        else if (s == "--no-tag") o.tag = false;
    }
    return o;
}

// Parse + apply CLI-style overrides exactly as src/pencil.cpp load_scene does.
static bool load(const std::string& path, ParsedScene& ps, const Opts& o) {
    std::vector<ParseError> errs;
    if (!parse_pecl(path, ps, errs)) {
        for (auto& e : errs) std::cerr << "line " << e.line << ": " << e.what << "\n";
        return false;
    }
    if (o.width > 0) { ps.width = o.width; ps.aspect = (float)ps.width / (float)ps.height; ps.finalize(); }
    if (o.height > 0) { ps.height = o.height; ps.aspect = (float)ps.width / (float)ps.height; ps.finalize(); }
    if (o.samples > 0) ps.samples = o.samples;
    if (o.depth > 0) ps.depth = o.depth;
    // wrap primitives and rebuild the BVH over the wrappers
    if (!o.tag) return true;
    auto& prims = ps.scene.primitives_;
    for (size_t i = 0; i < prims.size(); ++i) prims[i] = std::make_shared<Tagged>(prims[i], (int)i);
    ps.scene.build_bvh();
    return true;
}

static RenderSettings settings_for(const ParsedScene& ps, const Opts& o) {
    RenderSettings rs;
    rs.width = ps.width; rs.height = ps.height;
    rs.samples = ps.samples; rs.max_depth = ps.depth;
    rs.threads = o.threads;
    rs.pathtrace = o.pathtrace || ps.enable_pathtrace;
    rs.pathtrace_bounces = o.bounces;
    rs.mis = o.mis;
    return rs;
}

static void write_all(const std::string& path, const void* p, size_t n, std::ofstream* keep = nullptr) {
    std::ofstream f(path, std::ios::binary | std::ios::app);
    f.write((const char*)p, (std::streamsize)n);
    (void)keep;
}

static int cmd_render(int argc, char** argv) {
    Opts o = parse_opts(argc, argv, 4);
    ParsedScene ps;
    if (!load(argv[2], ps, o)) return 3;
    RenderSettings rs = settings_for(ps, o);
    Renderer r(ps.scene, ps.camera, rs);
    std::vector<Color> px; RenderStats st;
    const int W = rs.width, H = rs.height;
    const bool band = o.row0 >= 0;
    const int r0 = band ? std::max(0, o.row0) : 0, r1 = band ? std::min(H, o.row1) : H;
    if (r1 <= r0) return 2;
    uint64_t rays = 0;
    std::cout << "seconds";
    for (int rep = 0; rep < o.repeat; ++rep) {
        auto t0 = std::chrono::steady_clock::now();
        if (band) {
            // A band of rows, 16-pixel column tiles handed out to the threads; a
            // pixel's value does not depend on the tile it is rendered in (its RNG
            // is seeded from (i, j), src/renderer.hpp:313).
            px.assign((size_t)W * H, Color{0, 0, 0});
            r.ray_counter_.store(0);
            std::atomic<int> next{0};
            const int tiles = (W + 15) / 16;
            int T = o.threads > 0 ? o.threads : (int)std::thread::hardware_concurrency();
            if (T <= 0) T = 1;
            std::vector<std::thread> ws;
            for (int t = 0; t < T; ++t)
                ws.emplace_back([&]() {
                    for (int k; (k = next.fetch_add(1)) < tiles;)
                        r.render_tile_region(px, W, H, k * 16, r0, std::min(16, W - k * 16), r1 - r0);
                });
            for (auto& w : ws) w.join();
            rays = r.ray_counter_.load();
        } else {
            r.render(px, &st);
            rays = st.rays_traced;
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << " " << sec << std::flush;
    }
    std::cout << "\nrays " << rays << "\ntests";
    for (auto& t : g_tests) std::cout << " " << t.load();   // over all repeats
    std::cout << "\n";
    std::remove(argv[3]);
    int32_t hdr[2] = {W, r1 - r0};
    write_all(argv[3], hdr, sizeof(hdr));
    write_all(argv[3], px.data() + (size_t)r0 * W, (size_t)(r1 - r0) * W * sizeof(Color));
    write_all(argv[3], &rays, sizeof(rays));
    return 0;
}

static int cmd_hits(int argc, char** argv) {
    if (argc < 6) return 2;
    Opts o;
    ParsedScene ps;
    if (!load(argv[2], ps, o)) return 3;
    int W = atoi(argv[3]), H = atoi(argv[4]);
    std::remove(argv[5]);
    std::vector<float> rec((size_t)W * H * 10);
    for (int j = 0; j < H; ++j)
        for (int i = 0; i < W; ++i) {
            float u = (i + 0.5f) / (float)W;
            float v = 1.0f - (j + 0.5f) / (float)H;
            Ray ray = ps.camera.ray_at(u, v, nullptr);
            HitRecord h; g_last_id = -1;
            float* q = &rec[((size_t)j * W + i) * 10];
            if (ps.scene.hit(ray, 1e-4f, 1e9f, h)) {
                q[0] = h.t; int32_t id = g_last_id; memcpy(&q[1], &id, 4);
                q[2] = h.p.x; q[3] = h.p.y; q[4] = h.p.z;
                q[5] = h.normal.x; q[6] = h.normal.y; q[7] = h.normal.z;
                q[8] = (float)h.material_id; q[9] = h.front_face ? 1.0f : 0.0f;
            } else {
                q[0] = -1.0f; int32_t id = -1; memcpy(&q[1], &id, 4);
                for (int k = 2; k < 10; ++k) q[k] = 0.0f;
            }
        }
    int32_t hdr[2] = {W, H};
    write_all(argv[5], hdr, sizeof(hdr));
    write_all(argv[5], rec.data(), rec.size() * sizeof(float));
    return 0;
}

static void walk(const std::shared_ptr<BVHNode>& n, std::vector<int32_t>& leaves, std::vector<int32_t>& splits) {
    if (!n) return;
    if (n->is_leaf()) { leaves.push_back(std::static_pointer_cast<Tagged>(n->leaf)->id); return; }
    size_t idx = splits.size();
    splits.push_back(0);
    walk(n->left, leaves, splits);
    splits[idx] = (int32_t)leaves.size();   // split point = start of right range
    walk(n->right, leaves, splits);
}

static int cmd_bvh(int argc, char** argv) {
    if (argc < 4) return 2;
    Opts o;
    ParsedScene ps;
    if (!load(argv[2], ps, o)) return 3;
    std::vector<int32_t> leaves, splits;
    walk(ps.scene.bvh_.root_, leaves, splits);
    // cross-check: DFS leaf order equals BVH::storage_ order
    auto& st = ps.scene.bvh_.storage_;
    for (size_t i = 0; i < st.size(); ++i)
        if (std::static_pointer_cast<Tagged>(st[i])->id != leaves[i]) { std::cerr << "order mismatch\n"; return 4; }
    std::remove(argv[3]);
    int32_t hdr[2] = {(int32_t)leaves.size(), (int32_t)splits.size()};
    write_all(argv[3], hdr, sizeof(hdr));
    write_all(argv[3], leaves.data(), leaves.size() * 4);
    write_all(argv[3], splits.data(), splits.size() * 4);
    return 0;
}

// L2 loss of a render against a target HDR image, and central finite
// differences of it w.r.t. every material's albedo (3) and emission (3).
static double loss_of(const ParsedScene& ps, const RenderSettings& rs, const std::vector<float>& target) {
    Renderer r(ps.scene, ps.camera, rs);
    std::vector<Color> px;
    r.render(px, nullptr);
    double L = 0.0;
    for (size_t i = 0; i < px.size(); ++i) {
        double d0 = (double)px[i].x - target[i * 3 + 0];
        double d1 = (double)px[i].y - target[i * 3 + 1];
        double d2 = (double)px[i].z - target[i * 3 + 2];
        L += d0 * d0 + d1 * d1 + d2 * d2;
    }
    return L;
}

static int cmd_loss(int argc, char** argv) {
    if (argc < 5) return 2;
    Opts o = parse_opts(argc, argv, 5);
    ParsedScene ps;
    if (!load(argv[2], ps, o)) return 3;
    RenderSettings rs = settings_for(ps, o);
    std::ifstream tf(argv[3], std::ios::binary);
    std::vector<float> target((size_t)rs.width * rs.height * 3);
    tf.read((char*)target.data(), (std::streamsize)(target.size() * 4));
    if (!tf) { std::cerr << "short target\n"; return 4; }
    std::remove(argv[4]);
    double L = loss_of(ps, rs, target);
    write_all(argv[4], &L, sizeof(L));
    if (!o.fd) return 0;
    auto& mats = ps.scene.materials_;
    std::vector<double> g;
    for (size_t m = 0; m < mats.size(); ++m) {
        for (int which = 0; which < 2; ++which) {
            for (int c = 0; c < 3; ++c) {
                Color& field = which == 0 ? mats[m].albedo : mats[m].emission;
                float base = field[c];
                field[c] = base + o.h;
                double Lp = loss_of(ps, rs, target);
                field[c] = base - o.h;
                double Lm = loss_of(ps, rs, target);
                field[c] = base;
                g.push_back((Lp - Lm) / (2.0 * (double)o.h));
            }
        }
    }
    write_all(argv[4], g.data(), g.size() * sizeof(double));
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 3) { std::cerr << "usage: ref_harness render|hits|bvh|loss ...\n"; return 2; }
    std::string c = argv[1];
    if (c == "render" && argc >= 4) return cmd_render(argc, argv);
    if (c == "hits") return cmd_hits(argc, argv);
    if (c == "bvh") return cmd_bvh(argc, argv);
    if (c == "loss") return cmd_loss(argc, argv);
    std::cerr << "bad command\n";
    return 2;
}
