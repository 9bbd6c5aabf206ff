// renderer.hpp - Renderer / RenderSettings / RenderStats of the Pencil API,
// backed by the HIP core.
//
// Same entry point as the reference (src/renderer.hpp:153-198):
//   Renderer r(scene, camera, settings);  r.render(pixels, &stats);
// RenderSettings gains `backend` and `device`; the only backend this
// library implements is Backend::Hip.  The reference's CPU render loop is
// the reference itself -- asking this library for it is an error, never a
// silent fallback.  render() is synchronous, leaves W*H linear HDR colors
// in `out_pixels` (row 0 = top), and fills RenderStats with the
// reference-convention ray count plus the algorithmic query counts.
#pragma once

#include "scene.hpp"

#include <cstdio>
#include <string>
#include <vector>

namespace pencil {

enum class Backend : int { Cpu = 0, Hip = 1 };

struct RenderSettings {
    int width = 400;
    int height = 300;
    int samples = 4;
    int max_depth = 8;
    float shadow_bias = 1e-3f;
    int threads = 0;
    int soft_shadow_samples = 4;
    bool pathtrace = false;
    bool mis = false;            // trace_path_mis instead of trace_path (src/renderer.hpp:55)
    int pathtrace_bounces = 4;
    int tile_size = 64;
    int bokeh_blades = 0;        // > 0 overrides the camera's aperture shape (src/renderer.hpp:162-167)
    Backend backend = Backend::Hip;
    int device = 0;
};

struct RenderStats {
    uint64_t rays_traced = 0;       // trace()/trace_path() calls, as the reference counts
    uint64_t closest_queries = 0;   // every closest-hit query cast
    uint64_t anyhit_queries = 0;    // every shadow query cast
    double kernel_ms = 0.0;
    double upload_ms = 0.0;
    double build_ms = 0.0;
};

class Renderer {
public:
    Renderer(const Scene& scene, const Camera& camera, RenderSettings s)
        : scene_(scene), camera_(camera), settings_(s) {
        // a CLI --bokeh N wins over the scene's cmra { bd: N }; 0 keeps the scene's
        if (settings_.bokeh_blades > 0) camera_.set_bokeh_blades(settings_.bokeh_blades);
    }
    ~Renderer() {
        if (ctx_) {
            if (scene_id_ >= 0) pcl_scene_free(ctx_, scene_id_);
            pcl_destroy(ctx_);
        }
    }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    // Returns false (and leaves black pixels) on failure; error() says why.
    bool render(std::vector<Color>& out_pixels, RenderStats* stats = nullptr) {
        const int W = settings_.width, H = settings_.height;
        out_pixels.assign((size_t)W * H, Color{0, 0, 0});
        if (settings_.backend != Backend::Hip) {
            return fail("backend cpu is the reference renderer itself; this library renders with backend hip only");
        }
        if (!ctx_) {
            int rc = pcl_init(settings_.device, &ctx_);
            if (rc != PCL_OK) { ctx_ = nullptr; return fail("pcl_init failed (" + std::to_string(rc) + "): no HIP device?"); }
        }
        if (scene_id_ < 0) {
            std::vector<pcl_material> mats;
            std::vector<pcl_light> lights;
            pcl_scene_desc d = scene_.desc(mats, lights);
            int rc = pcl_scene_upload(ctx_, &d, &scene_id_);
            if (rc != PCL_OK) { scene_id_ = -1; return fail(std::string("scene upload: ") + pcl_last_error(ctx_)); }
            pcl_scene_info_get(ctx_, scene_id_, &info_);
        }
        pcl_settings ps{};
        ps.width = W; ps.height = H; ps.samples = settings_.samples; ps.max_depth = settings_.max_depth;
        ps.shadow_bias = settings_.shadow_bias; ps.soft_shadow_samples = settings_.soft_shadow_samples;
        ps.pathtrace = settings_.pathtrace ? (settings_.mis ? 2 : 1) : 0;
        ps.pathtrace_bounces = settings_.pathtrace_bounces;
        pcl_stats st{};
        static_assert(sizeof(Color) == 3 * sizeof(float), "Color is three packed floats");
        int rc = pcl_render(ctx_, scene_id_, &camera_.args(), &ps, reinterpret_cast<float*>(out_pixels.data()), &st);
        if (rc != PCL_OK) return fail(std::string("render: ") + pcl_last_error(ctx_));
        if (stats) {
            stats->rays_traced = st.rays_traced;
            stats->closest_queries = st.closest_queries;
            stats->anyhit_queries = st.anyhit_queries;
            stats->kernel_ms = st.kernel_ms;
            stats->upload_ms = info_.upload_ms;
            stats->build_ms = info_.build_ms;
        }
        return true;
    }

    const std::string& error() const { return error_; }

private:
    bool fail(const std::string& m) {
        error_ = m;
        std::fprintf(stderr, "pencil: %s\n", m.c_str());
        return false;
    }

    const Scene& scene_;
    Camera camera_;
    RenderSettings settings_;
    pcl_ctx* ctx_ = nullptr;
    int32_t scene_id_ = -1;
    pcl_scene_info info_{};
    std::string error_;
};

}  // namespace pencil
