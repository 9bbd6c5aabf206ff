// host_capi.cpp - C entry points over the host-side Pencil API, for callers
// that do not link C++ (the Python tests and bench): load a .pecl with the
// CLI's overrides, build the synthetic stress scene, and hand back the
// pcl_scene_desc / pcl_camera / pcl_settings that feed include/pencil_hip.h.
#include "image_io.hpp"
#include "pecl_parser.hpp"
#include "renderer.hpp"
#include "stress_scene.hpp"

#include <cstring>
#include <exception>
#include <memory>
#include <string>

namespace {

struct Handle {
    pencil::ParsedScene ps;
    std::vector<pcl_material> mats;
    std::vector<pcl_light> lights;
    pcl_scene_desc desc{};
    std::string unsupported;
    void refresh() { desc = ps.scene.desc(mats, lights); }
};

void copy_err(const std::string& s, char* buf, int len) {
    if (!buf || len <= 0) return;
    std::snprintf(buf, (size_t)len, "%s", s.c_str());
}

}  // namespace

// The parser converts numbers with std::stof, as the reference's does, so a
// malformed scene can throw; no exception crosses the C boundary.
#define PCLH_CATCH_PTR(buf, len)                                                             \
    catch (const std::exception& e_) { copy_err(std::string("error: ") + e_.what(), buf, len); return nullptr; } \
    catch (...) { copy_err("error", buf, len); return nullptr; }
#define PCLH_CATCH_INT(buf, len)                                                             \
    catch (const std::exception& e_) { copy_err(std::string("error: ") + e_.what(), buf, len); return -1; } \
    catch (...) { copy_err("error", buf, len); return -1; }

extern "C" {

// Parse a .pecl file and apply --width/--height/--samples/--depth overrides
// the way the CLI's load_scene does (src/pencil.cpp:246-269); 0 = keep.
void* pclh_load_pecl(const char* path, int width, int height, int samples, int depth, char* err, int errlen) try {
    auto h = std::make_unique<Handle>();
    std::vector<pencil::ParseError> errs;
    bool ok = pencil::parse_pecl(path, h->ps, errs);
    if (!ok) {
        std::string m = "failed to parse scene: " + std::string(path);
        for (auto& e : errs) m += "\n  line " + std::to_string(e.line) + ": " + e.what;
        copy_err(m, err, errlen);
        return nullptr;
    }
    auto& ps = h->ps;
    if (width > 0) { ps.width = width; ps.aspect = (float)ps.width / (float)ps.height; ps.finalize(); }
    if (height > 0) { ps.height = height; ps.aspect = (float)ps.width / (float)ps.height; ps.finalize(); }
    if (samples > 0) ps.samples = samples;
    if (depth > 0) ps.depth = depth;
    for (auto& u : ps.unsupported) h->unsupported += (h->unsupported.empty() ? "" : "; ") + u;
    h->refresh();
    return h.release();
} PCLH_CATCH_PTR(err, errlen)

void* pclh_make_stress(int freq, int n_spheres, uint32_t seed, int width, int height) try {
    auto h = std::make_unique<Handle>();
    pencil::StressParams P;
    P.freq = freq; P.n_spheres = n_spheres; P.seed = seed; P.width = width; P.height = height;
    pencil::make_stress_scene(P, h->ps);
    h->refresh();
    return h.release();
} PCLH_CATCH_PTR(nullptr, 0)

// mat_mix = 1 writes the batched workload's scene (pclh_make_mixed).
int pclh_write_stress(const char* dir, const char* name, int freq, int n_spheres, uint32_t seed, int width,
                      int height, int mat_mix, char* err, int errlen) try {
    pencil::StressParams P;
    P.freq = freq; P.n_spheres = n_spheres; P.seed = seed; P.width = width; P.height = height;
    P.mat_mix = mat_mix;
    std::string e;
    if (!pencil::write_stress_scene(P, dir, name, e)) { copy_err(e, err, errlen); return -1; }
    return 0;
} PCLH_CATCH_INT(err, errlen)

// The path-tracing workload: a Cornell box .pecl with the displaced
// icosphere (frequency freq, seed) placed inside at v * scale + (cx, cy, cz).
void* pclh_make_cornell_mesh(const char* pecl, int freq, uint32_t seed, float scale, float cx, float cy, float cz,
                             int width, int height, char* err, int errlen) try {
    void* hv = pclh_load_pecl(pecl, width, height, 0, 0, err, errlen);
    if (!hv) return nullptr;
    auto* h = static_cast<Handle*>(hv);
    pencil::StressParams P;
    P.freq = freq; P.seed = seed;
    P.mesh_scale = scale; P.mesh_offset[0] = cx; P.mesh_offset[1] = cy; P.mesh_offset[2] = cz;
    pencil::add_mesh(P, h->ps);
    h->refresh();
    return h;
} PCLH_CATCH_PTR(err, errlen)

// pclh_make_cornell_mesh's scene as files the reference parses: <dir>/<name>.pecl
// (the base .pecl plus a mesh{} block) and <dir>/<name>.obj.
int pclh_write_cornell_mesh(const char* pecl, const char* dir, const char* name, int freq, uint32_t seed,
                            float scale, float cx, float cy, float cz, char* err, int errlen) try {
    pencil::StressParams P;
    P.freq = freq; P.seed = seed;
    P.mesh_scale = scale; P.mesh_offset[0] = cx; P.mesh_offset[1] = cy; P.mesh_offset[2] = cz;
    std::string e;
    if (!pencil::write_mesh_in_scene(P, pecl, dir, name, e)) { copy_err(e, err, errlen); return -1; }
    return 0;
} PCLH_CATCH_INT(err, errlen)

// One scene of the batched workload: a displaced icosphere of frequency freq
// (radius 3) and n_spheres seeded spheres with a diffuse / metal / glass mix,
// the stress scene's camera and a soft point light.
void* pclh_make_mixed(int freq, int n_spheres, uint32_t seed, int width, int height) try {
    auto h = std::make_unique<Handle>();
    pencil::StressParams P;
    P.freq = freq; P.n_spheres = n_spheres; P.seed = seed; P.width = width; P.height = height; P.mat_mix = 1;
    pencil::make_stress_scene(P, h->ps);
    h->ps.scene.lights_mut()[0].radius = 0.5f;
    h->refresh();
    return h.release();
} PCLH_CATCH_PTR(nullptr, 0)

// The drop-in entry itself: pencil::Renderer::render (host/renderer.hpp, the
// reference's src/renderer.hpp:198 signature) on the scene with settings s
// and an optional --bokeh override, backend hip on device 0.  out_rgb:
// W*H*3 floats.  Returns 0, or -1 with the renderer's message in err.
int pclh_renderer_render(void* hv, const pcl_settings* s, int bokeh_blades, float* out_rgb, uint64_t* counts,
                         char* err, int errlen) try {
    auto* h = static_cast<Handle*>(hv);
    if (!h || !s || !out_rgb) return -1;
    pencil::RenderSettings rs;
    rs.width = s->width; rs.height = s->height; rs.samples = s->samples; rs.max_depth = s->max_depth;
    rs.shadow_bias = s->shadow_bias; rs.soft_shadow_samples = s->soft_shadow_samples;
    rs.pathtrace = s->pathtrace != 0; rs.mis = s->pathtrace == 2; rs.pathtrace_bounces = s->pathtrace_bounces;
    rs.bokeh_blades = bokeh_blades;
    rs.backend = pencil::Backend::Hip;
    pencil::Renderer r(h->ps.scene, h->ps.camera, rs);
    std::vector<pencil::Color> px;
    pencil::RenderStats st;
    if (!r.render(px, &st)) { copy_err(r.error(), err, errlen); return -1; }
    std::memcpy(out_rgb, px.data(), px.size() * sizeof(pencil::Color));
    if (counts) { counts[0] = st.rays_traced; counts[1] = st.closest_queries; counts[2] = st.anyhit_queries; }
    return 0;
} PCLH_CATCH_INT(err, errlen)

void pclh_free(void* h) { delete static_cast<Handle*>(h); }

// Pointers inside *out stay valid until pclh_free or the next material edit.
int pclh_desc(void* hv, pcl_scene_desc* out) {
    if (!hv || !out) return -1;
    *out = static_cast<Handle*>(hv)->desc;
    return 0;
}

int pclh_camera(void* hv, pcl_camera* out) {
    if (!hv || !out) return -1;
    *out = static_cast<Handle*>(hv)->ps.camera.args();
    return 0;
}

// RenderSettings as the CLI builds them (src/pencil.cpp:739-762).
int pclh_settings(void* hv, pcl_settings* out) {
    if (!hv || !out) return -1;
    const auto& ps = static_cast<Handle*>(hv)->ps;
    pencil::RenderSettings rs;
    *out = pcl_settings{};
    out->width = ps.width; out->height = ps.height; out->samples = ps.samples; out->max_depth = ps.depth;
    out->shadow_bias = rs.shadow_bias; out->soft_shadow_samples = rs.soft_shadow_samples;
    out->pathtrace = ps.enable_pathtrace ? (ps.enable_mis ? 2 : 1) : 0; out->pathtrace_bounces = rs.pathtrace_bounces;
    return 0;
}

int pclh_unsupported(void* hv, char* buf, int len) {
    if (!hv) return -1;
    const std::string& u = static_cast<Handle*>(hv)->unsupported;
    copy_err(u, buf, len);
    return (int)u.size();
}

int pclh_n_materials(void* hv) { return hv ? static_cast<Handle*>(hv)->ps.scene.material_count() : -1; }

// Overwrite albedo (which = 0) or emission (which = 1) of material m.
int pclh_set_material_color(void* hv, int m, int which, const float* rgb) {
    auto* h = static_cast<Handle*>(hv);
    if (!h || m < 0 || m >= h->ps.scene.material_count() || which < 0 || which > 1) return -1;
    auto& mat = h->ps.scene.materials_mut()[m];
    pencil::Color& c = which == 0 ? mat.albedo : mat.emission;
    c = {rgb[0], rgb[1], rgb[2]};
    h->refresh();
    return 0;
}

// color_to_rgb8 over an HDR buffer (ACES, given exposure).
int pclh_tonemap_rgb8(const float* hdr, int64_t npx, float exposure, uint8_t* out) {
    for (int64_t i = 0; i < npx; ++i)
        pencil::color_to_rgb8({hdr[i * 3], hdr[i * 3 + 1], hdr[i * 3 + 2]}, pencil::ToneMap::ACES, exposure, out + i * 3);
    return 0;
}

}  // extern "C"
