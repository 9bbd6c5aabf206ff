// pencil.cpp - command line front end, a drop-in for the reference CLI's
// render path (src/pencil.cpp:186-269, 635-790) with one new flag:
//   --backend hip   (the default here; `cpu` is the reference binary itself)
// Flags the HIP core has no renderer for (--bdpt, --volume, fog,
// bloom, denoise, ...) are refused with exit code 2 instead of being ignored.
#include "image_io.hpp"
#include "pecl_parser.hpp"
#include "renderer.hpp"

#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>

namespace {

std::string with_commas(uint64_t v) {
    std::string s = std::to_string(v), out;
    int n = 0;
    for (auto it = s.rbegin(); it != s.rend(); ++it) {
        if (n && n % 3 == 0) out.push_back(',');
        out.push_back(*it);
        ++n;
    }
    return std::string(out.rbegin(), out.rend());
}

void usage() {
    std::cout << "usage: pencil <scene.pecl> [out.png|out.ppm] [options]\n"
                 "  --backend hip|cpu  render core (hip: this library)\n"
                 "  --device N         HIP device\n"
                 "  --width N --height N --samples N --depth N --bounces N\n"
                 "  --pathtrace        path tracing with next-event estimation\n"
                 "  --mis              with --pathtrace: MIS-weighted light samples (trace_path_mis)\n"
                 "  --no-bvh           linear primitive scan\n"
                 "  --bokeh N          polygonal aperture with N blades (0 = the scene's)\n"
                 "  --exposure F --tonemap 0|1|2 --quiet\n";
}

}  // namespace

int main(int argc, char** argv) {
    std::string scene_path, out_path, backend = "hip";
    int ow = 0, oh = 0, os = 0, od = 0, bounces = 4, device = 0, tonemap = 1, bokeh = 0;
    bool quiet = false, pathtrace = false, no_bvh = false, mis = false;
    float exposure = 0.0f;
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) { std::cerr << "missing value for " << s << "\n"; std::exit(2); }
            return argv[++i];
        };
        if (s == "--help" || s == "-h") { usage(); return 0; }
        else if (s == "--backend") backend = next();
        else if (s == "--device") device = std::atoi(next());
        else if (s == "--width") ow = std::atoi(next());
        else if (s == "--height") oh = std::atoi(next());
        else if (s == "--samples") os = std::atoi(next());
        else if (s == "--depth") od = std::atoi(next());
        else if (s == "--bounces") bounces = std::atoi(next());
        else if (s == "--threads") next();   // host threads do not apply to the HIP core
        else if (s == "--pathtrace" || s == "--pt") pathtrace = true;
        else if (s == "--mis") mis = true;
        else if (s == "--no-bvh" || s == "--nobvh") no_bvh = true;
        else if (s == "--bokeh") bokeh = std::atoi(next());   // src/pencil.cpp:218
        else if (s == "--quiet") quiet = true;
        else if (s == "--exposure") exposure = std::strtof(next(), nullptr);
        else if (s == "--tonemap") tonemap = std::atoi(next());
        else if (s.rfind("--", 0) == 0) {
            std::cerr << "pencil: option " << s << " needs a renderer outside the HIP core\n";
            return 2;
        } else if (scene_path.empty()) scene_path = s;
        else if (out_path.empty()) out_path = s;
        else { std::cerr << "unknown argument: " << s << "\n"; return 2; }
    }
    if (scene_path.empty()) { usage(); return 0; }
    if (backend != "hip") {
        std::cerr << "pencil: backend '" << backend << "' is not in this library (the CPU path is the reference renderer)\n";
        return 2;
    }
    pencil::ParsedScene ps;
    std::vector<pencil::ParseError> errs;
    if (!pencil::parse_pecl(scene_path, ps, errs)) {
        std::cerr << "failed to parse scene: " << scene_path << "\n";
        for (const auto& e : errs) std::cerr << "  line " << e.line << ": " << e.what << "\n";
        return 3;
    }
    if (!ps.unsupported.empty()) {
        std::cerr << "pencil: scene needs features outside the HIP core:";
        for (auto& u : ps.unsupported) std::cerr << " [" << u << "]";
        std::cerr << "\n";
        return 2;
    }
    if (ow > 0) { ps.width = ow; ps.aspect = (float)ps.width / (float)ps.height; ps.finalize(); }
    if (oh > 0) { ps.height = oh; ps.aspect = (float)ps.width / (float)ps.height; ps.finalize(); }
    if (os > 0) ps.samples = os;
    if (od > 0) ps.depth = od;
    if (no_bvh) ps.scene.use_bvh = false;
    if (out_path.empty()) {
        size_t dot = scene_path.find_last_of('.');
        out_path = (dot == std::string::npos ? scene_path : scene_path.substr(0, dot)) + ".ppm";
    }
    pencil::RenderSettings rs;
    rs.width = ps.width; rs.height = ps.height; rs.samples = ps.samples; rs.max_depth = ps.depth;
    rs.pathtrace = pathtrace || ps.enable_pathtrace;
    rs.mis = mis || ps.enable_mis;   // src/pencil.cpp:749
    rs.pathtrace_bounces = bounces;
    rs.bokeh_blades = bokeh;
    rs.backend = pencil::Backend::Hip;
    rs.device = device;
    if (!quiet) {
        std::cout << "pencil (hip core) - " << scene_path << "\n"
                  << "  size:    " << ps.width << " x " << ps.height << "\n"
                  << "  aa:      " << ps.samples << "x" << ps.samples << " spp\n"
                  << "  depth:   " << (rs.pathtrace ? bounces : ps.depth) << (rs.pathtrace ? " bounces" : "") << "\n"
                  << "  lights:  " << ps.scene.light_count() << "\n"
                  << "  prims:   " << ps.scene.primitive_count() << "\n";
    }
    pencil::Renderer renderer(ps.scene, ps.camera, rs);
    std::vector<pencil::Color> pixels;
    pencil::RenderStats stats;
    auto t0 = std::chrono::steady_clock::now();
    if (!renderer.render(pixels, &stats)) return 5;
    auto t1 = std::chrono::steady_clock::now();
    double sec = std::chrono::duration<double>(t1 - t0).count();
    if (!quiet) {
        std::cout << "rendered in " << sec << " s (" << with_commas(stats.rays_traced) << " rays, "
                  << (sec > 1e-9 ? stats.rays_traced / sec / 1e6 : 0.0) << " Mrays/s; kernel "
                  << stats.kernel_ms << " ms, " << with_commas(stats.closest_queries + stats.anyhit_queries)
                  << " queries)\n";
    }
    std::string err;
    pencil::ToneMap tm = tonemap == 0 ? pencil::ToneMap::None : tonemap == 2 ? pencil::ToneMap::Reinhard : pencil::ToneMap::ACES;
    if (!pencil::write_image(out_path, pixels, rs.width, rs.height, err, tm, exposure)) {
        std::cerr << "error: " << err << "\n";
        return 4;
    }
    if (!quiet) std::cout << "wrote " << out_path << "\n";
    return 0;
}
