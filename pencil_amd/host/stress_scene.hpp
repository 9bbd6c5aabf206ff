// stress_scene.hpp - seeded synthetic scenes for the BVH stress workload.
//
// "bvh_stress" (1920x1080, 1 spp, depth 1, one hard point light: primary +
// one shadow ray per pixel): a geodesic icosphere of frequency F (20*F*F
// triangles; F = 229 gives 1,048,820, the closest icosphere count to 2^20)
// of radius 3, displaced radially by seeded 3-octave value noise of
// amplitude 0.1, plus N random spheres with centers U[-10,10]^3 and radii
// U[0.02,0.2].
//
// The same scene can be written as .pecl + .obj so the reference renderer
// parses exactly the same floats: vertices are printed with 9 significant
// digits (float round trip) and two unreferenced vertices at +-100 make the
// reference's fit_mesh an exact identity (scale 200 / extent 200 = 1, center
// 0; src/mesh.hpp:96-124), colors are integer percentages.
#pragma once

#include "pecl_parser.hpp"
#include "scene.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

namespace pencil {

struct StressParams {
    int freq = 229;          // icosphere frequency: 20*freq^2 triangles
    int n_spheres = 4096;
    uint32_t seed = 1234;
    int width = 1920, height = 1080;
    float radius = 3.0f, amplitude = 0.1f;
    // mesh placement, applied to the displaced vertices: v * mesh_scale + mesh_offset
    double mesh_scale = 1.0, mesh_offset[3] = {0.0, 0.0, 0.0};
    // sphere materials: 0 = all diffuse; 1 = a seeded mix of diffuse, metal
    // (fuzz) and glass spheres
    int mat_mix = 0;
};

namespace stress {

inline uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline double uniform(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// Lattice value in [-1, 1] from an integer hash of (x, y, z, seed).
inline double lattice(int64_t x, int64_t y, int64_t z, uint32_t seed) {
    uint64_t h = (uint64_t)x * 0x8DA6B343ull ^ (uint64_t)y * 0xD8163841ull ^ (uint64_t)z * 0xCB1AB31Full ^
                 (uint64_t)seed * 0x165667B1ull;
    uint64_t s = h;
    return uniform(s) * 2.0 - 1.0;
}

inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }

// Trilinear value noise with smoothstep weights.
inline double value_noise(double x, double y, double z, uint32_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

// 3 octaves, lacunarity 2, gain 0.5, normalized to [-1, 1].
inline double fbm3(double x, double y, double z, uint32_t seed) {
    double sum = 0.0, amp = 1.0, f = 1.5, norm = 0.0;
    for (int o = 0; o < 3; ++o) {
        sum += amp * value_noise(x * f, y * f, z * f, seed + 101u * (uint32_t)o);
        norm += amp;
        amp *= 0.5;
        f *= 2.0;
    }
    return sum / norm;
}

struct D3 { double x, y, z; };

// Displaced geodesic icosphere: vertices (deduplicated) and triangles.
inline void icosphere(const StressParams& P, std::vector<Vec3>& verts, std::vector<int32_t>& tris) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const D3 base[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                              {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                              {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    const int F = P.freq;
    std::unordered_map<uint64_t, int32_t> index;
    index.reserve((size_t)10 * F * F + 16);
    auto emit = [&](uint64_t key, D3 q) -> int32_t {
        auto it = index.find(key);
        if (it != index.end()) return it->second;
        double l = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
        double nx = q.x / l, ny = q.y / l, nz = q.z / l;
        double r = P.radius + P.amplitude * fbm3(nx * 1.0, ny * 1.0, nz * 1.0, P.seed);
        int32_t id = (int32_t)verts.size();
        const double s = P.mesh_scale;
        verts.push_back({(float)(nx * r * s + P.mesh_offset[0]), (float)(ny * r * s + P.mesh_offset[1]),
                         (float)(nz * r * s + P.mesh_offset[2])});
        index.emplace(key, id);
        return id;
    };
    // canonical point on edge (u < v) at k/F from u, or face-interior point
    auto edge_point = [&](int u, int v, int k) -> int32_t {
        if (k == 0) return emit((uint64_t)u, base[u]);
        if (k == F) return emit((uint64_t)v, base[v]);
        if (u > v) { int x = u; u = v; v = x; k = F - k; }
        double s = (double)k / F;
        D3 q{base[u].x + (base[v].x - base[u].x) * s, base[u].y + (base[v].y - base[u].y) * s,
             base[u].z + (base[v].z - base[u].z) * s};
        uint64_t key = (1ull << 60) | ((uint64_t)u << 40) | ((uint64_t)v << 32) | (uint64_t)k;
        return emit(key, q);
    };
    std::vector<int32_t> grid((size_t)(F + 1) * (F + 1));
    for (int f = 0; f < 20; ++f) {
        const int a = faces[f][0], b = faces[f][1], c = faces[f][2];
        for (int i = 0; i <= F; ++i) {
            for (int j = 0; i + j <= F; ++j) {
                int32_t id;
                if (j == 0) id = edge_point(a, b, i);
                else if (i == 0) id = edge_point(a, c, j);
                else if (i + j == F) id = edge_point(b, c, j);
                else {
                    double si = (double)i / F, sj = (double)j / F;
                    D3 q{base[a].x + (base[b].x - base[a].x) * si + (base[c].x - base[a].x) * sj,
                         base[a].y + (base[b].y - base[a].y) * si + (base[c].y - base[a].y) * sj,
                         base[a].z + (base[b].z - base[a].z) * si + (base[c].z - base[a].z) * sj};
                    uint64_t key = (2ull << 60) | ((uint64_t)f << 40) | ((uint64_t)i << 20) | (uint64_t)j;
                    id = emit(key, q);
                }
                grid[(size_t)i * (F + 1) + j] = id;
            }
        }
        for (int i = 0; i < F; ++i) {
            for (int j = 0; i + j < F; ++j) {
                int32_t p00 = grid[(size_t)i * (F + 1) + j], p10 = grid[(size_t)(i + 1) * (F + 1) + j];
                int32_t p01 = grid[(size_t)i * (F + 1) + j + 1];
                tris.push_back(p00); tris.push_back(p10); tris.push_back(p01);
                if (i + j + 2 <= F) {
                    int32_t p11 = grid[(size_t)(i + 1) * (F + 1) + j + 1];
                    tris.push_back(p10); tris.push_back(p11); tris.push_back(p01);
                }
            }
        }
    }
}

// kind: 0 diffuse, 1 metal (fuz percent in `fuz`), 2 glass (ior 1.5)
struct SphereSpec { Vec3 c; float r; int clr[3]; int kind = 0; int fuz = 0; };

inline std::vector<SphereSpec> spheres(const StressParams& P) {
    uint64_t s = 0x5EED0000ull ^ P.seed;
    std::vector<SphereSpec> out;
    for (int k = 0; k < P.n_spheres; ++k) {
        SphereSpec q;
        q.c = {(float)(uniform(s) * 20.0 - 10.0), (float)(uniform(s) * 20.0 - 10.0), (float)(uniform(s) * 20.0 - 10.0)};
        q.r = (float)(0.02 + uniform(s) * 0.18);
        for (int c = 0; c < 3; ++c) q.clr[c] = 20 + (int)(splitmix(s) % 71u);
        if (P.mat_mix) {
            const uint32_t u = (uint32_t)(splitmix(s) % 8u);
            if (u < 2) { q.kind = 1; q.fuz = (int)(splitmix(s) % 30u); }
            else if (u == 2) q.kind = 2;
        }
        out.push_back(q);
    }
    return out;
}

}  // namespace stress

// The seeded spheres with their materials; `mat { clr, fuz | ior }` as the
// parser turns them into materials (src/pecl_parser.hpp:739-762).
inline void add_spheres(const StressParams& P, ParsedScene& ps) {
    for (const auto& s : stress::spheres(P)) {
        const Color c{(float)s.clr[0] / 100.0f, (float)s.clr[1] / 100.0f, (float)s.clr[2] / 100.0f};
        Material m = make_diffuse(c, 0.0f, 32.0f);
        if (s.kind == 1) m = make_metal(c, 1.0f - (float)s.fuz / 100.0f);
        else if (s.kind == 2) m = make_dielectric(c, 1.5f);
        ps.scene.add_material(m);
        ps.scene.add_primitive(make_sphere(s.c, s.r, ps.scene.material_count() - 1));
    }
}

// Appends the displaced icosphere (one diffuse 0.7 material) to a scene.
inline void add_mesh(const StressParams& P, ParsedScene& ps) {
    std::vector<Vec3> verts;
    std::vector<int32_t> tris;
    stress::icosphere(P, verts, tris);
    ps.scene.add_material(make_diffuse({0.7f, 0.7f, 0.7f}, 0.0f, 32.0f));
    const int mat = ps.scene.material_count() - 1;
    for (size_t t = 0; t + 2 < tris.size(); t += 3)
        ps.scene.add_primitive(make_triangle(verts[tris[t]], verts[tris[t + 1]], verts[tris[t + 2]], mat));
}

// Builds the scene in memory with exactly the floats the reference gets from
// the .pecl + .obj that write_stress_scene emits.
inline void make_stress_scene(const StressParams& P, ParsedScene& ps) {
    std::vector<Vec3> verts;
    std::vector<int32_t> tris;
    stress::icosphere(P, verts, tris);
    ps = ParsedScene{};
    // mesh { file, scale: 200 } with the +-100 guard vertices: fit_mesh gives
    // scale 1 and offset -0; the vertex transform below is the parser's.
    Material mesh_mat = make_diffuse({0.7f, 0.7f, 0.7f}, 0.0f, 32.0f);
    ps.scene.add_material(mesh_mat);
    const float scale = 200.0f / 200.0f;
    const Vec3 off = -Vec3{0.0f, 0.0f, 0.0f} * scale;
    const Vec3 user = {0, 0, 0};
    for (size_t t = 0; t + 2 < tris.size(); t += 3) {
        Point3 v0 = verts[tris[t]] * scale + off + user;
        Point3 v1 = verts[tris[t + 1]] * scale + off + user;
        Point3 v2 = verts[tris[t + 2]] * scale + off + user;
        ps.scene.add_primitive(make_triangle(v0, v1, v2, 0));
    }
    add_spheres(P, ps);
    ps.scene.add_light(make_point_light({10.0f, 15.0f, -10.0f}, {1.0f, 1.0f, 1.0f}, 0.0f));
    ps.eye = {0.0f, 4.0f, -18.0f};
    ps.look = {0.0f, 0.0f, 0.0f};
    ps.up = {0.0f, 1.0f, 0.0f};
    ps.fov = 50.0f;
    ps.width = P.width;
    ps.height = P.height;
    ps.samples = 1;
    ps.depth = 1;
    ps.aspect = (float)P.width / (float)P.height;
    ps.finalize();
}

// Writes <dir>/<name>.pecl and <dir>/<name>.obj for the reference renderer.
inline bool write_stress_scene(const StressParams& P, const std::string& dir, const std::string& name,
                               std::string& err) {
    std::vector<Vec3> verts;
    std::vector<int32_t> tris;
    stress::icosphere(P, verts, tris);
    const std::string obj = dir + "/" + name + ".obj", pecl = dir + "/" + name + ".pecl";
    FILE* f = std::fopen(obj.c_str(), "w");
    if (!f) { err = "cannot write " + obj; return false; }
    std::fprintf(f, "# displaced icosphere, frequency %d, seed %u\n", P.freq, P.seed);
    for (const auto& v : verts) std::fprintf(f, "v %.9g %.9g %.9g\n", v.x, v.y, v.z);
    std::fprintf(f, "v -100 -100 -100\nv 100 100 100\n");
    for (size_t t = 0; t + 2 < tris.size(); t += 3)
        std::fprintf(f, "f %d %d %d\n", tris[t] + 1, tris[t + 1] + 1, tris[t + 2] + 1);
    std::fclose(f);
    f = std::fopen(pecl.c_str(), "w");
    if (!f) { err = "cannot write " + pecl; return false; }
    std::fprintf(f, "# bvh stress scene (seed %u)\n", P.seed);
    std::fprintf(f, "cmra {\n  pos: <0, 4, -18>\n  aim: <0, 0, 0>\n  up: <0, 1, 0>\n  fov: 50\n}\n");
    std::fprintf(f, "mat { clr: <70, 70, 70> }\nmesh {\n  file: %s\n  scale: 200\n}\n", obj.c_str());
    for (const auto& s : stress::spheres(P)) {
        if (s.kind == 1)
            std::fprintf(f, "mat { clr: <%d, %d, %d>, fuz: %d }\n", s.clr[0], s.clr[1], s.clr[2], s.fuz);
        else if (s.kind == 2)
            std::fprintf(f, "mat { clr: <%d, %d, %d>, ior: 1.5 }\n", s.clr[0], s.clr[1], s.clr[2]);
        else
            std::fprintf(f, "mat { clr: <%d, %d, %d> }\n", s.clr[0], s.clr[1], s.clr[2]);
        std::fprintf(f, "sphere { pos: <%.9g, %.9g, %.9g>, rad: %.9g }\n", s.c.x, s.c.y, s.c.z, s.r);
    }
    // the batched workload's scenes (mat_mix) use a soft light of radius 0.5
    std::fprintf(f, "light {\n  type: point\n  pos: <10, 15, -10>\n  clr: <100, 100, 100>\n  radius: %s\n}\n",
                 P.mat_mix ? "0.5" : "0");
    std::fprintf(f, "render {\n  size: %d %d\n  spp: 1\n  depth: 1\n}\n", P.width, P.height);
    std::fclose(f);
    return true;
}

// Writes <dir>/<name>.pecl -- the text of the .pecl at `base` followed by a
// diffuse 0.7 material and a mesh{} block -- and <dir>/<name>.obj with the
// displaced icosphere placed as add_mesh places it (v * mesh_scale +
// mesh_offset), plus the +-100 guard vertices, so the reference parses the
// scene add_mesh builds on top of `base`, float for float.
inline bool write_mesh_in_scene(const StressParams& P, const std::string& base, const std::string& dir,
                                const std::string& name, std::string& err) {
    std::vector<Vec3> verts;
    std::vector<int32_t> tris;
    stress::icosphere(P, verts, tris);
    const std::string obj = dir + "/" + name + ".obj", pecl = dir + "/" + name + ".pecl";
    FILE* in = std::fopen(base.c_str(), "rb");
    if (!in) { err = "cannot read " + base; return false; }
    std::string text;
    char buf[4096];
    for (size_t n; (n = std::fread(buf, 1, sizeof(buf), in)) > 0;) text.append(buf, n);
    std::fclose(in);
    FILE* f = std::fopen(obj.c_str(), "w");
    if (!f) { err = "cannot write " + obj; return false; }
    std::fprintf(f, "# displaced icosphere, frequency %d, seed %u\n", P.freq, P.seed);
    for (const auto& v : verts) std::fprintf(f, "v %.9g %.9g %.9g\n", v.x, v.y, v.z);
    std::fprintf(f, "v -100 -100 -100\nv 100 100 100\n");
    for (size_t t = 0; t + 2 < tris.size(); t += 3)
        std::fprintf(f, "f %d %d %d\n", tris[t] + 1, tris[t + 1] + 1, tris[t + 2] + 1);
    std::fclose(f);
    f = std::fopen(pecl.c_str(), "w");
    if (!f) { err = "cannot write " + pecl; return false; }
    std::fwrite(text.data(), 1, text.size(), f);
    std::fprintf(f, "\nmat { clr: <70, 70, 70> }\nmesh {\n  file: %s\n  scale: 200\n}\n", obj.c_str());
    std::fclose(f);
    return true;
}

}  // namespace pencil
