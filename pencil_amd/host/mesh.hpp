// mesh.hpp - OBJ loader and fit transform (src/mesh.hpp:38-124).
#pragma once

#include "scene.hpp"

#include <array>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace pencil {

struct MeshData {
    std::vector<Point3> vertices;
    std::vector<std::array<Point3, 3>> triangles;
    Point3 bbox_min = {1e30f, 1e30f, 1e30f};
    Point3 bbox_max = {-1e30f, -1e30f, -1e30f};
};

// load_obj (:38-107): v lines, f lines with v / v/vt / v//vn / v/vt/vn tokens,
// negative indices, fan triangulation; the bbox spans every v line.
inline bool load_obj(const std::string& path, MeshData& out, std::string& err) {
    std::ifstream f(path);
    if (!f) { err = "cannot open obj: " + path; return false; }
    std::string line;
    int lineno = 0;
    std::vector<int> idx;
    while (std::getline(f, line)) {
        ++lineno;
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "v") {
            float x, y, z;
            if (!(ss >> x >> y >> z)) { err = "bad v line " + std::to_string(lineno); return false; }
            out.vertices.push_back({x, y, z});
        } else if (tag == "f") {
            idx.clear();
            std::string tok;
            while (ss >> tok) {
                size_t i = 0;
                while (i < tok.size() && (tok[i] == '-' || std::isdigit((unsigned char)tok[i]))) ++i;
                if (i == 0) { err = "bad face token at line " + std::to_string(lineno); return false; }
                int v = std::atoi(tok.c_str());
                if (v < 0) v = (int)out.vertices.size() + v + 1;
                if (v < 1 || v > (int)out.vertices.size()) {
                    err = "invalid vertex index " + std::to_string(v) + " at line " + std::to_string(lineno);
                    return false;
                }
                idx.push_back(v - 1);
            }
            if (idx.size() < 3) continue;
            for (size_t i = 1; i + 1 < idx.size(); ++i)
                out.triangles.push_back({out.vertices[idx[0]], out.vertices[idx[i]], out.vertices[idx[i + 1]]});
        }
    }
    for (const auto& v : out.vertices) {
        out.bbox_min = {std::fmin(out.bbox_min.x, v.x), std::fmin(out.bbox_min.y, v.y), std::fmin(out.bbox_min.z, v.z)};
        out.bbox_max = {std::fmax(out.bbox_max.x, v.x), std::fmax(out.bbox_max.y, v.y), std::fmax(out.bbox_max.z, v.z)};
    }
    return true;
}

struct MeshTransform {
    float scale = 1.0f;
    Vec3 offset = {0, 0, 0};
};

// fit_mesh (:115-124)
inline MeshTransform fit_mesh(const MeshData& m, float target_size) {
    Vec3 ext = m.bbox_max - m.bbox_min;
    float longest = std::fmax(std::fmax(ext.x, ext.y), ext.z);
    if (longest <= 0.0f) return {};
    MeshTransform t;
    t.scale = target_size / longest;
    Vec3 center = (m.bbox_min + m.bbox_max) * 0.5f;
    t.offset = -center * t.scale;
    return t;
}

}  // namespace pencil
