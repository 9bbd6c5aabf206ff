// scene.hpp - host-side scene model of the Pencil API, feeding the HIP core.
//
// Mirrors the reference's data contract (src/vec3.hpp, src/material.hpp,
// src/light.hpp, src/scene.hpp, src/camera.hpp) with the same names and
// defaults, but stores primitives as constructor-argument records
// (pcl_prim) in insertion order: the BVH, intersection and shading live on the
// device behind include/pencil_hip.h.
#pragma once

#include "pencil_hip.h"

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace pencil {

// src/vec3.hpp:24-92 (host arithmetic only; same left-to-right dot order)
struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    constexpr Vec3() = default;
    constexpr Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec3 operator-() const { return {-x, -y, -z}; }
    Vec3 operator*(float s) const { return {x * s, y * s, z * s}; }
    Vec3 operator/(float s) const { return {x / s, y / s, z / s}; }
    Vec3 operator*(const Vec3& o) const { return {x * o.x, y * o.y, z * o.z}; }
    Vec3& operator+=(const Vec3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    float dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
    Vec3 cross(const Vec3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
    float length_squared() const { return x * x + y * y + z * z; }
    float length() const { return std::sqrt(length_squared()); }
    Vec3 normalized() const { float l = length(); return (l > 1e-12f) ? (*this / l) : Vec3{0, 0, 0}; }
};
using Point3 = Vec3;
using Color = Vec3;

inline float clampf(float v, float lo, float hi) { return (v < lo) ? lo : (v > hi) ? hi : v; }

// src/material.hpp:17-119
enum class MaterialKind : int { Diffuse = 0, Metal = 1, Dielectric = 2, Principled = 3 };

struct Material {
    MaterialKind kind = MaterialKind::Diffuse;
    Color albedo = {0.8f, 0.8f, 0.8f};
    float reflectivity = 0.0f;
    float ior = 1.0f;
    float specular = 0.0f;
    float shininess = 32.0f;
    Color specular_color = {1.0f, 1.0f, 1.0f};
    Color emission = {0.0f, 0.0f, 0.0f};
    // Features outside the HIP core; the parser records them so the core can
    // refuse a scene that needs them instead of rendering it wrongly.
    int refdis = 0;
    bool has_texture = false;
};

inline Material make_diffuse(Color albedo, float spec = 0.0f, float shininess = 32.0f,
                             Color spec_color = {1.0f, 1.0f, 1.0f}) {
    Material m;
    m.kind = MaterialKind::Diffuse; m.albedo = albedo; m.specular = spec;
    m.shininess = shininess; m.specular_color = spec_color;
    return m;
}

inline Material make_metal(Color albedo, float reflectivity) {
    Material m;
    m.kind = MaterialKind::Metal; m.albedo = albedo; m.reflectivity = reflectivity;
    return m;
}
inline Material make_dielectric(Color albedo, float ior) {
    Material m;
    m.kind = MaterialKind::Dielectric; m.albedo = albedo; m.ior = ior;
    return m;
}

// src/light.hpp:15-46
enum class LightKind : int { Point = 0, Directional = 1 };

struct Light {
    LightKind kind = LightKind::Point;
    Vec3 position_or_direction = {0.0f, 5.0f, 0.0f};
    Color intensity = {1.0f, 1.0f, 1.0f};
    float radius = 0.0f;
    bool is_emissive = false;
};

inline Light make_point_light(Vec3 pos, Color intensity, float radius = 0.0f, bool is_emissive = false) {
    Light l; l.kind = LightKind::Point; l.position_or_direction = pos;
    l.intensity = intensity; l.radius = radius; l.is_emissive = is_emissive;
    return l;
}
inline Light make_directional_light(Vec3 dir, Color intensity) {
    Light l; l.kind = LightKind::Directional; l.position_or_direction = dir.normalized();
    l.intensity = intensity; l.radius = 0.0f;
    return l;
}

// Primitive constructors (src/primitives.hpp) recorded as pcl_prim records.
inline pcl_prim make_sphere(Point3 c, float r, int mat) {
    pcl_prim p{}; p.kind = PCL_PRIM_SPHERE; p.material_id = mat;
    p.p[0] = c.x; p.p[1] = c.y; p.p[2] = c.z; p.p[3] = r; return p;
}
inline pcl_prim make_triangle(Point3 a, Point3 b, Point3 c, int mat) {
    pcl_prim p{}; p.kind = PCL_PRIM_TRIANGLE; p.material_id = mat;
    const Point3 v[3] = {a, b, c};
    for (int i = 0; i < 3; ++i) { p.p[i * 3] = v[i].x; p.p[i * 3 + 1] = v[i].y; p.p[i * 3 + 2] = v[i].z; }
    return p;
}
inline pcl_prim make_cube(Point3 c, float half, int mat) {
    pcl_prim p{}; p.kind = PCL_PRIM_CUBE; p.material_id = mat;
    p.p[0] = c.x; p.p[1] = c.y; p.p[2] = c.z; p.p[3] = half; return p;
}
inline pcl_prim make_plane(Point3 pt, Vec3 n, int mat) {
    pcl_prim p{}; p.kind = PCL_PRIM_PLANE; p.material_id = mat;
    p.p[0] = pt.x; p.p[1] = pt.y; p.p[2] = pt.z; p.p[3] = n.x; p.p[4] = n.y; p.p[5] = n.z; return p;
}
inline pcl_prim make_axial(int kind, Point3 c, Vec3 axis, float radius, float height, int mat) {
    pcl_prim p{}; p.kind = kind; p.material_id = mat;
    p.p[0] = c.x; p.p[1] = c.y; p.p[2] = c.z; p.p[3] = axis.x; p.p[4] = axis.y; p.p[5] = axis.z;
    p.p[6] = radius; p.p[7] = height; return p;
}

// src/scene.hpp:21-105
class Scene {
public:
    void add_material(const Material& m) { materials_.push_back(m); }
    void add_light(const Light& l) { lights_.push_back(l); }
    void add_primitive(const pcl_prim& p) { primitives_.push_back(p); }
    void pop_primitive() { if (!primitives_.empty()) primitives_.pop_back(); }

    int material_count() const { return (int)materials_.size(); }
    int light_count() const { return (int)lights_.size(); }
    size_t primitive_count() const { return primitives_.size(); }
    const std::vector<Material>& materials() const { return materials_; }
    std::vector<Material>& materials_mut() { return materials_; }
    std::vector<Light>& lights_mut() { return lights_; }
    const std::vector<Light>& lights() const { return lights_; }
    const std::vector<pcl_prim>& primitives() const { return primitives_; }
    const Material& material(int id) const { return materials_[id]; }

    bool use_bvh = true;
    Color background = {0.05f, 0.06f, 0.09f};
    Color ambient = {0.06f, 0.06f, 0.08f};

    // Flatten into the C ABI description.  The arrays referenced by the
    // returned struct live in this Scene (and `mats`/`lights` scratch).
    pcl_scene_desc desc(std::vector<pcl_material>& mats, std::vector<pcl_light>& lights) const {
        mats.clear(); lights.clear();
        for (const auto& m : materials_) {
            pcl_material d{};
            d.kind = (int)m.kind;
            for (int c = 0; c < 3; ++c) {
                d.albedo[c] = m.albedo[c]; d.specular_color[c] = m.specular_color[c]; d.emission[c] = m.emission[c];
            }
            d.reflectivity = m.reflectivity; d.ior = m.ior; d.specular = m.specular; d.shininess = m.shininess;
            // features outside the core are forwarded as an unsupported kind
            if (m.refdis != 0 || m.has_texture) d.kind = PCL_MAT_PRINCIPLED;
            mats.push_back(d);
        }
        for (const auto& l : lights_) {
            pcl_light d{};
            d.kind = (int)l.kind;
            for (int c = 0; c < 3; ++c) { d.pos_or_dir[c] = l.position_or_direction[c]; d.intensity[c] = l.intensity[c]; }
            d.radius = l.radius; d.is_emissive = l.is_emissive ? 1 : 0;
            lights.push_back(d);
        }
        pcl_scene_desc s{};
        s.prims = primitives_.data(); s.n_prims = (int32_t)primitives_.size();
        s.materials = mats.data(); s.n_materials = (int32_t)mats.size();
        s.lights = lights.data(); s.n_lights = (int32_t)lights.size();
        for (int c = 0; c < 3; ++c) { s.background[c] = background[c]; s.ambient[c] = ambient[c]; }
        s.use_bvh = use_bvh ? 1 : 0;
        return s;
    }

private:
    std::vector<Material> materials_;
    std::vector<Light> lights_;
    std::vector<pcl_prim> primitives_;
};

// src/camera.hpp:25-98: the look-at / frustum / DOF math runs on the device;
// the host keeps the constructor arguments.
class Camera {
public:
    Camera() = default;
    Camera(Point3 eye, Point3 look, Vec3 up_hint, float vfov_deg, float aspect,
           float aperture = 0.0f, float focus_dist = 1.0f) {
        for (int c = 0; c < 3; ++c) { args_.eye[c] = eye[c]; args_.look[c] = look[c]; args_.up[c] = up_hint[c]; }
        args_.vfov_deg = vfov_deg; args_.aspect = aspect; args_.aperture = aperture;
        args_.focus_dist = focus_dist; args_.bokeh_blades = 0;
    }
    void set_aspect(float a) { args_.aspect = a; }
    void set_bokeh_blades(int n) { args_.bokeh_blades = n; }
    int bokeh_blades() const { return args_.bokeh_blades; }
    float aperture() const { return args_.aperture; }
    float focus_dist() const { return args_.focus_dist; }
    const pcl_camera& args() const { return args_; }

private:
    pcl_camera args_{{0, 0, 0}, {0, 0, -1}, {0, 1, 0}, 60.0f, 1.0f, 0.0f, 1.0f, 0};
};

}  // namespace pencil
