// pecl_parser.hpp - .pecl scene parser (the subset the HIP core renders).
//
// Same grammar, keys, defaults and per-block material semantics as the
// reference parser (src/pecl_parser.hpp:38-1894): a tokenizer of braces,
// colons, commas, angle brackets, identifiers and numbers with '#' comments
// (:166-284); `<x, y, z>` vectors (:335-361); 0-100 colors (:405-419);
// material patches applied onto the current material (:487-781); one
// material per primitive block (:1619-1838); `render {}` settings
// (:1268-1388); terrain grids of triangles (:1180-1207, :1541-1576).  Blocks
// whose features live outside the core (fog, sky, implicit surfaces,
// textures, presets, dispersion, ...) are recorded in
// ParsedScene::unsupported so the renderer can refuse them.
#pragma once

#include "mesh.hpp"
#include "noise.hpp"
#include "scene.hpp"

#include <cctype>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace pencil {

struct ParsedScene {
    Scene scene;
    Camera camera;

    Point3 eye = {0.0f, 1.5f, -5.0f};
    Point3 look = {0.0f, 0.0f, 0.0f};
    Vec3 up = {0.0f, 1.0f, 0.0f};
    float fov = 60.0f;
    float aspect = 4.0f / 3.0f;
    float aperture = 0.0f;
    float focus_dist = 1.0f;
    int bokeh_blades = 0;

    int width = 400;
    int height = 300;
    int samples = 4;
    int depth = 8;
    bool enable_pathtrace = false;
    bool enable_mis = false;     // render { mis: true } (src/pecl_parser.hpp, rs.mis)

    // Features the reference supports that the HIP core does not render.
    std::vector<std::string> unsupported;

    // scratch fields reused while parsing each primitive block (:85-97)
    Vec3 _tmp_v1 = {0, 0, 0}, _tmp_v2 = {0, 0, 0}, _tmp_v3 = {0, 0, 0}, _tmp_v4 = {0, 0, 0};
    float _tmp_f1 = 1.0f, _tmp_f2 = 1.0f;
    int _tmp_i1 = 0, _tmp_i2 = 32;   // terrain octaves / grid (:86-87); the grid survives the
                                     // end-of-block reset (:1847-1851), the octaves do not
    std::string _pending_seed;       // terrain seed text (:91), cleared by the terrain block
    std::string _pending_mesh_file, _pending_include, _base_dir;
    bool _skip_finalize = false;
    Material _current_mat = make_diffuse({0.7f, 0.7f, 0.7f}, 0.0f, 32.0f);
    std::vector<Vec3> _group_offsets;

    Vec3 group_offset() const {
        Vec3 t{0, 0, 0};
        for (const auto& g : _group_offsets) t += g;
        return t;
    }

    // ParsedScene::finalize (:136-144); the BVH is built by the core on upload.
    void finalize() {
        if (focus_dist <= 0.0f) focus_dist = (eye - look).length();
        camera = Camera(eye, look, up, fov, aspect, aperture, focus_dist);
        camera.set_bokeh_blades(bokeh_blades);
    }
};

struct ParseError {
    int line = 0;
    std::string what;
};

enum class TokKind { LBrace, RBrace, Colon, Comma, AngleOpen, AngleClose, Ident, Number, End };

struct Token {
    TokKind kind = TokKind::End;
    std::string text;
    int line = 0;
};

class Lexer {
public:
    explicit Lexer(const std::string& src) : src_(src) {}

    Token next() {
        skip();
        if (pos_ >= src_.size()) return {TokKind::End, "", line_};
        char c = src_[pos_];
        switch (c) {
        case '{': pos_++; return {TokKind::LBrace, "{", line_};
        case '}': pos_++; return {TokKind::RBrace, "}", line_};
        case ':': pos_++; return {TokKind::Colon, ":", line_};
        case ',': pos_++; return {TokKind::Comma, ",", line_};
        case '<': pos_++; return {TokKind::AngleOpen, "<", line_};
        case '>': pos_++; return {TokKind::AngleClose, ">", line_};
        default: break;
        }
        bool dig_next = pos_ + 1 < src_.size() &&
                        (std::isdigit((unsigned char)src_[pos_ + 1]) || src_[pos_ + 1] == '.');
        if (std::isdigit((unsigned char)c) || c == '.' || (c == '-' && dig_next)) return number();
        if (std::isalpha((unsigned char)c) || c == '_') {
            size_t s = pos_;
            while (pos_ < src_.size() && (std::isalnum((unsigned char)src_[pos_]) || src_[pos_] == '_')) pos_++;
            return {TokKind::Ident, src_.substr(s, pos_ - s), line_};
        }
        pos_++;
        return {TokKind::Ident, std::string(1, c), line_};
    }

    // A file-path value: a double-quoted string (everything up to the next
    // quote, which is consumed) or a bare run ending before the first newline,
    // comma or closing brace; carriage returns are dropped and trailing blanks
    // trimmed in both forms (the reference grammar's path values).
    std::string read_path() {
        skip();
        const bool quoted = pos_ < src_.size() && src_[pos_] == '"';
        if (quoted) ++pos_;
        const size_t end = quoted ? src_.find('"', pos_) : src_.find_first_of("\n,}", pos_);
        const size_t stop = end == std::string::npos ? src_.size() : end;
        std::string out;
        for (size_t k = pos_; k < stop; ++k)
            if (src_[k] != '\r') out += src_[k];
        pos_ = (quoted && end != std::string::npos) ? end + 1 : stop;
        const size_t last = out.find_last_not_of(" \t");
        out.erase(last == std::string::npos ? 0 : last + 1);
        return out;
    }
    int line() const { return line_; }

private:
    const std::string& src_;
    size_t pos_ = 0;
    int line_ = 1;

    void skip() {
        while (pos_ < src_.size()) {
            char c = src_[pos_];
            if (c == ' ' || c == '\t' || c == '\r') pos_++;
            else if (c == '\n') { pos_++; line_++; }
            else if (c == '#') { while (pos_ < src_.size() && src_[pos_] != '\n') pos_++; }
            else break;
        }
    }
    Token number() {
        size_t s = pos_;
        if (src_[pos_] == '-') pos_++;
        bool dot = false, ex = false;
        while (pos_ < src_.size()) {
            char c = src_[pos_];
            if (std::isdigit((unsigned char)c)) pos_++;
            else if (c == '.' && !dot) { dot = true; pos_++; }
            else if ((c == 'e' || c == 'E') && !ex) {
                ex = true; pos_++;
                if (pos_ < src_.size() && (src_[pos_] == '+' || src_[pos_] == '-')) pos_++;
            } else break;
        }
        return {TokKind::Number, src_.substr(s, pos_ - s), line_};
    }
};

inline bool parse_vec3_strict(Lexer& lx, Vec3& out, ParseError& err) {
    Token t = lx.next();
    if (t.kind != TokKind::AngleOpen) { err = {t.line, "expected '<' for vector, got '" + t.text + "'"}; return false; }
    float v[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        Token n = lx.next();
        if (n.kind != TokKind::Number) { err = {n.line, "expected number in vector"}; return false; }
        v[i] = std::stof(n.text);
        if (i < 2) {
            Token s = lx.next();
            if (s.kind != TokKind::Comma) { err = {s.line, "expected ',' between vector components"}; return false; }
        }
    }
    Token c = lx.next();
    if (c.kind != TokKind::AngleClose) { err = {c.line, "expected '>' to close vector"}; return false; }
    out = {v[0], v[1], v[2]};
    return true;
}

inline bool parse_scalar(Lexer& lx, float& out, ParseError& err) {
    Token t = lx.next();
    if (t.kind != TokKind::Number) { err = {t.line, "expected number, got '" + t.text + "'"}; return false; }
    out = std::stof(t.text);
    return true;
}

inline bool parse_size2(Lexer& lx, int& a, int& b, ParseError& err) {
    Token t1 = lx.next();
    if (t1.kind != TokKind::Number) { err = {t1.line, "expected two numbers, got '" + t1.text + "'"}; return false; }
    Token t2 = lx.next();
    if (t2.kind != TokKind::Number) { err = {t2.line, "expected second number, got '" + t2.text + "'"}; return false; }
    a = (int)std::stof(t1.text); b = (int)std::stof(t2.text);
    return true;
}

inline bool parse_ident(Lexer& lx, std::string& out, ParseError& err) {
    Token t = lx.next();
    if (t.kind != TokKind::Ident) { err = {t.line, "expected identifier, got '" + t.text + "'"}; return false; }
    out = t.text;
    return true;
}

// :405-419
inline Color color_from_100(const Vec3& v) {
    return {clampf(v.x, 0.0f, 100.0f) / 100.0f, clampf(v.y, 0.0f, 100.0f) / 100.0f, clampf(v.z, 0.0f, 100.0f) / 100.0f};
}
inline Color emission_from_100(const Vec3& v) { return {v.x / 100.0f, v.y / 100.0f, v.z / 100.0f}; }

// :421-465 (fields the core renders, plus flags for the rest)
struct MaterialPatch {
    bool has_clr = false; Color clr;
    bool has_emt = false; Color emt;
    bool has_rfl = false; float rfl = 0;
    bool has_rfr = false; float rfr = 0;
    bool has_ior = false; float ior = 0;
    bool has_spc = false; float spc = 0;
    bool has_shn = false; float shn = 0;
    bool has_fuz = false; float fuz = 0;
    bool has_typ = false; std::string typ;
    bool has_radius = false; float radius = 0;
    bool has_refdis = false; int refdis = 0;
    bool has_principled = false;
    std::string outside;   // first key needing a feature outside the core
};

// try_parse_material_key, :487-709
inline bool try_parse_material_key(const std::string& key, Lexer& lx, MaterialPatch& p, ParseError& err) {
    auto scalar100 = [&](bool& has, float& dst) {
        float f; if (!parse_scalar(lx, f, err)) return false;
        has = true; dst = clampf(f, 0.0f, 100.0f) / 100.0f; return true;
    };
    if (key == "clr" || key == "color") {
        Vec3 v; if (!parse_vec3_strict(lx, v, err)) return false;
        p.has_clr = true; p.clr = color_from_100(v); return true;
    }
    if (key == "emt" || key == "emission") {
        Vec3 v; if (!parse_vec3_strict(lx, v, err)) return false;
        p.has_emt = true; p.emt = emission_from_100(v); return true;
    }
    if (key == "rfl" || key == "reflectivity" || key == "reflection") return scalar100(p.has_rfl, p.rfl);
    if (key == "rfr") return scalar100(p.has_rfr, p.rfr);
    if (key == "ior") { float f; if (!parse_scalar(lx, f, err)) return false; p.has_ior = true; p.ior = f; return true; }
    if (key == "spc" || key == "specular") return scalar100(p.has_spc, p.spc);
    if (key == "shn" || key == "shininess") { float f; if (!parse_scalar(lx, f, err)) return false; p.has_shn = true; p.shn = f; return true; }
    if (key == "fuz" || key == "fuzz") return scalar100(p.has_fuz, p.fuz);
    if (key == "typ" || key == "type") { std::string s; if (!parse_ident(lx, s, err)) return false; p.has_typ = true; p.typ = s; return true; }
    if (key == "refdis" || key == "dispersion") {
        float f; if (!parse_scalar(lx, f, err)) return false;
        p.has_refdis = true; p.refdis = (int)f; return true;
    }
    if (key == "shdw" || key == "shadow") { std::string s; return parse_ident(lx, s, err); }  // no effect on shading
    if (key == "radius" || key == "rad_l") { float f; if (!parse_scalar(lx, f, err)) return false; p.has_radius = true; p.radius = f; return true; }
    if (key == "metallic" || key == "metal" || key == "rough" || key == "roughness" || key == "spec_amt" ||
        key == "spec_tint" || key == "specular_tint" || key == "sheen" || key == "sheen_tint" || key == "clearcoat" ||
        key == "cc" || key == "cc_gloss" || key == "clearcoat_gloss" || key == "transmission" || key == "transmit") {
        float f; if (!parse_scalar(lx, f, err)) return false;
        p.has_principled = true; return true;
    }
    if (key == "sss" || key == "sss_r" || key == "sss_radius") { float f; return parse_scalar(lx, f, err); }  // needs render{sss}
    if (key == "sss_c" || key == "sss_color") { Vec3 v; return parse_vec3_strict(lx, v, err); }
    if (key == "tex" || key == "texture" || key == "nrml" || key == "normal" || key == "normal_map" || key == "bump" ||
        key == "disp" || key == "displacement" || key == "disp_map" || key == "disp_proc") {
        std::string s = lx.read_path();
        if (p.outside.empty()) p.outside = key;
        return true;
    }
    if (key == "disp_amt" || key == "disp_amount" || key == "disp_scale") { float f; return parse_scalar(lx, f, err); }
    if (key == "prst" || key == "preset") {
        std::string s; if (!parse_ident(lx, s, err)) return false;
        if (p.outside.empty()) p.outside = "preset " + s;
        return true;
    }
    return false;
}

// apply_patch, :711-781
inline Material apply_patch(const Material& base, const MaterialPatch& p) {
    Material m = base;
    if (p.has_typ) {
        if (p.typ == "metal") m.kind = MaterialKind::Metal;
        else if (p.typ == "glass" || p.typ == "dielectric") m.kind = MaterialKind::Dielectric;
        else if (p.typ == "principled" || p.typ == "pbr") m.kind = MaterialKind::Principled;
        else m.kind = MaterialKind::Diffuse;
    }
    if (p.has_principled) m.kind = MaterialKind::Principled;
    if (p.has_clr) m.albedo = p.clr;
    if (p.has_emt) m.emission = p.emt;
    if (p.has_rfl) {
        if (p.rfl > 0.0f && m.kind == MaterialKind::Diffuse) m.kind = MaterialKind::Metal;
        m.reflectivity = p.rfl;
    }
    if (p.has_rfr && p.rfr > 0.0f) {
        m.kind = MaterialKind::Dielectric;
        if (!p.has_ior) m.ior = 1.5f;
    }
    if (p.has_ior) { m.kind = MaterialKind::Dielectric; m.ior = p.ior; }
    if (p.has_spc) m.specular = p.spc;
    if (p.has_shn) m.shininess = p.shn;
    if (p.has_fuz) {
        if (m.kind != MaterialKind::Metal) m.kind = MaterialKind::Metal;
        m.reflectivity = 1.0f - p.fuz;
    }
    if (p.has_refdis) m.refdis = p.refdis;
    if (!p.outside.empty()) m.has_texture = true;
    return m;
}

inline void skip_braced(Lexer& lx) {
    int depth = 1;
    while (depth > 0) {
        Token t = lx.next();
        if (t.kind == TokKind::End) return;
        if (t.kind == TokKind::LBrace) depth++;
        else if (t.kind == TokKind::RBrace) depth--;
    }
}

inline bool parse_pecl(const std::string& path, ParsedScene& out, std::vector<ParseError>& errors);

inline bool parse_pecl_source(const std::string& src, ParsedScene& out, std::vector<ParseError>& errors) {
    Lexer lx(src);
    bool ok = true;
    while (true) {
        Token t = lx.next();
        if (t.kind == TokKind::End) break;
        if (t.kind != TokKind::Ident) {
            errors.push_back({t.line, "unexpected token '" + t.text + "'"});
            while (t.kind != TokKind::End && t.kind != TokKind::RBrace) t = lx.next();
            ok = false;
            continue;
        }
        const std::string block = t.text;
        if (block == "bg" || block == "background" || block == "amb" || block == "ambient") {
            Token colon = lx.next();
            if (colon.kind != TokKind::Colon) { errors.push_back({colon.line, "expected ':' after " + block}); ok = false; continue; }
            Vec3 v; ParseError e;
            if (parse_vec3_strict(lx, v, e)) {
                if (block[0] == 'b') out.scene.background = color_from_100(v);
                else out.scene.ambient = color_from_100(v);
            } else errors.push_back(e);
            continue;
        }
        Token lb = lx.next();
        if (lb.kind != TokKind::LBrace) { errors.push_back({lb.line, "expected '{' after " + block}); ok = false; continue; }

        MaterialPatch patch;
        std::string light_typ;
        while (true) {
            Token kt = lx.next();
            if (kt.kind == TokKind::RBrace) break;
            if (kt.kind == TokKind::End) { errors.push_back({kt.line, "unexpected end of file inside " + block}); ok = false; break; }
            if (kt.kind == TokKind::Comma) continue;
            if (kt.kind != TokKind::Ident) {
                errors.push_back({kt.line, "expected key, got '" + kt.text + "'"});
                ok = false;
                while (kt.kind != TokKind::End && kt.kind != TokKind::RBrace) kt = lx.next();
                if (kt.kind == TokKind::RBrace) break;
                continue;
            }
            const std::string key = kt.text;
            if (key == "func" || key == "function" || key == "anim" || key == "animation" || key == "phys" || key == "physics") {
                // sub-blocks: func bodies belong to implicit surfaces (outside
                // the core); anim/phys change nothing on a single frame 0.
                Token b = lx.next();
                if (b.kind == TokKind::LBrace) skip_braced(lx);
                if (key == "func" || key == "function") out.unsupported.push_back("implicit surface func");
                continue;
            }
            Token colon = lx.next();
            if (colon.kind != TokKind::Colon) {
                errors.push_back({colon.line, "expected ':' after " + key + " in " + block + ", got '" + colon.text + "'"});
                ok = false; continue;
            }
            ParseError e;
            bool consumed = false;
            bool non_material = (block == "fog" || block == "sky" || block == "skybox" || block == "env" ||
                                 block == "include" || block == "tern" || block == "terrain" || block == "gsky" ||
                                 block == "gen_sky" || block == "grp" || block == "group" || block == "render");
            if (!non_material && try_parse_material_key(key, lx, patch, e)) consumed = true;
            if (!consumed && block == "cmra") {
                if (key == "pos") consumed = parse_vec3_strict(lx, out.eye, e);
                else if (key == "aim" || key == "look" || key == "at") consumed = parse_vec3_strict(lx, out.look, e);
                else if (key == "up") consumed = parse_vec3_strict(lx, out.up, e);
                else if (key == "fov") consumed = parse_scalar(lx, out.fov, e);
                else if (key == "aspect") consumed = parse_scalar(lx, out.aspect, e);
                else if (key == "aperture" || key == "ap") consumed = parse_scalar(lx, out.aperture, e);
                else if (key == "focus" || key == "focal" || key == "fdist") consumed = parse_scalar(lx, out.focus_dist, e);
                else if (key == "bd" || key == "bokeh" || key == "blades") {
                    float f; if (parse_scalar(lx, f, e)) { out.bokeh_blades = (int)f; consumed = true; }
                }
            }
            if (!consumed && block == "mat") {
                errors.push_back({kt.line, "unknown material key: " + key});
                ok = false; lx.next();
                continue;
            }
            if (!consumed && block == "sphere") {
                if (key == "pos") consumed = parse_vec3_strict(lx, out._tmp_v1, e);
                else if (key == "rad") consumed = parse_scalar(lx, out._tmp_f1, e);
            }
            if (!consumed && block == "cube") {
                if (key == "pos") consumed = parse_vec3_strict(lx, out._tmp_v1, e);
                else if (key == "size") consumed = parse_scalar(lx, out._tmp_f1, e);
            }
            if (!consumed && block == "plane") {
                if (key == "pos") consumed = parse_vec3_strict(lx, out._tmp_v1, e);
                else if (key == "nrm" || key == "normal" || key == "dir") consumed = parse_vec3_strict(lx, out._tmp_v2, e);
            }
            if (!consumed && (block == "tri" || block == "triangle")) {
                if (key == "v0") consumed = parse_vec3_strict(lx, out._tmp_v1, e);
                else if (key == "v1") consumed = parse_vec3_strict(lx, out._tmp_v2, e);
                else if (key == "v2") consumed = parse_vec3_strict(lx, out._tmp_v3, e);
            }
            if (!consumed && block == "quad") {
                if (key == "a") consumed = parse_vec3_strict(lx, out._tmp_v1, e);
                else if (key == "b") consumed = parse_vec3_strict(lx, out._tmp_v2, e);
                else if (key == "c") consumed = parse_vec3_strict(lx, out._tmp_v3, e);
                else if (key == "d") consumed = parse_vec3_strict(lx, out._tmp_v4, e);
            }
            if (!consumed && (block == "cyl" || block == "cylinder")) {
                if (key == "pos" || key == "center") consumed = parse_vec3_strict(lx, out._tmp_v1, e);
                else if (key == "ax" || key == "axis") consumed = parse_vec3_strict(lx, out._tmp_v2, e);
                else if (key == "rad" || key == "radius") consumed = parse_scalar(lx, out._tmp_f1, e);
                else if (key == "h" || key == "height") consumed = parse_scalar(lx, out._tmp_f2, e);
            }
            if (!consumed && block == "cone") {
                if (key == "apex") consumed = parse_vec3_strict(lx, out._tmp_v1, e);
                else if (key == "ax" || key == "axis") consumed = parse_vec3_strict(lx, out._tmp_v2, e);
                else if (key == "rad" || key == "radius") consumed = parse_scalar(lx, out._tmp_f1, e);
                else if (key == "h" || key == "height") consumed = parse_scalar(lx, out._tmp_f2, e);
            }
            if (!consumed && (block == "tern" || block == "terrain")) {   // :1180-1207
                if (key == "size") {
                    int w, h;
                    if (parse_size2(lx, w, h, e)) { out._tmp_v1 = {(float)w, 0, (float)h}; consumed = true; }
                } else if (key == "height" || key == "h") consumed = parse_scalar(lx, out._tmp_f1, e);
                else if (key == "scale" || key == "freq") consumed = parse_scalar(lx, out._tmp_f2, e);
                else if (key == "octaves" || key == "oct") {
                    float f; if (parse_scalar(lx, f, e)) { out._tmp_i1 = (int)f; consumed = true; }
                } else if (key == "trgrid" || key == "grid") {
                    float f; if (parse_scalar(lx, f, e)) { out._tmp_i2 = (int)f; consumed = true; }
                } else if (key == "seed") {   // the token's text, a name or a number, is hashed
                    Token v = lx.next();
                    if (v.kind == TokKind::Ident || v.kind == TokKind::Number) out._pending_seed = v.text;
                    consumed = true;
                }
            }
            if (!consumed && block == "mesh") {
                if (key == "file" || key == "path" || key == "obj") { out._pending_mesh_file = lx.read_path(); consumed = true; }
                else if (key == "scale" || key == "fit") consumed = parse_scalar(lx, out._tmp_f1, e);
                else if (key == "pos" || key == "translate") consumed = parse_vec3_strict(lx, out._tmp_v3, e);
            }
            if (!consumed && block == "include") {
                if (key == "file" || key == "path") { out._pending_include = lx.read_path(); consumed = true; }
            }
            if (!consumed && (block == "grp" || block == "group")) {
                if (key == "pos") consumed = parse_vec3_strict(lx, out._tmp_v1, e);
                else if (key == "scale" || key == "s") consumed = parse_scalar(lx, out._tmp_f1, e);
            }
            if (!consumed && block == "light") {
                if (key == "typ" || key == "type") {
                    std::string s;
                    if (parse_ident(lx, s, e)) {
                        if (s == "point" || s == "dir" || s == "directional" || s == "sun") { light_typ = s; consumed = true; }
                        else e = {kt.line, "unknown light type: " + s};
                    }
                } else if (key == "pos") consumed = parse_vec3_strict(lx, out._tmp_v1, e);
                else if (key == "dir") consumed = parse_vec3_strict(lx, out._tmp_v2, e);
                else if (key == "intensity") {
                    Vec3 v; if (parse_vec3_strict(lx, v, e)) { patch.has_clr = true; patch.clr = color_from_100(v); consumed = true; }
                }
            }
            if (!consumed && block == "render") {
                if (key == "size") {
                    int w, h;
                    if (parse_size2(lx, w, h, e)) { out.width = w; out.height = h; out.aspect = (float)w / (float)h; consumed = true; }
                } else if (key == "spp" || key == "samples") {
                    float f; if (parse_scalar(lx, f, e)) { out.samples = (int)f; consumed = true; }
                } else if (key == "depth") {
                    float f; if (parse_scalar(lx, f, e)) { out.depth = (int)f; consumed = true; }
                } else if (key == "no_bvh" || key == "nobvh") {
                    std::string s; if (parse_ident(lx, s, e)) { out.scene.use_bvh = !(s == "true" || s == "yes" || s == "1"); consumed = true; }
                } else if (key == "fps" || key == "frames" || key == "anim_frames") {
                    float f; consumed = parse_scalar(lx, f, e);
                } else if (key == "pt" || key == "pathtrace" || key == "cautype" || key == "caustics_type" ||
                           key == "spctrl" || key == "spectral" || key == "sss" || key == "bdpt" ||
                           key == "volume" || key == "volumetric" || key == "mis") {
                    Token v = lx.next();
                    bool on = (v.kind == TokKind::Ident) ? (v.text == "true" || v.text == "yes" || v.text == "bi" ||
                                                            v.text == "bidirectional" || v.text == "uni" ||
                                                            v.text == "unidirectional" || v.text == "ppm" ||
                                                            v.text == "rglr" || v.text == "regular")
                                                         : (std::atof(v.text.c_str()) != 0.0);
                    consumed = (v.kind == TokKind::Ident || v.kind == TokKind::Number);
                    if (key == "pt" || key == "pathtrace") out.enable_pathtrace = on;
                    else if (key == "mis") out.enable_mis = on;
                    else if (key == "bdpt" && (v.text == "uni" || v.text == "unidirectional")) out.enable_pathtrace = true;
                    else if (on) out.unsupported.push_back("render " + key);
                }
            }
            if (!consumed && (block == "fog" || block == "sky" || block == "skybox" || block == "env" ||
                              block == "tern" || block == "terrain" || block == "gsky" || block == "gen_sky" ||
                              block == "surf")) {
                // outside the core: consume the value, the block is refused below
                Token v = lx.next();
                if (v.kind == TokKind::AngleOpen) { while (v.kind != TokKind::AngleClose && v.kind != TokKind::End) v = lx.next(); }
                consumed = true;
            }
            if (!consumed) {
                if (e.what.empty()) e = {kt.line, "unknown key '" + key + "' in " + block + " block"};
                errors.push_back(e);
                ok = false;
                lx.next();
            }
        }

        const Vec3 goff = out.group_offset();
        auto add_material = [&]() {
            Material m = apply_patch(out._current_mat, patch);
            out.scene.add_material(m);
            return out.scene.material_count() - 1;
        };
        if (block == "cmra") {
            out.aspect = (out.aspect == 0.0f) ? 1.0f : out.aspect;
        } else if (block == "mat") {
            Material fresh = make_diffuse({0.7f, 0.7f, 0.7f}, 0.0f, 32.0f);
            out._current_mat = apply_patch(fresh, patch);
        } else if (block == "sphere") {
            int mid = add_material();
            float rad = out._tmp_f1;
            out.scene.add_primitive(make_sphere(out._tmp_v1 + goff, rad, mid));
            const Material& m = out.scene.material(mid);
            if (m.emission.length_squared() > 1e-6f)   // :1690-1693
                out.scene.add_light(make_point_light(out._tmp_v1, m.emission, rad * 0.5f, true));
        } else if (block == "cube") {
            int mid = add_material();
            out.scene.add_primitive(make_cube(out._tmp_v1 + goff, out._tmp_f1, mid));
        } else if (block == "plane") {
            int mid = add_material();
            out.scene.add_primitive(make_plane(out._tmp_v1, out._tmp_v2, mid));
        } else if (block == "tri" || block == "triangle") {
            int mid = add_material();
            out.scene.add_primitive(make_triangle(out._tmp_v1, out._tmp_v2, out._tmp_v3, mid));
        } else if (block == "quad") {   // make_quad, primitives.hpp:390-396
            int mid = add_material();
            out.scene.add_primitive(make_triangle(out._tmp_v1, out._tmp_v2, out._tmp_v3, mid));
            out.scene.add_primitive(make_triangle(out._tmp_v1, out._tmp_v3, out._tmp_v4, mid));
        } else if (block == "cyl" || block == "cylinder" || block == "cone") {
            int mid = add_material();
            Vec3 axis = (out._tmp_v2.length_squared() < 1e-9f) ? Vec3{0, 1, 0} : out._tmp_v2;
            out.scene.add_primitive(make_axial(block == "cone" ? PCL_PRIM_CONE : PCL_PRIM_CYLINDER,
                                               out._tmp_v1, axis, out._tmp_f1, out._tmp_f2, mid));
        } else if (block == "mesh") {   // :1758-1789
            int mid = add_material();
            if (out._pending_mesh_file.empty()) {
                errors.push_back({0, "mesh block requires a file: ... entry"}); ok = false;
            } else {
                // the path as written, relative to the working directory (the
                // reference does not try the scene's directory for meshes,
                // unlike includes)
                const std::string mpath = out._pending_mesh_file;
                MeshData md; std::string merr;
                if (!load_obj(mpath, md, merr)) { errors.push_back({0, merr}); ok = false; }
                else {
                    float target = (out._tmp_f1 > 0.0f) ? out._tmp_f1 : 2.0f;
                    MeshTransform tf = fit_mesh(md, target);
                    Vec3 off = out._tmp_v3;
                    for (const auto& tri : md.triangles) {
                        Point3 v0 = tri[0] * tf.scale + tf.offset + off;
                        Point3 v1 = tri[1] * tf.scale + tf.offset + off;
                        Point3 v2 = tri[2] * tf.scale + tf.offset + off;
                        out.scene.add_primitive(make_triangle(v0, v1, v2, mid));
                    }
                }
            }
            out._pending_mesh_file.clear();
        } else if (block == "light") {   // :1823-1837
            Light l;
            // "type" is taken by the material-key pass first (:1618, :487-709)
            const std::string& lt = patch.has_typ ? patch.typ : light_typ;
            if (lt == "dir" || lt == "directional" || lt == "sun") {
                l = make_directional_light(out._tmp_v2, patch.has_clr ? patch.clr : Color{1, 1, 1});
            } else {
                float r = patch.has_radius ? patch.radius : 0.5f;
                l = make_point_light(out._tmp_v1, patch.has_clr ? patch.clr : Color{1, 1, 1}, r);
            }
            out.scene.add_light(l);
        } else if (block == "grp" || block == "group") {
            out._group_offsets.push_back(out._tmp_v1);
            if (out._tmp_f1 != 1.0f) out.unsupported.push_back("group scale");
        } else if (block == "include" || block == "render") {
            // include handled below; render consumed into fields
        } else if (block == "tern" || block == "terrain") {
            // :1541-1576: a grid x grid field of quads over [-w/2, w/2] x [-d/2, d/2],
            // each split into two triangles, heights fbm(x f, z f) * h
            const int mid = add_material();
            const float tw = out._tmp_v1.x > 0 ? out._tmp_v1.x : 20.0f;
            const float td = out._tmp_v1.z > 0 ? out._tmp_v1.z : 20.0f;
            const float th = out._tmp_f1 > 0 ? out._tmp_f1 : 3.0f;
            const float freq = out._tmp_f2 > 0 ? out._tmp_f2 : 0.1f;
            const int octaves = out._tmp_i1 > 0 ? out._tmp_i1 : 4;
            const uint32_t seed = out._pending_seed.empty() ? (uint32_t)(12345 + mid * 7)
                                                            : GradientNoise2::seed_of(out._pending_seed);
            const GradientNoise2 noise(seed);
            const int grid = out._tmp_i2 > 0 ? out._tmp_i2 : 32;
            for (int zi = 0; zi < grid; ++zi) {
                for (int xi = 0; xi < grid; ++xi) {
                    const float x0 = (float)xi / grid * tw - tw / 2, z0 = (float)zi / grid * td - td / 2;
                    const float x1 = (float)(xi + 1) / grid * tw - tw / 2, z1 = (float)(zi + 1) / grid * td - td / 2;
                    const Point3 a{x0, noise.fbm(x0 * freq, z0 * freq, octaves) * th, z0};
                    const Point3 b{x1, noise.fbm(x1 * freq, z0 * freq, octaves) * th, z0};
                    const Point3 c{x0, noise.fbm(x0 * freq, z1 * freq, octaves) * th, z1};
                    const Point3 d{x1, noise.fbm(x1 * freq, z1 * freq, octaves) * th, z1};
                    out.scene.add_primitive(make_triangle(a, b, d, mid));
                    out.scene.add_primitive(make_triangle(a, d, c, mid));
                }
            }
            out._pending_seed.clear();
        } else if (block == "fog" || block == "sky" || block == "skybox" || block == "env" ||
                   block == "gsky" || block == "gen_sky" || block == "surf") {
            out.unsupported.push_back(block + " block");
        } else {
            errors.push_back({t.line, "unknown block '" + block + "' (skipped)"});
            ok = false;
        }
        if (!patch.outside.empty() && block != "mat") out.unsupported.push_back(patch.outside);

        out._tmp_v1 = {0, 0, 0}; out._tmp_v2 = {0, 0, 0}; out._tmp_v3 = {0, 0, 0}; out._tmp_v4 = {0, 0, 0};
        out._tmp_f1 = 1.0f; out._tmp_f2 = 1.0f; out._tmp_i1 = 0;

        if (!out._pending_include.empty()) {   // :1859-1883
            std::string inc = out._pending_include;
            out._pending_include.clear();
            std::string full = inc;
            std::ifstream test(inc);
            if (!test && !out._base_dir.empty()) full = out._base_dir + "/" + inc;
            std::string saved = out._base_dir;
            size_t slash = full.find_last_of('/');
            out._base_dir = (slash != std::string::npos) ? full.substr(0, slash) : "";
            bool saved_skip = out._skip_finalize;
            out._skip_finalize = true;
            std::vector<ParseError> ie;
            parse_pecl(full, out, ie);
            for (auto& x : ie) errors.push_back(x);
            out._base_dir = saved;
            out._skip_finalize = saved_skip;
        }
    }
    if (out.width > 0 && out.height > 0) out.aspect = (float)out.width / (float)out.height;
    if (!out._skip_finalize) out.finalize();
    return ok;
}

// parse_pecl, :784
inline bool parse_pecl(const std::string& path, ParsedScene& out, std::vector<ParseError>& errors) {
    std::ifstream f(path);
    if (!f) { errors.push_back({0, "Cannot open scene file: " + path}); return false; }
    std::string src((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t slash = path.find_last_of('/');
    if (slash != std::string::npos && out._base_dir.empty()) out._base_dir = path.substr(0, slash);
    return parse_pecl_source(src, out, errors);
}

}  // namespace pencil
