// image_io.hpp - HDR -> 8-bit encode and PPM/PNG output.
//
// color_to_rgb8 follows src/image_io.hpp:46-59 and src/tonemap.hpp:27-61:
// exposure (2^stops), per-channel ACES (Narkowicz 2015) or Reinhard, clamp,
// gamma 1/2.2, (uint8)(x * 255 + 0.5).  PNG output is written with stored
// (uncompressed) deflate blocks: the pixels are what matter, the reference's
// stb encoder only compresses better.
#pragma once

#include "scene.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace pencil {

enum class ToneMap { None, ACES, Reinhard, Exponential };

inline float aces(float x) {
    float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return std::max(0.0f, (x * (a * x + b)) / (x * (c * x + d) + e));
}

inline Color tonemap_color(Color c, ToneMap mode, float exposure) {
    float e = std::pow(2.0f, exposure);
    c = c * e;
    switch (mode) {
    case ToneMap::ACES: return {aces(c.x), aces(c.y), aces(c.z)};
    case ToneMap::Reinhard: return {c.x / (1.0f + c.x), c.y / (1.0f + c.y), c.z / (1.0f + c.z)};
    case ToneMap::Exponential: return {1.0f - std::exp(-c.x), 1.0f - std::exp(-c.y), 1.0f - std::exp(-c.z)};
    default: return c;
    }
}

inline void color_to_rgb8(const Color& c, ToneMap tm, float exposure, uint8_t* rgb) {
    Color m = tonemap_color(c, tm, exposure);
    float v[3] = {clampf(m.x, 0.0f, 1.0f), clampf(m.y, 0.0f, 1.0f), clampf(m.z, 0.0f, 1.0f)};
    for (int k = 0; k < 3; ++k) rgb[k] = (uint8_t)(std::pow(v[k], 1.0f / 2.2f) * 255.0f + 0.5f);
}

namespace detail {
inline uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    for (size_t i = 0; i < n; ++i) {
        c ^= p[i];
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
    }
    return c;
}
inline void be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24)); v.push_back((uint8_t)(x >> 16)); v.push_back((uint8_t)(x >> 8)); v.push_back((uint8_t)x);
}
inline void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    be32(out, (uint32_t)data.size());
    std::vector<uint8_t> td(type, type + 4);
    td.insert(td.end(), data.begin(), data.end());
    out.insert(out.end(), td.begin(), td.end());
    be32(out, crc32(td.data(), td.size()) ^ 0xFFFFFFFFu);
}
}  // namespace detail

inline bool write_png_rgb8(const std::string& path, const uint8_t* rgb, int W, int H) {
    std::vector<uint8_t> raw;
    raw.reserve((size_t)H * (W * 3 + 1));
    for (int j = 0; j < H; ++j) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb + (size_t)j * W * 3, rgb + (size_t)(j + 1) * W * 3);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t x : raw) { a = (a + x) % 65521u; b = (b + a) % 65521u; }
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        size_t n = std::min<size_t>(65535, raw.size() - off);
        bool last = off + n >= raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)n); z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)~n); z.push_back((uint8_t)(~n >> 8));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
        if (last) break;
    }
    detail::be32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    detail::be32(ihdr, (uint32_t)W); detail::be32(ihdr, (uint32_t)H);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    detail::chunk(out, "IHDR", ihdr);
    detail::chunk(out, "IDAT", z);
    detail::chunk(out, "IEND", {});
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return std::fclose(f) == 0 && ok;
}

// write_image, src/image_io.hpp:63-113: format from the extension (.ppm/.png).
inline bool write_image(const std::string& path, const std::vector<Color>& pixels, int W, int H, std::string& err,
                        ToneMap tm = ToneMap::ACES, float exposure = 0.0f) {
    std::string ext = path.substr(path.find_last_of('.') == std::string::npos ? path.size() : path.find_last_of('.') + 1);
    for (auto& c : ext) c = (char)std::tolower((unsigned char)c);
    if (ext != "ppm" && ext != "png") { err = "unknown image format: " + path + " (use .ppm or .png)"; return false; }
    std::vector<uint8_t> rgb((size_t)W * H * 3);
    for (size_t i = 0; i < (size_t)W * H; ++i) color_to_rgb8(pixels[i], tm, exposure, &rgb[i * 3]);
    if (ext == "png") {
        if (!write_png_rgb8(path, rgb.data(), W, H)) { err = "cannot write " + path; return false; }
        return true;
    }
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { err = "cannot open output: " + path; return false; }
    std::fprintf(f, "P6\n%d %d\n255\n", W, H);
    bool ok = std::fwrite(rgb.data(), 1, rgb.size(), f) == rgb.size();
    if (std::fclose(f) != 0 || !ok) { err = "cannot write " + path; return false; }
    return true;
}

}  // namespace pencil
