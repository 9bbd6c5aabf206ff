// noise.hpp - 2-D gradient noise and its fractal sum for the terrain block.
//
// The terrain heights must be the reference's to the bit (its triangles are
// the scene's geometry), so this follows PerlinNoise (src/noise.hpp:16-91)
// step for step: a 256-entry permutation shuffled from the top down by the
// LCG s = s * 1103515245 + 12345 (index (s / 65536) mod (i + 1)), duplicated
// to 512; eight gradients chosen by hash & 7; the quintic fade; and the
// float operations in the reference's order (the host build contracts no
// a*b+c).  String seeds hash with 32-bit FNV-1a (0 maps to 1).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>

namespace pencil {

class GradientNoise2 {
public:
    explicit GradientNoise2(uint32_t seed) {
        uint32_t s = seed ? seed : 1u;
        for (int i = 0; i < 256; ++i) p_[i] = (uint8_t)i;
        for (int i = 255; i > 0; --i) {
            s = s * 1103515245u + 12345u;
            const int j = (int)((s / 65536u) % (uint32_t)(i + 1));
            const uint8_t a = p_[i];
            p_[i] = p_[j];
            p_[j] = a;
        }
        for (int i = 0; i < 256; ++i) p_[256 + i] = p_[i];
    }

    // noise2D (:31-47), in [-1, 1]
    float at(float x, float y) const {
        const float fx = std::floor(x), fy = std::floor(y);
        const int xi = (int)fx & 255, yi = (int)fy & 255;
        const float xf = x - fx, yf = y - fy;
        const float u = fade(xf), v = fade(yf);
        const int h00 = p_[p_[xi] + yi], h01 = p_[p_[xi] + yi + 1];
        const int h10 = p_[p_[xi + 1] + yi], h11 = p_[p_[xi + 1] + yi + 1];
        const float bottom = mix(slope(h00, xf, yf), slope(h10, xf - 1, yf), u);
        const float top = mix(slope(h01, xf, yf - 1), slope(h11, xf - 1, yf - 1), u);
        return mix(bottom, top, v);
    }

    // fbm (:50-62): octaves of at() at doubling frequency and halving
    // amplitude, over the sum of the amplitudes
    float fbm(float x, float y, int octaves) const {
        float sum = 0.0f, amp = 1.0f, freq = 1.0f, norm = 0.0f;
        for (int o = 0; o < octaves; ++o) {
            sum += at(x * freq, y * freq) * amp;
            norm += amp;
            amp *= 0.5f;
            freq *= 2.0f;
        }
        return sum / norm;
    }

    // hash_seed (:65-72)
    static uint32_t seed_of(const std::string& text) {
        uint32_t h = 2166136261u;
        for (unsigned char c : text) {
            h ^= (uint32_t)c;
            h *= 16777619u;
        }
        return h ? h : 1u;
    }

private:
    uint8_t p_[512];

    static float fade(float t) { return t * t * t * (t * (t * 6 - 15) + 10); }
    static float mix(float a, float b, float t) { return a + t * (b - a); }
    static float slope(int hash, float x, float y) {
        const int h = hash & 7;
        const float u = h < 4 ? x : y, v = h < 4 ? y : x;
        return ((h & 1) ? -u : u) + ((h & 2) ? -2.0f * v : 2.0f * v);
    }
};

}  // namespace pencil
