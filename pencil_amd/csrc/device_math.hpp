// device_math.hpp - float vector math with the reference's evaluation order.
//
// Every operation is a single IEEE float op in the order the reference's
// Vec3 (src/vec3.hpp:24-92) performs it; the whole library is compiled with
// -ffp-contract=off so no a*b+c is fused, division and sqrt are correctly
// rounded (HIP's default), and denormals are kept.  With that, the device
// reproduces the reference's x86-64 SSE float results bit for bit for
// + - * / sqrt.  cos/sin/pow are evaluated in double and rounded once: this
// is the correctly rounded float result except in astronomically rare cases,
// where the reference's libm (glibc 2.35 cosf/sinf/powf, < 1 ulp but not
// always correctly rounded) may differ by one ulp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PCL_HD __host__ __device__ __forceinline__
#define PCL_D __device__ __forceinline__

// Explicit address spaces.  Pointers that reach a kernel through memory (a
// scene table, a struct) lose the compiler's knowledge that they point to
// global memory or LDS, and it then emits generic FLAT instructions, which
// are slower and stall on both memory counters.  Loads from scene data and
// the LDS traversal stack go through these qualifiers instead.
#define PCL_GLOBAL __attribute__((address_space(1)))
#define PCL_LDS __attribute__((address_space(3)))

template <class T>
__device__ __forceinline__ const PCL_GLOBAL T* gptr(const T* p) { return (const PCL_GLOBAL T*)p; }

typedef float pcl_vf4 __attribute__((ext_vector_type(4)));
typedef int pcl_vi4 __attribute__((ext_vector_type(4)));
// 16-byte global loads (one dwordx4) of a 16-byte aligned address.
__device__ __forceinline__ float4 gload_f4(const void* p) {
    const pcl_vf4 v = *(const PCL_GLOBAL pcl_vf4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int4 gload_i4(const void* p) {
    const pcl_vi4 v = *(const PCL_GLOBAL pcl_vi4*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}

namespace pcl {

struct f3 {
    float x, y, z;
    PCL_HD f3() : x(0.f), y(0.f), z(0.f) {}
    PCL_HD f3(float a, float b, float c) : x(a), y(b), z(c) {}
    PCL_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    PCL_HD f3 operator+(const f3& o) const { return f3(x + o.x, y + o.y, z + o.z); }
    PCL_HD f3 operator-(const f3& o) const { return f3(x - o.x, y - o.y, z - o.z); }
    PCL_HD f3 operator-() const { return f3(-x, -y, -z); }
    PCL_HD f3 operator*(float s) const { return f3(x * s, y * s, z * s); }
    PCL_HD f3 operator/(float s) const { return f3(x / s, y / s, z / s); }
    PCL_HD f3 operator*(const f3& o) const { return f3(x * o.x, y * o.y, z * o.z); }
    PCL_HD f3& operator+=(const f3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    PCL_HD float dot(const f3& o) const { return x * o.x + y * o.y + z * o.z; }
    PCL_HD f3 cross(const f3& o) const { return f3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
    PCL_HD float length_squared() const { return x * x + y * y + z * z; }
    PCL_HD float length() const { return sqrtf(length_squared()); }
    PCL_HD f3 normalized() const {
        float l = length();
        return (l > 1e-12f) ? (*this / l) : f3(0.f, 0.f, 0.f);
    }
    PCL_HD f3 reflect(const f3& n) const { return *this - n * (2.0f * this->dot(n)); }
};

PCL_HD float clampf(float v, float lo, float hi) { return (v < lo) ? lo : (v > hi) ? hi : v; }

// libm-equivalent transcendentals: one double evaluation, one rounding.
// Kept out of line: each double-precision expansion is a few KB of code, and
// inlined at every call site it would push the kernels far past the
// instruction cache.
#define PCL_DN __device__ __attribute__((noinline))
PCL_DN float cosf_cr(float x) { return (float)cos((double)x); }
PCL_DN float sinf_cr(float x) { return (float)sin((double)x); }
PCL_DN float powf_cr(float x, float y) { return (float)pow((double)x, (double)y); }
PCL_DN float atan2f_cr(float y, float x) { return (float)atan2((double)y, (double)x); }

// schlick (src/vec3.hpp:83-87)
PCL_D float schlick(float c, float r) {
    float r0 = (1.0f - r) / (1.0f + r);
    r0 = r0 * r0;
    return r0 + (1.0f - r0) * powf_cr(1.0f - c, 5.0f);
}

// xorshift32 stream, one per pixel (src/rng.hpp:13-45).
struct Rng {
    uint32_t s;
    PCL_D explicit Rng(uint32_t v) : s(v) {}
    PCL_D uint32_t u32() { s ^= s << 13u; s ^= s >> 17u; s ^= s << 5u; return s; }
    PCL_D float unit() { return (float)u32() / 4294967296.0f; }
    PCL_D float sym() { return unit() * 2.0f - 1.0f; }
    PCL_D void skip(uint32_t n) { for (uint32_t k = 0; k < n; ++k) u32(); }   // n draws, values unused
    // Shirley-Chiu concentric mapping (rng.hpp:28-44)
    PCL_D void unit_disc(float& dx, float& dy) {
        float r1 = unit(), r2 = unit();
        float sx = 2.0f * r1 - 1.0f, sy = 2.0f * r2 - 1.0f;
        if (sx == 0.0f && sy == 0.0f) { dx = 0.f; dy = 0.f; return; }
        float theta, r;
        if (fabsf(sx) > fabsf(sy)) { r = sx; theta = (3.14159265f / 4.0f) * (sy / sx); }
        else { r = sy; theta = (3.14159265f / 2.0f) - (3.14159265f / 4.0f) * (sx / sy); }
        dx = r * cosf_cr(theta);
        dy = r * sinf_cr(theta);
    }
};

struct Ray {
    f3 o, d;
    PCL_HD f3 at(float t) const { return o + d * t; }
};

}  // namespace pcl
