// Device-side binary32 vector algebra in the reference's operation order
// (reference math/tVector.ipp, math/math.h). Every HIP source of this
// library is compiled with -ffp-contract=off, so a*b+c stays two roundings,
// as in the reference's SSE build; correctly rounded division and sqrt are
// hipcc's defaults.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace mpt {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3f4(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
// TVector / scalar multiplies by the reciprocal (tVector.ipp:103-111)
__device__ __forceinline__ V3 divs(V3 a, float s) {
    const float inv = 1.0f / s;
    return a * inv;
}
// dot and squared length accumulate from zero (tVector.ipp:321-328, 427-434)
__device__ __forceinline__ float dot(V3 a, V3 b) {
    float r = 0.0f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
__device__ __forceinline__ float absdot(V3 a, V3 b) { return fabsf(dot(a, b)); }
__device__ __forceinline__ float len2(V3 a) { return dot(a, a); }
__device__ __forceinline__ float len(V3 a) { return sqrtf(len2(a)); }
__device__ __forceinline__ V3 normalize(V3 a) {
    const float inv = 1.0f / len(a);
    return a * inv;
}
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ bool is_zero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ __forceinline__ float max_comp(V3 a) {
    float m = a.x;
    if (a.y > m) m = a.y;
    if (a.z > m) m = a.z;
    return m;
}
// reflect about n: 2 |v.n| n - v (tVector.ipp:462-465)
__device__ __forceinline__ V3 reflect(V3 v, V3 n) { return n * (2.0f * absdot(v, n)) - v; }
__device__ __forceinline__ V3 lerp(V3 a, V3 b, float t) {
    const float c = 1.0f - t;
    return V3{c * a.x + t * b.x, c * a.y + t * b.y, c * a.z + t * b.z};
}
// math::min / max / clamp (math.h:116-129)
__device__ __forceinline__ float mmin(float a, float b) { return (a < b) ? a : b; }
__device__ __forceinline__ float mmax(float a, float b) { return (a > b) ? a : b; }
__device__ __forceinline__ float mclamp(float v, float lo, float hi) { return mmin(mmax(v, lo), hi); }
__device__ __forceinline__ float msign(float v) { return static_cast<float>((0.0f < v) - (v < 0.0f)); }

// constants as static_cast<real>(double) (math/constant.h)
constexpr float K_PI = static_cast<float>(3.1415926535897932);
constexpr float K_TWO_PI = static_cast<float>(6.2831853071795864);
constexpr float K_INV_PI = static_cast<float>(0.3183098861837906);
constexpr float K_INV_TWO_PI = static_cast<float>(0.1591549430918953);
constexpr float K_RAY_EPS = static_cast<float>(0.001);

// Philox4x32-10 (Salmon et al., SC 2011): counter-based, so any path vertex
// can re-derive its uniforms from (pixel, sample, bounce, slot).
__device__ __forceinline__ uint4 philox(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k.x += 0x9E3779B9u;
            k.y += 0xBB67AE85u;
        }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}
__device__ __forceinline__ float unit24(uint32_t x) { return static_cast<float>(x >> 8) * (1.0f / 16777216.0f); }

}  // namespace mpt
