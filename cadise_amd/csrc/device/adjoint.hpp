// Device pieces of the path-replay adjoint: the differentiable texel of a
// vertex, the roughness log-derivative of the GGX BSDF, and the per-wave LDS
// tables that pre-reduce gradient contributions before the FP32 atomics.
// The reference has no derivatives; the functions differentiate the BSDFs of
// shade.hpp (conductorMicrofacet.cpp:32-67, isotropicGgx.cpp:11-69,
// tRoughnessMapper.h:18-40).
#pragma once

#include "shade.hpp"

namespace mpt {

// Texture element index and channel count of the texel a vertex reads:
// albedo (Lambertian) or roughness (GGX); (-1, 0) for constant parameters.
// The texel is found by the same arithmetic as tex_lookup.
__device__ __forceinline__ int2 vertex_param(const DevScene& sc, const DevMaterial& m, V3 uvw) {
    const int32_t id = m.type == 0 ? m.albedo_tex : m.rough_tex;
    if (id < 0) return make_int2(-1, 0);
    const DevTexture t = tex_resolve(sc, id, uvw);
    float u, v;
    if (t.wrap_clamp) {
        u = mclamp(uvw.x, 0.f, 1.f);
        v = mclamp(uvw.y, 0.f, 1.f);
    } else {
        u = uvw.x - floorf(uvw.x);
        v = uvw.y - floorf(uvw.y);
    }
    int32_t x = static_cast<int32_t>(u * static_cast<float>(t.width));
    int32_t y = static_cast<int32_t>(v * static_cast<float>(t.height));
    x = (x < t.width) ? x : x - 1;
    y = (y < t.height) ? y : y - 1;
    const int64_t e = t.offset + (static_cast<int64_t>(y) * t.width + x) * t.channels;
    return make_int2(static_cast<int32_t>(e), t.channels);
}

// d ln f / d r for the GGX conductor at (V = h.wi, L, N = h.Ns), r the
// roughness texel value: F does not depend on alpha; D and both G1 factors do;
// alpha = max(max(r, 1e-3)^2, 1e-3).
__device__ __forceinline__ float ggx_dlnf_dr(const DevScene& sc, const DevMaterial& m, const SurfaceHit& h, V3 L) {
    const float r = m.rough_tex < 0 ? m.roughness : tex_lookup(sc, m.rough_tex, h.uvw).x;
    const float safe = mmax(r, 0.001f);
    const float a = mmax(safe * safe, 0.001f);
    const float dadr = (r > 0.001f && safe * safe > 0.001f) ? 2.0f * r : 0.0f;
    const V3 N = h.Ns, V = h.wi;
    const float vn = dot(V, N), ln = dot(L, N);
    const V3 H = normalize((V + L) * msign(vn));
    const float c = dot(N, H), c2 = c * c;
    const float tan2 = (1.0f - c2) / c2;
    float d = 2.0f / a - 4.0f * a / (a * a + tan2);
    const float t2v = (1.0f - vn * vn) / (vn * vn), sv = sqrtf(1.0f + a * a * t2v);
    const float t2l = (1.0f - ln * ln) / (ln * ln), sl = sqrtf(1.0f + a * a * t2l);
    d -= a * t2v / (sv * (1.0f + sv));
    d -= a * t2l / (sl * (1.0f + sl));
    return d * dadr;
}

// Per-wave open-addressing table in LDS: (key, 3 floats). Every lane of a wave
// inserts at most one key per round, so NSLOT >= 64 always has room; linear
// probing with atomicCAS on the key, LDS float atomics on the values.
template <int NSLOT>
struct WaveTable {
    int*   keys;   // NSLOT, -1 = empty
    float* vals;   // 3 * NSLOT

    __device__ __forceinline__ void clear_lane(int lane) {
        for (int i = lane; i < NSLOT; i += 64) {
            keys[i] = -1;
            vals[3 * i] = vals[3 * i + 1] = vals[3 * i + 2] = 0.0f;
        }
    }
    __device__ __forceinline__ void insert(int key, V3 v) {
        uint32_t h = (static_cast<uint32_t>(key) * 0x9E3779B1u) >> 16;
        for (int probe = 0; probe < NSLOT; ++probe) {
            const uint32_t i = (h + probe) & (NSLOT - 1);
            const int prev = atomicCAS(&keys[i], -1, key);
            if (prev == -1 || prev == key) {
                atomicAdd(&vals[3 * i], v.x);
                atomicAdd(&vals[3 * i + 1], v.y);
                atomicAdd(&vals[3 * i + 2], v.z);
                return;
            }
        }
    }
};

}  // namespace mpt
