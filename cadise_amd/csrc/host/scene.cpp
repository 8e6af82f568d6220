// Host-side setup math whose results feed the device kernels bit-for-bit:
// camera matrices, area-light radiance, power-cluster CDF. The operation
// order is the reference's (binary32, no contraction: built with
// -ffp-contract=off) because the device reproduces the reference's rays.
#include "scene.hpp"

#include <cfloat>

#include "bvh.hpp"

#include <cmath>
#include <cstring>

namespace mpt {

namespace {

inline F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 scale(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len2(F3 a) {
    float r = 0.f;
    r += a.x * a.x;
    r += a.y * a.y;
    r += a.z * a.z;
    return r;
}
inline float len(F3 a) { return std::sqrt(len2(a)); }
inline F3 normalize(F3 a) {
    const float inv = 1.f / len(a);
    return scale(a, inv);
}
inline F3 load(const float* p) { return {p[0], p[1], p[2]}; }

// TMatrix4::operator*= (tMatrix4.ipp:102-116): accumulate from zero, k ascending
void mat_mul_inplace(float* a, const float* b) {
    float t[16];
    for (int r = 0; r < 4; ++r) {
        for (int c = 0; c < 4; ++c) {
            float acc = 0.f;
            for (int k = 0; k < 4; ++k) acc += a[4 * r + k] * b[4 * k + c];
            t[4 * r + c] = acc;
        }
    }
    std::memcpy(a, t, sizeof(t));
}

}  // namespace

void camera_matrices(const HostScene& s, float f2c[16], float c2w[16]) {
    // lookAt: z = -direction, x = up x z, y = z x x; columns x, y, z, position
    const F3 nz = normalize(F3{-s.cam_dir.x, -s.cam_dir.y, -s.cam_dir.z});
    const F3 nx = normalize(cross(s.cam_up, nz));
    const F3 ny = cross(nz, nx);
    const float m[16] = {nx.x, ny.x, nz.x, s.cam_pos.x, nx.y, ny.y, nz.y, s.cam_pos.y,
                         nx.z, ny.z, nz.z, s.cam_pos.z, 0.f,  0.f,  0.f,  1.f};
    std::memcpy(c2w, m, sizeof(m));

    // Sensor size in float64 (Camera::_getAspectRatio, _getSensorSizeXy), the
    // half field of view converted to float before degree->radian, and the
    // film->camera matrix T(-W/2, H/2, -offset) * S(W/resX, -H/resY, 1) in float.
    const double aspect = static_cast<double>(static_cast<size_t>(s.width)) /
                          static_cast<double>(static_cast<size_t>(s.height));
    const double sw = static_cast<double>(s.sensor_width_mm);
    const double sh = sw / aspect;
    const float resx = static_cast<float>(static_cast<size_t>(s.width));
    const float resy = static_cast<float>(static_cast<size_t>(s.height));
    const float half_rad = static_cast<float>(static_cast<double>(s.fov) * 0.5) * static_cast<float>(0.0174532925199432);
    const double offset = (sw * 0.5) / static_cast<double>(std::tan(half_rad));
    float f[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const float tr[16] = {1, 0, 0, static_cast<float>(-(sw * 0.5)), 0, 1, 0, static_cast<float>(sh * 0.5),
                          0, 0, 1, static_cast<float>(-offset),     0, 0, 0, 1};
    const float sc[16] = {static_cast<float>(sw / static_cast<double>(resx)), 0, 0, 0,
                          0, static_cast<float>(-sh / static_cast<double>(resy)), 0, 0,
                          0, 0, 1, 0,
                          0, 0, 0, 1};
    mat_mul_inplace(f, tr);
    mat_mul_inplace(f, sc);
    std::memcpy(f2c, f, sizeof(f));
}

float prim_area(const HostScene& s, int64_t i) {
    const float* v = &s.prim_v[9 * i];
    if (s.prim_type[i] == PRIM_SPHERE) return static_cast<float>(12.566370614359172) * v[3] * v[3];  // sphere.cpp:142-144
    if (s.prim_type[i] == PRIM_RECTANGLE) {
        const F3 ea = sub(load(v), load(v + 3)), eb = sub(load(v + 6), load(v + 3));
        return len(ea) * len(eb);
    }
    const F3 e1 = sub(load(v + 3), load(v)), e2 = sub(load(v + 6), load(v));
    return 0.5f * len(cross(e1, e2));
}

std::vector<F3> light_radiance(const HostScene& s) {
    std::vector<F3> out(s.lights.size());
    for (size_t i = 0; i < s.lights.size(); ++i) {
        const Light& l = s.lights[i];
        if (l.type == LIGHT_ENVIRONMENT) continue;   // radiance from its texture
        float sum = 0.f;
        sum += l.color.x;
        sum += l.color.y;
        sum += l.color.z;
        const F3 unit = scale(l.color, 1.f / sum);         // color / color.sum()
        const F3 total = scale(unit, l.watt);
        const float area = prim_area(s, l.prim);
        const F3 per_area = scale(total, 1.f / area);       // / area
        out[i] = scale(per_area, static_cast<float>(0.3183098861837906));  // * inv_pi
    }
    return out;
}

void distribution_cdf(const float* value, size_t n, float* cdf) {
    const float delta = 1.f / static_cast<float>(n);
    cdf[0] = 0.f;
    for (size_t i = 1; i <= n; ++i) cdf[i] = cdf[i - 1] + value[i - 1] * delta;
    const float sum = cdf[n];
    if (sum > 0.f) {
        const float inv = 1.f / sum;
        for (size_t i = 1; i <= n; ++i) cdf[i] *= inv;
    } else {
        for (size_t i = 1; i <= n; ++i) cdf[i] = static_cast<float>(i) * delta;
    }
    cdf[n] = 1.f;
}

EnvTables env_tables(const Texture& t) {
    EnvTables out;
    const size_t w = static_cast<size_t>(t.width), h = static_cast<size_t>(t.height);
    std::vector<float> weighted(w * h);
    const float pi = static_cast<float>(3.1415926535897932);
    for (size_t iy = 0; iy < h; ++iy) {
        const float v = (static_cast<float>(iy) + 0.5f) / static_cast<float>(h);
        // image rows run bottom to top, spherical theta top to bottom
        const float sin_theta = std::sin((1.0f - v) * pi);
        for (size_t ix = 0; ix < w; ++ix) {
            // nearest lookup at the texel centre is texel (ix, iy)
            const float* p = &t.data[3 * (iy * w + ix)];
            const float lum = 0.0126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
            weighted[iy * w + ix] = lum * sin_theta;
            out.background_flux += weighted[iy * w + ix];
        }
    }
    out.cdf.assign((h + 1) + h * (w + 1), 0.f);
    std::vector<float> rows(h, 0.f);
    for (size_t iy = 0; iy < h; ++iy) {
        for (size_t ix = 0; ix < w; ++ix) rows[iy] += weighted[iy * w + ix];
        distribution_cdf(&weighted[iy * w], w, &out.cdf[(h + 1) + iy * (w + 1)]);
    }
    distribution_cdf(rows.data(), h, out.cdf.data());
    return out;
}

float env_scene_radius(const HostScene& s) {
    float lo[3], hi[3];
    reference_root_bound(s, lo, hi);
    const bool empty = (lo[0] == FLT_MAX && lo[1] == FLT_MAX && lo[2] == FLT_MAX) ||
                       (hi[0] == FLT_MIN && hi[1] == FLT_MIN && hi[2] == FLT_MIN);
    if (empty) return 5000.f;
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    float l2 = 0.f;
    l2 += ex * ex;
    l2 += ey * ey;
    l2 += ez * ez;
    return std::sqrt(l2) * 0.5f;
}

std::vector<float> power_cdf(const HostScene& s, const std::vector<F3>& rad) {
    const size_t n = s.lights.size();
    std::vector<float> cdf(n + 1, 0.f);
    if (n == 0) return cdf;
    const float delta = 1.f / static_cast<float>(n);
    for (size_t i = 1; i <= n; ++i) {
        const Light& l = s.lights[i - 1];
        float flux;
        if (l.type == LIGHT_ENVIRONMENT) {
            // EnvironmentLight::_updateApproxmiatedFlux (environmentLight.cpp:138-140).
            // The scene-bound radius reaches the environment light only when it is
            // the last light declared: RenderDatabase::_setUpLight re-marks every
            // light declared after it as "the environment light"
            // (renderDatabase.cpp:181-189, 100-111), which keeps the default 5000.
            const float r = (i == n) ? env_scene_radius(s) : 5000.f;
            flux = env_tables(s.textures[l.env_tex]).background_flux * static_cast<float>(12.566370614359172) * r * r;
        } else {
            const F3 le = rad[i - 1];
            const float lum = 0.0126f * le.x + 0.7152f * le.y + 0.0722f * le.z;
            flux = lum * prim_area(s, l.prim) * static_cast<float>(3.1415926535897932);
            if (flux <= 0.f) flux = 1.f;
        }
        cdf[i] = cdf[i - 1] + flux * delta;
    }
    const float total = cdf[n];
    if (total > 0.f) {
        const float inv = 1.f / total;
        for (size_t i = 1; i <= n; ++i) cdf[i] *= inv;
    } else {
        for (size_t i = 1; i <= n; ++i) cdf[i] = static_cast<float>(i) * delta;
    }
    cdf[n] = 1.f;
    return cdf;
}

int32_t strata_base(int32_t spp) {
    const size_t lower = static_cast<size_t>(std::floor(std::sqrt(static_cast<float>(static_cast<size_t>(spp)))));
    return static_cast<int32_t>(static_cast<size_t>(std::sqrt(static_cast<double>(lower * lower))));
}

}  // namespace mpt
