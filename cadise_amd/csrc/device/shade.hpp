// Shading-stage device functions: surface detail, textures, the two BSDFs of
// this path, area-light sampling and MIS. Each follows the reference function
// cited beside it in operation order (binary32, no contraction), so a path
// shaded here and by the reference from the same uniforms differs only by the
// last-ulp differences of acosf / cosf / sinf / atanf / expf / powf.
#pragma once

#include "dmath.hpp"
#include "dscene.hpp"

namespace mpt {

struct SurfaceHit {
    int32_t prim;
    V3 P, Ng, Ns, uvw, wi;
};

// Triangle::_positionToBarycentric (triangle.cpp:233-262); v0 = eAB and v1 = eAC
// are the stored edges (vB - vA and vC - vA computed once, as the reference does).
__device__ __forceinline__ V3 barycentric(V3 A, V3 e1, V3 e2, V3 P) {
    const V3 v2 = P - A;
    const float d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
    const float d20 = dot(v2, e1), d21 = dot(v2, e2);
    const float den = d00 * d11 - d01 * d01;
    if (den == 0.0f) return v3(0.f, 0.f, 0.f);
    const float inv = 1.0f / den;
    const float u = (d11 * d20 - d01 * d21) * inv;
    const float v = (d00 * d21 - d01 * d20) * inv;
    return v3(1.0f - u - v, u, v);
}

__device__ __forceinline__ V3 bary_mix(V3 b, V3 a, V3 c, V3 d) { return (a * b.x + c * b.y) + d * b.z; }

__device__ __forceinline__ V3 load3(const float* p) { return v3(p[0], p[1], p[2]); }

// SphericalMapper::mappingToUvw (sphericalMapper.cpp:14-28): y up, phi from +z
// towards +x.
__device__ __forceinline__ V3 spherical_uvw(V3 d) {
    const V3 u = normalize(d);
    const float theta = acosf(mclamp(u.y, -1.0f, 1.0f));
    float phi = atan2f(u.x, u.z);
    if (phi < 0.0f) phi += K_TWO_PI;
    return v3(phi * K_INV_TWO_PI, (K_PI - theta) * K_INV_PI, 0.0f);
}

// Scene::isIntersecting's surface detail (scene.cpp:30-48) with
// Triangle::evaluateSurfaceDetail (triangle.cpp:125-160),
// Rectangle::evaluateSurfaceDetail (rectangle.cpp:102-132) and
// Sphere::evaluateSurfaceDetail (sphere.cpp:109-130).
__device__ __forceinline__ SurfaceHit surface_detail(const DevScene& sc, int32_t leaf, V3 o, V3 d, float t) {
    SurfaceHit h;
    const float4 r0 = sc.tris[3 * leaf + 0];
    const float4 r1 = sc.tris[3 * leaf + 1];
    const float4 r2 = sc.tris[3 * leaf + 2];
    const uint32_t bits = __float_as_uint(r0.w);
    h.prim = static_cast<int32_t>(bits & 0x7FFFFFFFu);
    h.wi = -d;
    h.P = o + d * t;
    const V3 A = v3f4(r0), e1 = v3f4(r1), e2 = v3f4(r2);
    if ((bits & 0x80000000u) && r1.w != 0.0f) {
        // Sphere::evaluateSurfaceDetail (sphere.cpp:109-130): the outward normal, uvw
        // by the spherical mapper of it
        const V3 N = normalize(h.P - A);
        h.Ng = N;
        h.Ns = N;
        h.uvw = spherical_uvw(N);
        return h;
    }
    V3 N = cross(e1, e2);
    N = is_zero(N) ? v3(0.f, 1.f, 0.f) : normalize(N);
    if (bits & 0x80000000u) {
        h.Ng = N;
        h.Ns = N;
        const V3 onp = h.P - A;
        const float p1 = dot(onp, normalize(e1)) / len(e1);
        const float p2 = dot(onp, normalize(e2)) / len(e1);   // |eA| for both: the reference's quirk
        // uvw of corners A(1,0,0) B(0,0,0) C(0,1,0) D(1,1,0)
        const V3 l1 = lerp(v3(0.f, 0.f, 0.f), v3(0.f, 1.f, 0.f), p2);
        const V3 l2 = lerp(v3(1.f, 0.f, 0.f), v3(1.f, 1.f, 0.f), p2);
        h.uvw = lerp(l1, l2, p1);
        return h;
    }
    const int4 info = sc.prim_info[h.prim];
    const V3 b = barycentric(A, e1, e2, h.P);
    if (info.w) {
        const float* u = sc.prim_uvw + 9 * static_cast<int64_t>(h.prim);
        h.uvw = bary_mix(b, load3(u), load3(u + 3), load3(u + 6));
    } else {
        h.uvw = bary_mix(b, v3(0.f, 0.f, 0.f), v3(1.f, 0.f, 0.f), v3(0.f, 1.f, 0.f));
    }
    V3 Ns;
    if (info.z) {
        const float* n = sc.prim_n + 9 * static_cast<int64_t>(h.prim);
        V3 na = load3(n), nb = load3(n + 3), nc = load3(n + 6);
        if (is_zero(na)) na = N;
        if (is_zero(nb)) nb = N;
        if (is_zero(nc)) nc = N;
        Ns = bary_mix(b, na, nb, nc);
    } else {
        Ns = bary_mix(b, N, N, N);
    }
    h.Ns = is_zero(Ns) ? N : normalize(Ns);
    h.Ng = N;
    return h;
}

// TConstantTexture / TNearestPixelSampler with repeat or clamp wrap
// (tConstantTexture.ipp:14-19, tNearestPixelSampler.ipp:13-31,
// repeatUvwWrapper.cpp:9-15, clampUvwWrapper.cpp:8-12)
// The image a texture reads at uvw: checkerboards resolved
// (TCheckerboardTexture::evaluate, tCheckerboardTexture.ipp:28-40; the host
// bounds the nesting).
__device__ __forceinline__ DevTexture tex_resolve(const DevScene& sc, int32_t id, V3 uvw) {
    DevTexture t = sc.texs[id];
    for (int k = 0; k < 9 && t.odd >= 0; ++k) {
        const int32_t uc = static_cast<int32_t>(floorf(uvw.x / t.odd_size));
        const int32_t vc = static_cast<int32_t>(floorf(uvw.y / t.even_size));
        t = sc.texs[(uc + vc) % 2 == 0 ? t.even : t.odd];
    }
    return t;
}

__device__ __forceinline__ V3 tex_lookup(const DevScene& sc, int32_t id, V3 uvw) {
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
    const float* px = sc.tex_data + t.offset + (static_cast<int64_t>(y) * t.width + x) * t.channels;
    return t.channels == 1 ? v3(px[0], px[0], px[0]) : v3(px[0], px[1], px[2]);
}

__device__ __forceinline__ V3 mat_albedo(const DevScene& sc, const DevMaterial& m, V3 uvw) {
    return m.albedo_tex < 0 ? v3(m.albedo[0], m.albedo[1], m.albedo[2]) : tex_lookup(sc, m.albedo_tex, uvw);
}
// TRoughnessMapper<SQUARE>::map (tRoughnessMapper.h:18-40)
__device__ __forceinline__ float mat_alpha(const DevScene& sc, const DevMaterial& m, V3 uvw) {
    const float r = m.rough_tex < 0 ? m.roughness : tex_lookup(sc, m.rough_tex, uvw).x;
    const float safe = mmax(r, 0.001f);
    return mmax(safe * safe, 0.001f);
}

// IsotropicGgx::distributionD / shadowingMaskingG (isotropicGgx.cpp:11-69),
// Microfacet::_isShadowingMaskingValid (microfacet.cpp:11-23)
__device__ __forceinline__ float ggx_D(float a, V3 N, V3 H) {
    const float c = dot(N, H);
    if (c <= 0.0f) return 0.0f;
    const float a2 = a * a, c2 = c * c, c4 = c2 * c2;
    const float tan2 = (1.0f - c2) / c2;
    const float den = K_PI * c4 * (a2 + tan2) * (a2 + tan2);
    return a2 / den;
}
__device__ __forceinline__ float ggx_G(float a, V3 V, V3 L, V3 N, V3 H) {
    if (!((dot(V, H) * dot(V, N) > 0.0f) && (dot(L, H) * dot(L, N) > 0.0f))) return 0.0f;
    const float a2 = a * a;
    const float vn = dot(V, N), vn2 = vn * vn;
    const float g1v = 2.0f / (1.0f + sqrtf(1.0f + a2 * ((1.0f - vn2) / vn2)));
    const float ln = dot(L, N), ln2 = ln * ln;
    const float g1l = 2.0f / (1.0f + sqrtf(1.0f + a2 * ((1.0f - ln2) / ln2)));
    return g1v * g1l;
}
// SchlickConductorFresnel::evaluateReflectance (schlickConductorFresnel.cpp:28-41)
__device__ __forceinline__ V3 schlick(const DevMaterial& m, float cosi) {
    const float c = 1.0f - fabsf(cosi);
    const V3 f0 = v3(m.f0[0], m.f0[1], m.f0[2]);
    const V3 f0c = v3(1.f, 1.f, 1.f) - f0;
    return f0 + ((f0c * c) * (c * c)) * (c * c);
}

// VanillaDielectricFresnel::evaluateReflectance (vanillaDielectricFresnel.cpp:14-45)
__device__ __forceinline__ float dielectric_fresnel(float eta_outer, float eta_inner, float cos_i) {
    float ei = eta_outer, et = eta_inner, c = cos_i;
    if (c < 0.0f) {
        const float t = ei;
        ei = et;
        et = t;
        c *= -1.0f;
    }
    const float r = ei / et;
    const float st2 = r * r * (1.0f - c * c);
    if (st2 >= 1.0f) return 1.0f;
    const float ct = sqrtf(1.0f - st2);
    const float rpa = (et * c - ei * ct) / (et * c + ei * ct);
    const float rpe = (ei * c - et * ct) / (ei * c + et * ct);
    return 0.5f * (rpa * rpa + rpe * rpe);
}

// TVector::canRefract (tVector.ipp:467-503) of v about n
__device__ __forceinline__ bool refract(V3 v, V3 n, float eta_outer, float eta_inner, V3& out) {
    float ei = eta_outer, et = eta_inner, sf = 1.0f;
    const float ci = dot(v, n);
    if (ci < 0.0f) {
        const float t = ei;
        ei = et;
        et = t;
        sf = -1.0f;
    }
    const float r = ei / et;
    const float st2 = r * r * (1.0f - ci * ci);
    if (st2 >= 1.0f) return false;
    const float ct = sqrtf(1.0f - st2);
    V3 d = v * (-r);
    d = d + n * (((r * sf) * ci - ct) * sf);
    out = normalize(d);
    return true;
}

// PhongBsdf (phongBsdf.cpp:14-100): (R.V)^n with R = L reflected about Ns, times
// (n + 2) / 2pi (value) or (n + 1) / 2pi (pdf); zero across the surface
__device__ __forceinline__ float phong_power(float n, V3 V, V3 L, V3 Ns) {
    const V3 R = reflect(L, Ns);
    return powf(mmax(dot(R, V), 0.0f), n);
}

// Lambertian / GGX / Phong value and pdf for (wi = h.wi, wo = L); the mixed
// (plastic) BSDF's are ratio-weighted sums (mixedBsdf.cpp:35-63, 157-187);
// specular BSDFs and the absorber evaluate to zero.
// (lambertianDiffuse.cpp:27-44, 83-96; conductorMicrofacet.cpp:32-67, 123-156;
// phongBsdf.cpp:22-39, 83-100)
__device__ __forceinline__ void bsdf_eval_pdf(const DevScene& sc, const DevMaterial& m, const SurfaceHit& h, V3 L,
                                              V3& f, float& pdf) {
    const V3 Ns = h.Ns, V = h.wi;
    const float vn = dot(V, Ns), ln = dot(L, Ns);
    f = v3(0.f, 0.f, 0.f);
    pdf = 0.0f;
    if (m.type >= DM_MIRROR) return;
    if (vn * ln <= 0.0f) return;
    if (m.type == DM_LAMBERT) {
        f = mat_albedo(sc, m, h.uvw) * K_INV_PI;
        pdf = absdot(L, Ns) * K_INV_PI;
        return;
    }
    if (m.type == DM_PHONG || m.type == DM_PLASTIC) {
        const float pw = phong_power(m.exponent, V, L, Ns);
        const float fb = pw * ((m.exponent + 2.0f) * K_INV_TWO_PI);
        const float pb = pw * ((m.exponent + 1.0f) * K_INV_TWO_PI);
        if (m.type == DM_PHONG) {
            f = v3(fb, fb, fb);
            pdf = pb;
            return;
        }
        const float r = m.ratio, rc = 1.0f - r;
        const float avg = ((r + r) + r) / 3.0f;
        const V3 fa = mat_albedo(sc, m, h.uvw) * K_INV_PI;
        const float pa = absdot(L, Ns) * K_INV_PI;
        f = fa * r + v3(fb, fb, fb) * rc;
        pdf = pa * avg + pb * (1.0f - avg);
        return;
    }
    V3 H = (V + L) * msign(vn);
    if (is_zero(H)) return;
    H = normalize(H);
    const float a = mat_alpha(sc, m, h.uvw);
    const V3 F = schlick(m, dot(L, H));
    const float D = ggx_D(a, Ns, H);
    const float G = ggx_G(a, V, L, Ns, H);
    f = divs((F * G) * D, 4.0f * fabsf(vn * ln));
    const float pdfL = fabsf(fabsf(D * dot(Ns, H)) / (4.0f * dot(L, H)));
    pdf = isfinite(pdfL) ? pdfL : 0.0f;
}

// math::build_coordinate_system (math.cpp:36-58): Ns as the y axis
__device__ __forceinline__ void frame(V3 y, V3& z, V3& x) {
    if (fabsf(y.x) > fabsf(y.y)) z = divs(v3(-y.z, 0.f, y.x), sqrtf(y.x * y.x + y.z * y.z));
    else z = divs(v3(0.f, y.z, -y.y), sqrtf(y.y * y.y + y.z * y.z));
    x = cross(y, z);
}

struct BsdfSample {
    V3 f, L;
    float pdf;
    bool valid;
};

// Draws of a material's sample: (u0, u1) = sample.x, .y of the two-uniform
// BSDFs; uc = the lobe choice (plastic: mixedBsdf.cpp:81; glass:
// specularDielectric.cpp:49).
__device__ __forceinline__ bool mat_needs_choice(int32_t type) { return type == DM_PLASTIC || type == DM_GLASS; }
__device__ __forceinline__ bool mat_specular(int32_t type) {
    return type == DM_MIRROR || type == DM_GLASS || type == DM_TRANSMISSION;
}

// LambertianDiffuse::evaluateSample (lambertianDiffuse.cpp:46-81) with
// Hemisphere::cosineWeightedSampling (hemisphere.cpp:30-48)
__device__ __forceinline__ BsdfSample lambert_sample(const DevScene& sc, const DevMaterial& m, const SurfaceHit& h,
                                                     V3 x, V3 z, float u0, float u1) {
    BsdfSample s;
    const V3 Ns = h.Ns;
    const float phi = K_TWO_PI * u0;
    const float theta = acosf(1.0f - 2.0f * u1) * 0.5f;
    const float ct = cosf(theta);
    const float st = sqrtf(1.0f - ct * ct);
    const V3 l = v3(sinf(phi) * st, ct, cosf(phi) * st);
    V3 L = normalize((x * l.x + Ns * l.y) + z * l.z);
    if (dot(h.wi, Ns) <= 0.0f) L = -L;
    s.f = mat_albedo(sc, m, h.uvw) * K_INV_PI;
    s.L = L;
    s.pdf = ct * K_INV_PI;
    s.valid = !is_zero(s.f) && s.pdf > 0.0f;
    return s;
}

// PhongBsdf::evaluateSample (phongBsdf.cpp:41-81) with
// Hemisphere::cosineExpWeightedSampling (hemisphere.cpp:50-68): the lobe is
// sampled about Ns and weighed about the mirror direction, as the reference does
__device__ __forceinline__ BsdfSample phong_sample(const DevMaterial& m, const SurfaceHit& h, V3 x, V3 z, float u0,
                                                   float u1) {
    BsdfSample s;
    s.f = v3(0.f, 0.f, 0.f);
    s.L = v3(0.f, 0.f, 0.f);
    s.pdf = 0.0f;
    s.valid = false;
    const V3 Ns = h.Ns, V = h.wi;
    const float phi = K_TWO_PI * u0;
    const float ct = powf(u1, 1.0f / (m.exponent + 1.0f));
    const float st = sqrtf(1.0f - ct * ct);
    const V3 l = v3(sinf(phi) * st, ct, cosf(phi) * st);
    const V3 L = normalize((x * l.x + Ns * l.y) + z * l.z);
    if (dot(V, Ns) * dot(L, Ns) <= 0.0f) return s;
    const float pw = phong_power(m.exponent, V, L, Ns);
    const float fv = pw * ((m.exponent + 2.0f) * K_INV_TWO_PI);
    s.f = v3(fv, fv, fv);
    s.L = L;
    s.pdf = pw * ((m.exponent + 1.0f) * K_INV_TWO_PI);
    s.valid = !is_zero(s.f) && s.pdf > 0.0f;
    return s;
}

// Transmitted direction, value and pdf 1 of SpecularTransmission::evaluateSample
// (specularTransmission.cpp:35-78) and SpecularDielectric's refraction branch
// (specularDielectric.cpp:83-112); the radiance-mode factor (etaT / etaI)^2.
__device__ __forceinline__ bool transmit(const DevScene& sc, const DevMaterial& m, const SurfaceHit& h, BsdfSample& s) {
    float ei = m.ior_outer, et = m.ior_inner;
    V3 L;
    if (!refract(h.wi, h.Ns, ei, et, L)) return false;
    const float ci = dot(L, h.Ns);
    const float R = dielectric_fresnel(m.ior_outer, m.ior_inner, ci);
    if (ci < 0.0f) {
        const float t = ei;
        ei = et;
        et = t;
    }
    const float btdf = (et * et) / (ei * ei);
    const float T = 1.0f - R;
    s.f = divs((mat_albedo(sc, m, h.uvw) * T) * btdf, fabsf(ci));
    s.L = L;
    s.pdf = 1.0f;
    return true;
}

// Bsdf::evaluateSample of every material on this path, from uniforms
// (u0, u1) = sample.(x, y) and the lobe choice uc (lambertianDiffuse.cpp:46-81;
// conductorMicrofacet.cpp:69-121 with isotropicGgx.cpp:71-100; phongBsdf.cpp:41-81;
// mixedBsdf.cpp:65-155; specularReflection.cpp:32-57; specularDielectric.cpp:35-123;
// specularTransmission.cpp:35-78; absorberBsdf.cpp: never valid).
__device__ __forceinline__ BsdfSample bsdf_sample(const DevScene& sc, const DevMaterial& m, const SurfaceHit& h,
                                                  float u0, float u1, float uc) {
    BsdfSample s;
    s.f = v3(0.f, 0.f, 0.f);
    s.L = v3(0.f, 0.f, 0.f);
    s.pdf = 0.0f;
    s.valid = false;
    const V3 Ns = h.Ns, V = h.wi;
    if (m.type == DM_ABSORB) return s;
    if (m.type >= DM_MIRROR) {
        const float vn = dot(V, Ns);
        float R = 0.0f;
        bool reflect_lobe = m.type == DM_MIRROR;
        float pdf = 1.0f;
        if (m.type == DM_GLASS) {
            R = dielectric_fresnel(m.ior_outer, m.ior_inner, vn);
            const float prob = ((R + R) + R) / 3.0f;   // reflectance.average()
            reflect_lobe = uc < prob;
            pdf = reflect_lobe ? prob : 1.0f - prob;
        }
        if (reflect_lobe) {
            const V3 L = reflect(V, Ns * ((vn > 0.0f) ? 1.0f : -1.0f));
            const V3 alb = mat_albedo(sc, m, h.uvw);
            if (m.type == DM_MIRROR) {
                const float ln = dot(L, Ns);
                const V3 F = m.fresnel == 0 ? schlick(m, ln) : v3(1.f, 1.f, 1.f) * dielectric_fresnel(m.ior_outer, m.ior_inner, ln);
                s.f = divs(F * alb, fabsf(ln));
            } else {
                s.f = divs(alb * R, absdot(L, Ns));
            }
            s.L = L;
            s.pdf = pdf;
        } else {
            if (!transmit(sc, m, h, s)) return s;
            s.pdf = pdf;
        }
        s.valid = !is_zero(s.f) && s.pdf > 0.0f;
        return s;
    }
    V3 z, x;
    frame(Ns, z, x);
    if (m.type == DM_LAMBERT) return lambert_sample(sc, m, h, x, z, u0, u1);
    if (m.type == DM_PHONG) return phong_sample(m, h, x, z, u0, u1);
    if (m.type == DM_PLASTIC) {
        // MixedBsdf(A = Lambertian, B = Phong); choosing B weighs B's value and pdf
        // by the ratio and A's by its complement, as the reference does
        const float r = m.ratio, rc = 1.0f - r;
        const float avg = ((r + r) + r) / 3.0f;
        if (uc < avg) {
            const BsdfSample a = lambert_sample(sc, m, h, x, z, u0, u1);
            if (!a.valid) return s;
            const float pw = (dot(V, Ns) * dot(a.L, Ns) <= 0.0f) ? 0.0f : phong_power(m.exponent, V, a.L, Ns);
            const float fb = pw * ((m.exponent + 2.0f) * K_INV_TWO_PI);
            const float pb = pw * ((m.exponent + 1.0f) * K_INV_TWO_PI);
            s.f = a.f * r + v3(fb, fb, fb) * rc;
            s.pdf = avg * a.pdf + (1.0f - avg) * pb;
            s.L = a.L;
        } else {
            const BsdfSample b = phong_sample(m, h, x, z, u0, u1);
            if (!b.valid) return s;
            const bool same = dot(V, Ns) * dot(b.L, Ns) > 0.0f;
            const V3 fa = same ? mat_albedo(sc, m, h.uvw) * K_INV_PI : v3(0.f, 0.f, 0.f);
            const float pa = same ? absdot(b.L, Ns) * K_INV_PI : 0.0f;
            s.f = b.f * r + fa * rc;
            s.pdf = avg * b.pdf + (1.0f - avg) * pa;
            s.L = b.L;
        }
        s.valid = !is_zero(s.f) && s.pdf > 0.0f;
        return s;
    }
    // GGX conductor
    const float nf = (dot(V, Ns) > 0.0f) ? 1.0f : -1.0f;
    const float a = mat_alpha(sc, m, h.uvw);
    const float s0 = mclamp(u0, 0.0f, 0.9999f), s1 = mclamp(u1, 0.0f, 0.9999f);
    const float phi = K_TWO_PI * s0;
    const float theta = atanf(a * sqrtf(s1 / (1.0f - s1)));
    V3 hl;
    if (!isfinite(theta)) {
        hl = v3(0.f, 1.f, 0.f);
    } else {
        const float ct = cosf(theta);
        const float st = sqrtf(1.0f - ct * ct);
        hl = v3(sinf(phi) * st, ct, cosf(phi) * st);
    }
    const V3 H = normalize((x * hl.x + Ns * hl.y) + z * hl.z);
    const V3 L = reflect(V, H * nf);
    const float vn = dot(V, Ns), ln = dot(L, Ns), lh = dot(L, H), nh = dot(Ns, H);
    if (vn * ln <= 0.0f) return s;
    const V3 F = schlick(m, lh);
    const float D = ggx_D(a, Ns, H);
    const float G = ggx_G(a, V, L, Ns, H);
    const float pdfL = fabsf(fabsf(D * nh) / (4.0f * lh));
    if (!isfinite(pdfL)) return s;
    s.f = divs((F * G) * D, 4.0f * fabsf(vn * ln));
    s.L = L;
    s.pdf = pdfL;
    s.valid = !is_zero(s.f) && s.pdf > 0.0f;   // BsdfSample::isValid (bsdfSample.cpp:39-41)
    return s;
}

// AreaLight::_canEmit (areaLight.cpp:24-28)
__device__ __forceinline__ bool can_emit(const DevLight& l, V3 dir, V3 N) { return (dot(dir, N) >= 0.0f) != (l.back_face != 0); }

struct LightSample {
    V3 pos, radiance;
    float pdfW;
    bool valid;
};

// SingleAreaLight::evaluateDirectSample (singleAreaLight.cpp:54-85) with the
// primitive's position sample: Triangle (triangle.cpp:162-193, sampleWarp.cpp:10-19)
// or Rectangle (rectangle.cpp:134-169). The rectangle's rejection loop over
// (s, t) is replaced by t = u * short/long, the accepted t's distribution.
__device__ __forceinline__ LightSample light_sample(const DevLight& l, V3 target, float u0, float u1) {
    LightSample r;
    r.valid = false;
    r.pdfW = 0.0f;
    r.pos = v3(0.f, 0.f, 0.f);
    r.radiance = v3(0.f, 0.f, 0.f);
    const V3 A = load3(l.a), e1 = load3(l.e1), e2 = load3(l.e2);
    V3 P, Ns;
    V3 N = cross(e1, e2);
    N = is_zero(N) ? v3(0.f, 1.f, 0.f) : normalize(N);
    if (l.type == 1) {
        V3 ea = e1, eb = e2;
        if (len(ea) < len(eb)) {
            const V3 t = ea;
            ea = eb;
            eb = t;
        }
        const float lw = len(ea), sw = len(eb);
        const float s = u0, t = u1 * (sw / lw);
        P = (A + ea * s) + normalize(eb) * (t * len(ea));
        Ns = N;
    } else {
        const float sq = sqrtf(u0);
        P = (A + e1 * (1.0f - sq)) + e2 * (sq * u1);
        if (l.has_normals) {
            const V3 b = barycentric(A, e1, e2, P);
            const V3 n = bary_mix(b, load3(l.n), load3(l.n + 3), load3(l.n + 6));
            Ns = is_zero(n) ? N : n;   // not renormalized (triangle.cpp:182-183)
        } else {
            const V3 b = barycentric(A, e1, e2, P);
            const V3 n = bary_mix(b, N, N, N);
            Ns = is_zero(n) ? N : n;
        }
    }
    if (!(l.pdfA > 0.0f)) return r;
    const V3 ev = target - P;
    if (!can_emit(l, ev, Ns)) return r;
    const float d2 = len2(ev);
    const float c = absdot(divs(ev, sqrtf(d2)), Ns);
    if (c <= 0.0f) return r;
    r.pos = P;
    r.radiance = load3(l.radiance);
    r.pdfW = l.pdfA * d2 / c;
    r.valid = !is_zero(r.radiance) && r.pdfW > 0.0f;   // DirectLightSample::isValid
    return r;
}

// SingleAreaLight::emittance (singleAreaLight.cpp:39-52)
__device__ __forceinline__ V3 light_emittance(const DevLight& l, const SurfaceHit& h) {
    return can_emit(l, h.wi, h.Ns) ? load3(l.radiance) : v3(0.f, 0.f, 0.f);
}
// SingleAreaLight::evaluateDirectPdfW (singleAreaLight.cpp:87-109)
__device__ __forceinline__ float light_pdfW(const DevLight& l, const SurfaceHit& h, V3 target) {
    const V3 ev = target - h.P;
    if (!can_emit(l, h.wi, h.Ns)) return 0.0f;
    const float c = absdot(h.wi, h.Ns);
    if (c <= 0.0f) return 0.0f;
    return l.pdfA * len2(ev) / c;
}

// ---- environment light -----------------------------------------------------------
// The background of a scene with an environment light is an absorbing infinite
// sphere that every unbounded ray missing the scene hits (scene.cpp:30-63,
// infiniteSphere.cpp:35-47); its surface position is the ray's direction.


// Distribution1D::continuousToDiscrete (distribution1D.cpp:115-125): the last
// i in [0, n) with cdf[i] <= u (a binary search over the nondecreasing CDF).
__device__ __forceinline__ int32_t cdf_index(const float* cdf, int32_t n, float u) {
    int32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (cdf[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Distribution1D::sampleContinuous (distribution1D.cpp:60-80)
__device__ __forceinline__ float cdf_sample(const float* cdf, int32_t n, float u, float& pdf, int32_t& index) {
    const float delta = 1.0f / static_cast<float>(n);
    const int32_t i = cdf_index(cdf, n, u);
    index = i;
    pdf = (cdf[i + 1] - cdf[i]) / delta;
    float d = u - cdf[i];
    if (cdf[i + 1] > cdf[i]) d /= cdf[i + 1] - cdf[i];
    return mclamp((static_cast<float>(i) + d) * delta, 0.0f, 1.0f);
}

// Distribution2D::pdfContinuous (distribution2D.cpp:46-53)
__device__ __forceinline__ float env_pdf_uv(const DevScene& sc, const DevLight& l, float u, float v) {
    const float* m = sc.env_cdf + l.env_cdf;
    const int32_t iy = cdf_index(m, l.env_h, v);
    const float pdf_y = (m[iy + 1] - m[iy]) / (1.0f / static_cast<float>(l.env_h));
    const float* c = m + (l.env_h + 1) + static_cast<int64_t>(iy) * (l.env_w + 1);
    const int32_t ix = cdf_index(c, l.env_w, u);
    const float pdf_x = (c[ix + 1] - c[ix]) / (1.0f / static_cast<float>(l.env_w));
    return pdf_x * pdf_y;
}

// EnvironmentLight::emittance at the background hit of direction d
// (environmentLight.cpp:58-64)
__device__ __forceinline__ V3 env_emittance(const DevScene& sc, const DevLight& l, V3 d) {
    return tex_lookup(sc, l.env_tex, spherical_uvw(d));
}

// EnvironmentLight::evaluateDirectPdfW at the background hit of direction d
// (environmentLight.cpp:90-104)
__device__ __forceinline__ float env_pdfW(const DevScene& sc, const DevLight& l, V3 d) {
    const V3 uvw = spherical_uvw(d);
    const float sin_theta = sinf((1.0f - uvw.y) * K_PI);
    if (sin_theta <= 0.0f) return 0.0f;
    return env_pdf_uv(sc, l, uvw.x, uvw.y) / (((2.0f * K_PI) * K_PI) * sin_theta);
}

// EnvironmentLight::evaluateDirectSample with sample (sx, sy)
// (environmentLight.cpp:66-88), Distribution2D::sampleContinuous
// (distribution2D.cpp:30-44) and InfiniteSphere::uvwToPosition
// (infiniteSphere.cpp:61-78, z up: the frame disagrees with the lookup's).
__device__ __forceinline__ LightSample env_light_sample(const DevScene& sc, const DevLight& l, float sx, float sy) {
    LightSample r;
    r.valid = false;
    r.pdfW = 0.0f;
    r.pos = v3(0.f, 0.f, 0.f);
    r.radiance = v3(0.f, 0.f, 0.f);
    const float* m = sc.env_cdf + l.env_cdf;
    float pdf_y, pdf_x;
    int32_t iy, ix;
    const float v = cdf_sample(m, l.env_h, sy, pdf_y, iy);
    const float* c = m + (l.env_h + 1) + static_cast<int64_t>(iy) * (l.env_w + 1);
    const float u = cdf_sample(c, l.env_w, sx, pdf_x, ix);
    const float uv_pdf = pdf_x * pdf_y;
    const float theta = (1.0f - v) * K_PI;
    const float phi = u * K_TWO_PI;
    const float ct = cosf(theta);
    const float st = sqrtf(1.0f - ct * ct);
    const V3 dir = v3(cosf(phi) * st, sinf(phi) * st, ct);
    const V3 pos = normalize(dir) * 1000000.0f;
    const float sin_theta = sinf((1.0f - v) * K_PI);
    if (sin_theta <= 0.0f) return r;
    r.pos = pos;
    r.radiance = tex_lookup(sc, l.env_tex, v3(u, v, 0.0f));
    r.pdfW = uv_pdf / (((2.0f * K_PI) * K_PI) * sin_theta);
    r.valid = !is_zero(r.radiance) && r.pdfW > 0.0f;
    return r;
}

// The picked light's direct sample: area or environment.
__device__ __forceinline__ LightSample any_light_sample(const DevScene& sc, const DevLight& l, V3 target, float u0,
                                                        float u1) {
    return l.type == LIGHT_TYPE_ENV ? env_light_sample(sc, l, u0, u1) : light_sample(l, target, u0, u1);
}

// TMis<POWER>::weight (tMis.h:21-47)
__device__ __forceinline__ float mis_power(float a, float b) {
    const float a2 = a * a, b2 = b * b;
    return a2 / (a2 + b2);
}

}  // namespace mpt
