/*
 * Oracle (test infrastructure): scene assembly, surface details, BSDFs,
 * area lights, camera, sampler, film, and the PtEstimator bounce loop.
 * Each function cites the reference function it restates.
 */
#include "or_internal.h"
#include "seqrng.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---- random numbers --------------------------------------------------------- */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC 2011). */
void or_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline float u32_to_unit(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

/* Slots of the Philox layout within one bounce b: counter (pixel, sample, 2b + slot/4, 0);
 * the lobe choices of two-lobe samples (plastic, glass) at slots 8 and 9: counter
 * (pixel, sample, 2b, 1). */
enum {
    SL_PICK = 0, SL_LPOS0 = 1, SL_LPOS1 = 2, SL_RR = 3, SL_MIS0 = 4, SL_MIS1 = 5, SL_CONT0 = 6, SL_CONT1 = 7,
    SL_MISC = 8, SL_CONTC = 9
};

typedef struct {
    int      mode;
    uint64_t state;
    uint32_t key[2];
    uint32_t pixel, sample;
    int64_t  draws;
} or_rng;

static float rng_u(or_rng* r, int bounce, int slot) {
    r->draws++;
    if (r->mode == OR_RNG_SEQ) return seqrng_next(&r->state);
    const uint32_t ctr[4] = {r->pixel, r->sample, (uint32_t)(2 * bounce + (slot & 7) / 4), slot >= 8 ? 1u : 0u};
    uint32_t out[4];
    or_philox(ctr, r->key, out);
    return u32_to_unit(out[slot & 3]);
}

/* Vector2R(Random::nextReal(), Random::nextReal()): g++ evaluates the two
 * constructor arguments right to left, so the first draw lands in y. */
static void rng_u2(or_rng* r, int bounce, int slot_x, int slot_y, float* x, float* y) {
    if (r->mode == OR_RNG_SEQ) {
        *y = rng_u(r, bounce, slot_y);
        *x = rng_u(r, bounce, slot_x);
    } else {
        *x = rng_u(r, bounce, slot_x);
        *y = rng_u(r, bounce, slot_y);
    }
}

/* ---- scene assembly ---------------------------------------------------------- */

static float prim_area(const or_prim* p) {
    if (p->type == OR_PRIM_SPHERE) return (float)12.566370614359172 * p->radius * p->radius; /* sphere.cpp:142-144 */
    if (p->type == OR_PRIM_RECTANGLE) return v3_len(p->e1) * v3_len(p->e2); /* rectangle.cpp:179-181 */
    return 0.5f * v3_len(v3_cross(p->e1, p->e2));                            /* triangle.cpp:199-201 */
}

/* 4x4 float product as TMatrix4::operator*= (tMatrix4.ipp:102-116) */
static void mat_mul(float* a, const float* b) {
    float tmp[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            float acc = 0.0f;
            for (int k = 0; k < 4; ++k) acc += a[4 * r + k] * b[4 * k + c];
            tmp[4 * r + c] = acc;
        }
    memcpy(a, tmp, sizeof(tmp));
}

/* PerspectivePinholeCamera ctor + updateTransform (perspectivePinholeCamera.cpp:14-50),
 * TMatrix4::lookAt / translate / scale (tMatrix4.ipp:25-60), Camera::_getAspectRatio
 * (camera.cpp:16-18) */
static void camera_setup(or_scene* s) {
    const or_scene_desc* d = &s->desc;
    const v3 pos = v3_load(d->cam_pos), dir = v3_load(d->cam_dir), up = v3_load(d->cam_up);
    const v3 nz = v3_normalize(v3_neg(dir));
    const v3 nx = v3_normalize(v3_cross(up, nz));
    const v3 ny = v3_cross(nz, nx);
    const float c2w[16] = {nx.x, ny.x, nz.x, pos.x, nx.y, ny.y, nz.y, pos.y,
                           nx.z, ny.z, nz.z, pos.z, 0.0f, 0.0f, 0.0f, 1.0f};
    memcpy(s->camera_to_world, c2w, sizeof(c2w));
    s->cam_pos = pos;

    const double aspect  = (double)(size_t)d->width / (double)(size_t)d->height;
    const double sensorW = (double)d->sensor_width_mm;
    const double sensorH = sensorW / aspect;
    const float  resX = (float)(size_t)d->width, resY = (float)(size_t)d->height;
    const double fov = (double)d->fov_deg;
    const float  halfRad = (float)(fov * 0.5) * OR_DEG2RAD;
    const double offset = (sensorW * 0.5) / (double)tanf(halfRad);
    float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const float tr[16] = {1, 0, 0, (float)(-(sensorW * 0.5)), 0, 1, 0, (float)(sensorH * 0.5),
                          0, 0, 1, (float)(-offset), 0, 0, 0, 1};
    const float sc[16] = {(float)(sensorW / (double)resX), 0, 0, 0, 0, (float)(-sensorH / (double)resY), 0, 0,
                          0, 0, 1.0f, 0, 0, 0, 0, 1};
    mat_mul(m, tr);
    mat_mul(m, sc);
    memcpy(s->film_to_camera, m, sizeof(m));
}

/* ---- environment light ---------------------------------------------------------- */

/* Distribution1D ctor (distribution1D.cpp:10-41) */
static void dist1_init(or_dist1* d, const float* value, int64_t n) {
    d->n = n;
    d->cdf = (float*)calloc((size_t)n + 1, sizeof(float));
    d->delta = 1.0f / (float)(size_t)n;
    d->cdf[0] = 0.0f;
    for (int64_t i = 1; i <= n; ++i) d->cdf[i] = d->cdf[i - 1] + value[i - 1] * d->delta;
    const float sum = d->cdf[n];
    if (sum > 0.0f) {
        const float inv = 1.0f / sum;
        for (int64_t i = 1; i <= n; ++i) d->cdf[i] *= inv;
    } else {
        for (int64_t i = 1; i <= n; ++i) d->cdf[i] = (float)(size_t)i * d->delta;
    }
    d->cdf[n] = 1.0f;
}

/* Distribution1D::continuousToDiscrete (distribution1D.cpp:115-125): scan down from n - 1 */
static int64_t dist1_index(const or_dist1* d, float u) {
    for (int64_t i = d->n - 1; i >= 0; --i)
        if (d->cdf[i] <= u) return i;
    return 0;
}

/* Distribution1D::pdfContinuous(index) (distribution1D.cpp:105-107) */
static float dist1_pdf(const or_dist1* d, int64_t i) { return (d->cdf[i + 1] - d->cdf[i]) / d->delta; }

/* Distribution1D::sampleContinuous (distribution1D.cpp:59-81) */
static float dist1_sample(const or_dist1* d, float u, float* pdf, int64_t* index) {
    const int64_t i = dist1_index(d, u);
    *index = i;
    *pdf = dist1_pdf(d, i);
    float delta = u - d->cdf[i];
    if (d->cdf[i + 1] > d->cdf[i]) delta /= d->cdf[i + 1] - d->cdf[i];
    const float x = ((float)(size_t)i + delta) * d->delta;
    return m_clamp(x, 0.0f, 1.0f);
}

static v3 tex_eval(const or_texture* t, v3 uvw);

/* EnvironmentLight ctor (environmentLight.cpp:16-57) with RgbSpectrum::luminance
 * (rgbSpectrum.cpp:33-37) and Distribution2D ctor (distribution2D.cpp:10-28) */
static void env_init(or_light_state* L, const or_texture* t) {
    L->env = *t;
    const int64_t w = t->width > 0 ? t->width : 1, h = t->height > 0 ? t->height : 1;
    float* weighted = (float*)calloc((size_t)(w * h), sizeof(float));
    L->background_flux = 0.0f;
    for (int64_t iy = 0; iy < h; ++iy) {
        const float v = ((float)(size_t)iy + 0.5f) / (float)(size_t)h;
        const float sinTheta = sinf((1.0f - v) * OR_PI);
        for (int64_t ix = 0; ix < w; ++ix) {
            const float u = ((float)(size_t)ix + 0.5f) / (float)(size_t)w;
            const v3 rad = tex_eval(t, v3_make(u, v, 0.0f));
            const float lum = 0.0126f * rad.x + 0.7152f * rad.y + 0.0722f * rad.z;
            weighted[iy * w + ix] = lum * sinTheta;
            L->background_flux += weighted[iy * w + ix];
        }
    }
    float* rows = (float*)calloc((size_t)h, sizeof(float));
    L->conditional = (or_dist1*)calloc((size_t)h, sizeof(or_dist1));
    for (int64_t iy = 0; iy < h; ++iy) {
        for (int64_t ix = 0; ix < w; ++ix) rows[iy] += weighted[iy * w + ix];
        dist1_init(&L->conditional[iy], &weighted[iy * w], w);
    }
    dist1_init(&L->marginal, rows, h);
    L->radius = 5000.0f; /* environmentLight.cpp:55-56 */
    free(rows);
    free(weighted);
}

/* RenderDatabase::prepareRender (renderDatabase.cpp:100-111) with
 * _setUpLight (renderDatabase.cpp:181-189): the scene-bound radius goes to the
 * light recorded as the environment light, which is the last light declared
 * once an environment light exists; it reaches the environment light only when
 * that light is the last one. Bound = the BVH root's (bvhAccelerator.cpp:34-38),
 * isEmpty (tAabb.ipp:30-34), extent().length() (tVector.ipp:316-328). */
static void env_radius(or_scene* s) {
    if (s->env_light < 0 || s->env_light != s->nlights - 1) return;
    const or_aabb b = s->nodes[0].bound;
    const int empty = (b.mn.x == FLT_MAX && b.mn.y == FLT_MAX && b.mn.z == FLT_MAX) ||
                      (b.mx.x == FLT_MIN && b.mx.y == FLT_MIN && b.mx.z == FLT_MIN);
    if (empty) return;
    const v3 e = v3_sub(b.mx, b.mn);
    float l2 = 0.0f;
    l2 += e.x * e.x;
    l2 += e.y * e.y;
    l2 += e.z * e.z;
    s->lights[s->env_light].radius = sqrtf(l2) * 0.5f;
}

/* PowerLightCluster + Distribution1D (powerLightCluster.cpp:9-25, distribution1D.cpp:10-40);
 * approximateFlux: SingleAreaLight (singleAreaLight.cpp:168-184), EnvironmentLight
 * (environmentLight.cpp:137-139); luminance (rgbSpectrum.cpp:33-37) */
static void power_cdf_init(or_scene* s) {
    const or_scene_desc* desc = &s->desc;
    s->power_cdf = (float*)calloc((size_t)desc->nlights + 1, sizeof(float));
    if (desc->nlights <= 0) return;
    const float delta = 1.0f / (float)(size_t)desc->nlights;
    s->power_cdf[0] = 0.0f;
    for (int32_t i = 1; i <= desc->nlights; ++i) {
        or_light_state* L = &s->lights[i - 1];
        float flux;
        if (L->type == OR_LIGHT_ENVIRONMENT) {
            flux = L->background_flux * (float)12.566370614359172 * L->radius * L->radius;
        } else {
            const v3 le = L->radiance;
            const float lum = 0.0126f * le.x + 0.7152f * le.y + 0.0722f * le.z;
            flux = lum * prim_area(&s->prims[L->prim]) * OR_PI;
            if (flux <= 0.0f) flux = 1.0f;
        }
        L->flux = flux;
        s->power_cdf[i] = s->power_cdf[i - 1] + flux * delta;
    }
    const float sum = s->power_cdf[desc->nlights];
    if (sum > 0.0f) {
        const float inv = 1.0f / sum;
        for (int32_t i = 1; i <= desc->nlights; ++i) s->power_cdf[i] *= inv;
    } else {
        for (int32_t i = 1; i <= desc->nlights; ++i) s->power_cdf[i] = (float)(size_t)i * delta;
    }
    s->power_cdf[desc->nlights] = 1.0f;
}

or_scene* or_scene_create(const or_scene_desc* desc) {
    or_scene* s = (or_scene*)calloc(1, sizeof(or_scene));
    s->desc = *desc;
    s->nprims = desc->nprims;
    s->prims = (or_prim*)calloc((size_t)(desc->nprims > 0 ? desc->nprims : 1), sizeof(or_prim));
    for (int64_t i = 0; i < desc->nprims; ++i) {
        or_prim* p = &s->prims[i];
        const float* v = desc->prim_v + 9 * i;
        p->type = desc->prim_type[i];
        p->vA = v3_load(v); p->vB = v3_load(v + 3); p->vC = v3_load(v + 6);
        p->mat = desc->prim_mat[i];
        p->light = -1;
        if (p->type == OR_PRIM_SPHERE) {
            /* Sphere ctor (sphere.cpp:18-31): centre and radius */
            p->radius = v[3];
            continue;
        }
        if (p->type == OR_PRIM_RECTANGLE) {
            /* Rectangle ctor (rectangle.cpp:14-37) */
            p->e1 = v3_sub(p->vA, p->vB);
            p->e2 = v3_sub(p->vC, p->vB);
            p->vD = v3_add(v3_add(p->vB, p->e1), p->e2);
            p->uvwA = v3_make(1, 0, 0); p->uvwB = v3_make(0, 0, 0);
            p->uvwC = v3_make(0, 1, 0); p->uvwD = v3_make(1, 1, 0);
        } else {
            /* Triangle ctor (triangle.cpp:15-42) + TriangleMesh setters (triangleMesh.cpp:20-40) */
            p->e1 = v3_sub(p->vB, p->vA);
            p->e2 = v3_sub(p->vC, p->vA);
            v3 N = v3_cross(p->e1, p->e2);
            N = v3_is_zero(N) ? v3_make(0.0f, 1.0f, 0.0f) : v3_normalize(N);
            p->nA = p->nB = p->nC = N;
            p->uvwA = v3_make(0, 0, 0); p->uvwB = v3_make(1, 0, 0); p->uvwC = v3_make(0, 1, 0);
            if (desc->prim_n) {
                const float* n = desc->prim_n + 9 * i;
                const v3 a = v3_load(n), b = v3_load(n + 3), c = v3_load(n + 6);
                if (!v3_is_zero(a)) p->nA = a;
                if (!v3_is_zero(b)) p->nB = b;
                if (!v3_is_zero(c)) p->nC = c;
            }
            if (desc->prim_uvw && !isnan(desc->prim_uvw[9 * i])) {
                const float* u = desc->prim_uvw + 9 * i;
                p->uvwA = v3_load(u); p->uvwB = v3_load(u + 3); p->uvwC = v3_load(u + 6);
            }
        }
    }
    s->nmats = desc->nmats;
    s->mats = (or_material*)calloc((size_t)(desc->nmats > 0 ? desc->nmats : 1), sizeof(or_material));
    memcpy(s->mats, desc->mats, sizeof(or_material) * (size_t)desc->nmats);
    s->nlights = desc->nlights;
    s->env_light = -1;
    s->lights = (or_light_state*)calloc((size_t)(desc->nlights > 0 ? desc->nlights : 1), sizeof(or_light_state));
    for (int32_t i = 0; i < desc->nlights; ++i) {
        const or_light* L = &desc->lights[i];
        s->lights[i].type = L->type;
        s->lights[i].prim = -1;
        if (L->type == OR_LIGHT_ENVIRONMENT) {
            env_init(&s->lights[i], &L->env);
            s->env_light = i; /* the background sphere of the last environment light */
            continue;
        }
        /* SingleAreaLight ctor (singleAreaLight.cpp:20-37) */
        const v3 color = v3_load(L->color);
        const v3 unit = v3_divs(color, v3_sum(color));
        const v3 total = v3_scale(unit, L->watt);
        const float area = prim_area(&s->prims[L->prim]);
        s->lights[i].prim = L->prim;
        s->lights[i].radiance = v3_scale(v3_divs(total, area), OR_INV_PI);
        s->lights[i].back_face = L->back_face;
        s->prims[L->prim].light = i; /* Primitive::setAreaLight: the last binding wins */
    }
    camera_setup(s);
    /* StratifiedSampler ctor (stratifiedSampler.cpp:14-20), math::nearest_lower_square_number (math.cpp:18-23) */
    {
        const size_t lower = (size_t)floorf(sqrtf((float)(size_t)desc->spp));
        s->strata_base = (int)(size_t)sqrt((double)(lower * lower));
    }
    or_build_bvh(s);
    env_radius(s);
    power_cdf_init(s);
    return s;
}

void or_scene_destroy(or_scene* s) {
    if (!s) return;
    for (int32_t i = 0; i < s->nlights; ++i) {
        if (s->lights[i].type != OR_LIGHT_ENVIRONMENT) continue;
        for (int32_t r = 0; r < s->lights[i].env.height; ++r) free(s->lights[i].conditional[r].cdf);
        free(s->lights[i].conditional);
        free(s->lights[i].marginal.cdf);
    }
    free(s->prims); free(s->mats); free(s->lights); free(s->power_cdf);
    free(s->nodes); free(s->ordered);
    free(s);
}

/* ---- surface interaction ------------------------------------------------------- */

typedef struct {
    int32_t prim;
    v3 P, Ng, Ns, uvw, wi, wo;
} or_si;

static int background_hit(const or_scene* s, v3 d, float maxT, or_si* si);
static v3 spherical_uvw(v3 d);

/* Triangle::_positionToBarycentric (triangle.cpp:233-262) */
static v3 barycentric(const or_prim* p, v3 P) {
    const v3 v0 = v3_sub(p->vB, p->vA), v1 = v3_sub(p->vC, p->vA), v2 = v3_sub(P, p->vA);
    const float d00 = v3_dot(v0, v0), d01 = v3_dot(v0, v1), d11 = v3_dot(v1, v1);
    const float d20 = v3_dot(v2, v0), d21 = v3_dot(v2, v1);
    const float den = d00 * d11 - d01 * d01;
    if (den == 0.0f) return v3_make(0, 0, 0);
    const float inv = 1.0f / den;
    const float u = (d11 * d20 - d01 * d21) * inv;
    const float v = (d00 * d21 - d01 * d20) * inv;
    return v3_make(1.0f - u - v, u, v);
}

static v3 bary_mix(v3 b, v3 a, v3 bb, v3 c) {
    return v3_add(v3_add(v3_scale(a, b.x), v3_scale(bb, b.y)), v3_scale(c, b.z));
}

/* Scene::isIntersecting detail (scene.cpp:30-48), Triangle::evaluateSurfaceDetail
 * (triangle.cpp:125-160), Rectangle::evaluateSurfaceDetail (rectangle.cpp:102-132) */
static void surface_detail(const or_scene* s, int32_t prim, v3 o, v3 d, float t, or_si* si) {
    const or_prim* p = &s->prims[prim];
    si->prim = prim;
    si->wi = v3_neg(d);
    si->P = v3_add(o, v3_scale(d, t));
    if (p->type == OR_PRIM_SPHERE) {
        /* Sphere::evaluateSurfaceDetail (sphere.cpp:109-130) */
        const v3 N = v3_normalize(v3_sub(si->P, p->vA));
        si->Ng = si->Ns = N;
        si->uvw = spherical_uvw(N);
        return;
    }
    if (p->type == OR_PRIM_RECTANGLE) {
        v3 N = v3_cross(p->e1, p->e2);
        N = v3_is_zero(N) ? v3_make(0.0f, 1.0f, 0.0f) : v3_normalize(N);
        si->Ng = si->Ns = N;
        const v3 onPlane = v3_sub(si->P, p->vB);
        const float p1 = v3_dot(onPlane, v3_normalize(p->e1)) / v3_len(p->e1);
        const float p2 = v3_dot(onPlane, v3_normalize(p->e2)) / v3_len(p->e1);
        const v3 l1 = v3_lerp(p->uvwB, p->uvwC, p2);
        const v3 l2 = v3_lerp(p->uvwA, p->uvwD, p2);
        si->uvw = v3_lerp(l1, l2, p1);
    } else {
        v3 Ng = v3_cross(p->e1, p->e2);
        Ng = v3_is_zero(Ng) ? v3_make(0.0f, 1.0f, 0.0f) : v3_normalize(Ng);
        const v3 b = barycentric(p, si->P);
        si->uvw = bary_mix(b, p->uvwA, p->uvwB, p->uvwC);
        v3 Ns = bary_mix(b, p->nA, p->nB, p->nC);
        Ns = v3_is_zero(Ns) ? Ng : v3_normalize(Ns);
        si->Ns = Ns;
        si->Ng = Ng;
    }
}

static int scene_intersect(const or_scene* s, v3 o, v3 d, float minT, float maxT, or_si* si, float* tout) {
    int32_t prim = -1;
    float t = maxT;
    if (!or_closest(s, o, d, minT, &t, &prim)) return 0;
    surface_detail(s, prim, o, d, t, si);
    if (tout) *tout = t;
    return 1;
}

void or_intersect_detail(const or_scene* s, const float* rays, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        float* o = out + 14 * i;
        or_si si;
        float t;
        if (scene_intersect(s, v3_load(r), v3_load(r + 4), r[3], r[7], &si, &t)) {
            o[0] = (float)si.prim; o[1] = t;
            v3_store(o + 2, si.P); v3_store(o + 5, si.Ng); v3_store(o + 8, si.Ns); v3_store(o + 11, si.uvw);
        } else if (background_hit(s, v3_load(r + 4), r[7], &si)) {
            /* the background sphere (prim -2): maxT becomes its radius (infiniteSphere.cpp:36-45) */
            o[0] = -2.0f; o[1] = 1000000.0f;
            v3_store(o + 2, si.P); v3_store(o + 5, si.Ng); v3_store(o + 8, si.Ns); v3_store(o + 11, si.uvw);
        } else {
            memset(o, 0, 14 * sizeof(float));
            o[0] = -1.0f;
        }
    }
}

/* ---- textures (tConstantTexture.ipp:14-19, tNearestPixelSampler.ipp:13-31,
 *      repeatUvwWrapper.cpp:9-15, clampUvwWrapper.cpp:8-12) ------------------- */

/* TCheckerboardTexture::evaluate (tCheckerboardTexture.ipp:28-40): the even
 * texture where floor(u / oddSize) + floor(v / evenSize) is even (C++ %, so
 * negative odd sums count as odd), else the odd texture. */
static const or_texture* tex_resolve(const or_texture* t, v3 uvw) {
    while (t->odd && t->even) {
        const int32_t uc = (int32_t)floorf(uvw.x / t->odd_size);
        const int32_t vc = (int32_t)floorf(uvw.y / t->even_size);
        t = ((uc + vc) % 2 == 0) ? t->even : t->odd;
    }
    return t;
}

static v3 tex_eval(const or_texture* t, v3 uvw) {
    t = tex_resolve(t, uvw);
    if (!t->data) return v3_make(t->value[0], t->value[1], t->value[2]);
    float u, v;
    if (t->wrap_clamp) { u = m_clamp(uvw.x, 0.0f, 1.0f); v = m_clamp(uvw.y, 0.0f, 1.0f); }
    else { u = uvw.x - floorf(uvw.x); v = uvw.y - floorf(uvw.y); }
    int32_t x = (int32_t)(u * (float)t->width);
    int32_t y = (int32_t)(v * (float)t->height);
    x = (x < t->width) ? x : x - 1;
    y = (y < t->height) ? y : y - 1;
    const float* px = t->data + ((size_t)y * (size_t)t->width + (size_t)x) * (size_t)t->channels;
    if (t->channels == 1) return v3_make(px[0], px[0], px[0]);
    return v3_make(px[0], px[1], px[2]);
}

/* ---- BSDFs ---------------------------------------------------------------------- */

typedef struct { v3 f; v3 L; float pdf; int valid; } or_bsdf_sample;

/* math::build_coordinate_system (math.cpp:36-58) */
static void frame(v3 y, v3* z, v3* x) {
    if (fabsf(y.x) > fabsf(y.y)) {
        v3 zz = v3_make(-y.z, 0.0f, y.x);
        *z = v3_divs(zz, sqrtf(y.x * y.x + y.z * y.z));
    } else {
        v3 zz = v3_make(0.0f, y.z, -y.y);
        *z = v3_divs(zz, sqrtf(y.y * y.y + y.z * y.z));
    }
    *x = v3_cross(y, *z);
}
static v3 to_world(v3 x, v3 y, v3 z, v3 l) {
    return v3_add(v3_add(v3_scale(x, l.x), v3_scale(y, l.y)), v3_scale(z, l.z));
}

/* TRoughnessMapper<SQUARE>::map (tRoughnessMapper.h:18-40) */
static float rough_to_alpha(float r) {
    const float safe = m_max(r, 0.001f);
    return m_max(safe * safe, 0.001f);
}
/* IsotropicGgx::distributionD (isotropicGgx.cpp:11-34) */
static float ggx_D(float a, v3 N, v3 H) {
    const float NdotH = v3_dot(N, H);
    if (NdotH <= 0.0f) return 0.0f;
    const float a2 = a * a, c2 = NdotH * NdotH, c4 = c2 * c2;
    const float tan2 = (1.0f - c2) / c2;
    const float den = OR_PI * c4 * (a2 + tan2) * (a2 + tan2);
    return a2 / den;
}
/* IsotropicGgx::shadowingMaskingG (isotropicGgx.cpp:36-69), Microfacet::_isShadowingMaskingValid (microfacet.cpp:11-23) */
static float ggx_G(float a, v3 V, v3 L, v3 N, v3 H) {
    const float VdotH = v3_dot(V, H), VdotN0 = v3_dot(V, N), LdotH = v3_dot(L, H), LdotN0 = v3_dot(L, N);
    if (!((VdotH * VdotN0 > 0.0f) && (LdotH * LdotN0 > 0.0f))) return 0.0f;
    const float a2 = a * a;
    const float VdotN = v3_dot(V, N), VdotN2 = VdotN * VdotN;
    const float tanV2 = (1.0f - VdotN2) / VdotN2;
    const float G1V = 2.0f / (1.0f + sqrtf(1.0f + a2 * tanV2));
    const float LdotN = v3_dot(L, N), LdotN2 = LdotN * LdotN;
    const float tanL2 = (1.0f - LdotN2) / LdotN2;
    const float G1L = 2.0f / (1.0f + sqrtf(1.0f + a2 * tanL2));
    return G1V * G1L;
}
/* SchlickConductorFresnel::evaluateReflectance (schlickConductorFresnel.cpp:28-41) */
static v3 schlick(v3 f0, float cosI) {
    const float c = 1.0f - fabsf(cosI);
    const v3 f0c = v3_complement(f0);
    return v3_add(f0, v3_scale(v3_scale(v3_scale(f0c, c), c * c), c * c));
}

static float mat_alpha(const or_material* m, v3 uvw) { return rough_to_alpha(tex_eval(&m->roughness, uvw).x); }

/* VanillaDielectricFresnel::evaluateReflectance (vanillaDielectricFresnel.cpp:14-45) */
static float dielectric_fresnel(float iorOuter, float iorInner, float cosThetaI) {
    float etaI = iorOuter, etaT = iorInner, cosI = cosThetaI;
    if (cosI < 0.0f) {
        const float t = etaI; etaI = etaT; etaT = t;
        cosI *= -1.0f;
    }
    const float etaRatio = etaI / etaT;
    const float sinT2 = etaRatio * etaRatio * (1.0f - cosI * cosI);
    if (sinT2 >= 1.0f) return 1.0f;
    const float cosT = sqrtf(1.0f - sinT2);
    const float rPar = (etaT * cosI - etaI * cosT) / (etaT * cosI + etaI * cosT);
    const float rPerp = (etaI * cosI - etaT * cosT) / (etaI * cosI + etaT * cosT);
    return 0.5f * (rPar * rPar + rPerp * rPerp);
}

/* TVector::canRefract (tVector.ipp:467-503) */
static int can_refract(v3 v, v3 n, float iorOuter, float iorInner, v3* out) {
    float etaI = iorOuter, etaT = iorInner, signFactor = 1.0f;
    const float cosI = v3_dot(v, n);
    if (cosI < 0.0f) {
        const float t = etaI; etaI = etaT; etaT = t;
        signFactor = -1.0f;
    }
    const float etaRatio = etaI / etaT;
    const float sinT2 = etaRatio * etaRatio * (1.0f - cosI * cosI);
    if (sinT2 >= 1.0f) return 0;
    const float cosT = sqrtf(1.0f - sinT2);
    v3 dir = v3_scale(v, -etaRatio);
    dir = v3_add(dir, v3_scale(n, (etaRatio * signFactor * cosI - cosT) * signFactor));
    *out = v3_normalize(dir);
    return 1;
}

/* RgbSpectrum(r).average() (tVector.ipp:301-313) */
static float avg3(float r) { return ((r + r) + r) / 3.0f; }

/* PhongBsdf (phongBsdf.cpp:14-20): _brdfFactor, _pdfFactor */
static float phong_brdf_factor(const or_material* m) { return (m->exponent + 2.0f) * (float)0.1591549430918953; }
static float phong_pdf_factor(const or_material* m) { return (m->exponent + 1.0f) * (float)0.1591549430918953; }

/* LambertianDiffuse::evaluate (lambertianDiffuse.cpp:27-44) */
static v3 lambert_eval(const or_material* m, const or_si* si) {
    if (v3_dot(si->wi, si->Ns) * v3_dot(si->wo, si->Ns) <= 0.0f) return v3_make(0, 0, 0);
    return v3_scale(tex_eval(&m->albedo, si->uvw), OR_INV_PI);
}
/* LambertianDiffuse::evaluatePdfW (lambertianDiffuse.cpp:83-96) */
static float lambert_pdf(const or_si* si) {
    if (v3_dot(si->wi, si->Ns) * v3_dot(si->wo, si->Ns) <= 0.0f) return 0.0f;
    return v3_absdot(si->wo, si->Ns) * OR_INV_PI;
}
/* PhongBsdf::evaluate / evaluatePdfW (phongBsdf.cpp:22-39, 83-100) */
static float phong_eval(const or_material* m, const or_si* si) {
    const v3 Ns = si->Ns, V = si->wi, L = si->wo;
    const v3 R = v3_reflect(L, Ns);
    if (v3_dot(V, Ns) * v3_dot(L, Ns) <= 0.0f) return 0.0f;
    const float RdotV = m_max(v3_dot(R, V), 0.0f);
    return powf(RdotV, m->exponent) * phong_brdf_factor(m);
}
static float phong_pdf(const or_material* m, const or_si* si) {
    const v3 Ns = si->Ns, V = si->wi, L = si->wo;
    const v3 R = v3_reflect(L, Ns);
    if (v3_dot(V, Ns) * v3_dot(L, Ns) <= 0.0f) return 0.0f;
    const float RdotV = m_max(v3_dot(R, V), 0.0f);
    return powf(RdotV, m->exponent) * phong_pdf_factor(m);
}

/* ConductorMicrofacet::evaluate (conductorMicrofacet.cpp:32-67) */
static v3 ggx_eval(const or_material* m, const or_si* si) {
    const v3 Ns = si->Ns, V = si->wi, L = si->wo;
    const float VdotN = v3_dot(V, Ns), LdotN = v3_dot(L, Ns);
    if (VdotN * LdotN <= 0.0f) return v3_make(0, 0, 0);
    v3 H = v3_scale(v3_add(V, L), m_sign(VdotN));
    if (v3_is_zero(H)) return v3_make(0, 0, 0);
    H = v3_normalize(H);
    const float a = mat_alpha(m, si->uvw);
    const float LdotH = v3_dot(L, H);
    const v3 F = schlick(v3_load(m->f0), LdotH);
    const float D = ggx_D(a, Ns, H);
    const float G = ggx_G(a, V, L, Ns, H);
    return v3_divs(v3_scale(v3_scale(F, G), D), 4.0f * fabsf(VdotN * LdotN));
}
/* ConductorMicrofacet::evaluatePdfW (conductorMicrofacet.cpp:123-156) */
static float ggx_pdf(const or_material* m, const or_si* si) {
    const v3 Ns = si->Ns, V = si->wi, L = si->wo;
    const float VdotN = v3_dot(V, Ns), LdotN = v3_dot(L, Ns);
    if (VdotN * LdotN <= 0.0f) return 0.0f;
    v3 H = v3_scale(v3_add(V, L), m_sign(VdotN));
    if (v3_is_zero(H)) return 0.0f;
    H = v3_normalize(H);
    const float a = mat_alpha(m, si->uvw);
    const float D = ggx_D(a, Ns, H);
    const float NdotH = v3_dot(Ns, H), LdotH = v3_dot(L, H);
    const float pdfH = fabsf(D * NdotH);
    const float pdfL = fabsf(pdfH / (4.0f * LdotH));
    if (!isfinite(pdfL)) return 0.0f;
    return pdfL;
}

/* Bsdf::evaluate of each material; MixedBsdf::evaluate (mixedBsdf.cpp:35-63) for
 * plastic (A = Lambertian, B = Phong, ratio = diffuse-ratio); specular BSDFs and
 * the absorber evaluate to zero (specularReflection.cpp:25-30, specularDielectric.cpp:28-33,
 * specularTransmission.cpp:28-33, absorberBsdf.cpp:12-17) */
static v3 bsdf_eval(const or_material* m, const or_si* si) {
    switch (m->type) {
    case OR_MAT_LAMBERT: return lambert_eval(m, si);
    case OR_MAT_GGX: return ggx_eval(m, si);
    case OR_MAT_PHONG: { const float f = phong_eval(m, si); return v3_make(f, f, f); }
    case OR_MAT_PLASTIC: {
        const float r = m->ratio, rc = 1.0f - r;
        const v3 fa = lambert_eval(m, si);
        const float fb = phong_eval(m, si);
        return v3_add(v3_scale(fa, r), v3_make(fb * rc, fb * rc, fb * rc));
    }
    default: return v3_make(0, 0, 0);
    }
}

/* Bsdf::evaluatePdfW of each material; MixedBsdf::evaluatePdfW (mixedBsdf.cpp:157-187) */
static float bsdf_pdf(const or_material* m, const or_si* si) {
    switch (m->type) {
    case OR_MAT_LAMBERT: return lambert_pdf(si);
    case OR_MAT_GGX: return ggx_pdf(m, si);
    case OR_MAT_PHONG: return phong_pdf(m, si);
    case OR_MAT_PLASTIC: {
        const float avg = avg3(m->ratio);
        return lambert_pdf(si) * avg + phong_pdf(m, si) * (1.0f - avg);
    }
    default: return 0.0f;
    }
}

/* Materials whose lobes include a specular one (BsdfLobes::hasAtLeastOne in
 * ptEstimator.cpp:68-70): no NEE there, emission counted at the next hit. */
static int mat_specular(const or_material* m) {
    return m->type == OR_MAT_MIRROR || m->type == OR_MAT_GLASS || m->type == OR_MAT_TRANSMISSION;
}

/* Uniforms an evaluateSample draws, in the reference's order: Vector2R(next, next)
 * lands its first draw in y (g++ evaluates constructor arguments right to left);
 * MixedBsdf draws its lobe choice first (mixedBsdf.cpp:81), SpecularDielectric
 * its choice only (specularDielectric.cpp:49). */
static int bsdf_draws(const or_material* m) {
    switch (m->type) {
    case OR_MAT_LAMBERT: case OR_MAT_GGX: case OR_MAT_PHONG: return 2;
    case OR_MAT_PLASTIC: return 3;
    case OR_MAT_GLASS: return 1;
    default: return 0;
    }
}

/* LambertianDiffuse::evaluateSample (lambertianDiffuse.cpp:46-81) with
 * Hemisphere::cosineWeightedSampling (hemisphere.cpp:30-48); sample = (x, y) */
static or_bsdf_sample lambert_sample(const or_material* m, const or_si* si, float x0, float y0) {
    or_bsdf_sample out;
    memset(&out, 0, sizeof(out));
    const v3 Ns = si->Ns, V = si->wi;
    v3 z, x;
    frame(Ns, &z, &x);
    const float phi = OR_TWO_PI * x0;
    const float theta = acosf(1.0f - 2.0f * y0) * 0.5f;
    const float cosT = cosf(theta);
    const float sinT = sqrtf(1.0f - cosT * cosT);
    const v3 l = v3_make(sinf(phi) * sinT, cosT, cosf(phi) * sinT);
    const float pdf = cosT * OR_INV_PI;
    v3 L = v3_normalize(to_world(x, Ns, z, l));
    if (v3_dot(V, Ns) <= 0.0f) L = v3_neg(L);
    out.f = v3_scale(tex_eval(&m->albedo, si->uvw), OR_INV_PI);
    out.L = L;
    out.pdf = pdf;
    out.valid = !v3_is_zero(out.f) && out.pdf > 0.0f;
    return out;
}

/* PhongBsdf::evaluateSample (phongBsdf.cpp:41-81) with
 * Hemisphere::cosineExpWeightedSampling (hemisphere.cpp:50-68) */
static or_bsdf_sample phong_sample(const or_material* m, const or_si* si, float x0, float y0) {
    or_bsdf_sample out;
    memset(&out, 0, sizeof(out));
    const v3 Ns = si->Ns, V = si->wi;
    v3 z, x;
    frame(Ns, &z, &x);
    const float phi = OR_TWO_PI * x0;
    const float cosT = powf(y0, 1.0f / (m->exponent + 1.0f));
    const float sinT = sqrtf(1.0f - cosT * cosT);
    const v3 l = v3_make(sinf(phi) * sinT, cosT, cosf(phi) * sinT);
    const v3 L = v3_normalize(to_world(x, Ns, z, l));
    if (v3_dot(V, Ns) * v3_dot(L, Ns) <= 0.0f) return out;
    const v3 R = v3_reflect(L, Ns);
    const float RdotV = m_max(v3_dot(R, V), 0.0f);
    const float powerTerm = powf(RdotV, m->exponent);
    const float f = powerTerm * phong_brdf_factor(m);
    out.f = v3_make(f, f, f);
    out.L = L;
    out.pdf = powerTerm * phong_pdf_factor(m);
    out.valid = !v3_is_zero(out.f) && out.pdf > 0.0f;
    return out;
}

/* ConductorMicrofacet::evaluateSample (conductorMicrofacet.cpp:69-121) with
 * IsotropicGgx::sampleHalfVectorH (isotropicGgx.cpp:71-100) */
static or_bsdf_sample ggx_sample(const or_material* m, const or_si* si, float u0, float u1) {
    or_bsdf_sample out;
    memset(&out, 0, sizeof(out));
    const v3 Ns = si->Ns, V = si->wi;
    v3 z, x;
    frame(Ns, &z, &x);
    const float Nf = (v3_dot(V, Ns) > 0.0f) ? 1.0f : -1.0f;
    const float a = mat_alpha(m, si->uvw);
    const float s0 = m_clamp(u0, 0.0f, 0.9999f), s1 = m_clamp(u1, 0.0f, 0.9999f);
    const float phi = OR_TWO_PI * s0;
    const float theta = atanf(a * sqrtf(s1 / (1.0f - s1)));
    v3 h;
    if (!isfinite(theta)) h = v3_make(0.0f, 1.0f, 0.0f);
    else {
        const float cosT = cosf(theta);
        const float sinT = sqrtf(1.0f - cosT * cosT);
        h = v3_make(sinf(phi) * sinT, cosT, cosf(phi) * sinT);
    }
    const v3 H = v3_normalize(to_world(x, Ns, z, h));
    const v3 L = v3_reflect(V, v3_scale(H, Nf));
    const float VdotN = v3_dot(V, Ns), LdotN = v3_dot(L, Ns), LdotH = v3_dot(L, H), NdotH = v3_dot(Ns, H);
    if (VdotN * LdotN <= 0.0f) return out;
    const v3 F = schlick(v3_load(m->f0), LdotH);
    const float D = ggx_D(a, Ns, H);
    const float G = ggx_G(a, V, L, Ns, H);
    const float pdfH = fabsf(D * NdotH);
    const float pdfL = fabsf(pdfH / (4.0f * LdotH));
    if (!isfinite(pdfL)) return out;
    out.f = v3_divs(v3_scale(v3_scale(F, G), D), 4.0f * fabsf(VdotN * LdotN));
    out.L = L;
    out.pdf = pdfL;
    out.valid = !v3_is_zero(out.f) && out.pdf > 0.0f;
    return out;
}

/* Refraction of SpecularTransmission::evaluateSample (specularTransmission.cpp:35-78)
 * and SpecularDielectric's refraction branch (specularDielectric.cpp:83-112),
 * radiance transport; pdf set by the caller. */
static int transmit(const or_material* m, const or_si* si, or_bsdf_sample* out) {
    float etaI = m->ior_outer, etaT = m->ior_inner;
    v3 L;
    if (!can_refract(si->wi, si->Ns, etaI, etaT, &L)) return 0;
    const float cosThetaI = v3_dot(L, si->Ns);
    const float R = dielectric_fresnel(m->ior_outer, m->ior_inner, cosThetaI);
    if (cosThetaI < 0.0f) { const float t = etaI; etaI = etaT; etaT = t; }
    const float btdf = (etaT * etaT) / (etaI * etaI);
    const float LdotN = fabsf(cosThetaI);
    const float T = 1.0f - R;
    out->f = v3_divs(v3_scale(v3_scale(tex_eval(&m->albedo, si->uvw), T), btdf), LdotN);
    out->L = L;
    return 1;
}

/* Bsdf::evaluateSample of each material from its uniforms u[] in draw order
 * (bsdf_draws): LambertianDiffuse, ConductorMicrofacet, PhongBsdf as above;
 * MixedBsdf (mixedBsdf.cpp:65-155: choosing B weighs B by the ratio and A by its
 * complement, as written there); SpecularReflection (specularReflection.cpp:32-57);
 * SpecularDielectric (specularDielectric.cpp:35-123); SpecularTransmission;
 * AbsorberBsdf (never valid). */
static or_bsdf_sample bsdf_sample_u(const or_material* m, const or_si* si, const float* u) {
    or_bsdf_sample out;
    memset(&out, 0, sizeof(out));
    const v3 Ns = si->Ns, V = si->wi;
    switch (m->type) {
    case OR_MAT_LAMBERT: return lambert_sample(m, si, u[1], u[0]);
    case OR_MAT_GGX: return ggx_sample(m, si, u[1], u[0]);
    case OR_MAT_PHONG: return phong_sample(m, si, u[1], u[0]);
    case OR_MAT_PLASTIC: {
        const float r = m->ratio, rc = 1.0f - r;
        const float avg = avg3(r);
        or_si local = *si;
        if (u[0] < avg) {
            const or_bsdf_sample a = lambert_sample(m, si, u[2], u[1]);
            if (!a.valid) return out;
            local.wo = a.L;
            const float fb = phong_eval(m, &local);
            out.f = v3_add(v3_scale(a.f, r), v3_make(rc * fb, rc * fb, rc * fb));
            out.pdf = avg * a.pdf + (1.0f - avg) * phong_pdf(m, &local);
            out.L = a.L;
        } else {
            const or_bsdf_sample b = phong_sample(m, si, u[2], u[1]);
            if (!b.valid) return out;
            local.wo = b.L;
            const v3 fa = lambert_eval(m, &local);
            out.f = v3_add(v3_scale(b.f, r), v3_scale(fa, rc));
            out.pdf = avg * b.pdf + (1.0f - avg) * lambert_pdf(&local);
            out.L = b.L;
        }
        break;
    }
    case OR_MAT_MIRROR: {
        const float VdotN = v3_dot(V, Ns);
        const float NFactor = (VdotN > 0.0f) ? 1.0f : -1.0f;
        const v3 L = v3_reflect(V, v3_scale(Ns, NFactor));
        const float LdotN = v3_dot(L, Ns);
        v3 F;
        if (m->fresnel == 0) F = schlick(v3_load(m->f0), LdotN);
        else { const float R = dielectric_fresnel(m->ior_outer, m->ior_inner, LdotN); F = v3_make(R, R, R); }
        out.f = v3_divs(v3_mul(F, tex_eval(&m->albedo, si->uvw)), fabsf(LdotN));
        out.L = L;
        out.pdf = 1.0f;
        break;
    }
    case OR_MAT_GLASS: {
        const float VdotN = v3_dot(V, Ns);
        const float R = dielectric_fresnel(m->ior_outer, m->ior_inner, VdotN);
        const float prob = avg3(R);
        if (u[0] < prob) {
            const float NFactor = (VdotN > 0.0f) ? 1.0f : -1.0f;
            const v3 L = v3_reflect(V, v3_scale(Ns, NFactor));
            const float LdotN = v3_absdot(L, Ns);
            out.f = v3_divs(v3_scale(tex_eval(&m->albedo, si->uvw), R), LdotN);
            out.L = L;
            out.pdf = prob;
        } else {
            if (!transmit(m, si, &out)) return out;
            out.pdf = 1.0f - prob;
        }
        break;
    }
    case OR_MAT_TRANSMISSION:
        if (!transmit(m, si, &out)) return out;
        out.pdf = 1.0f;
        break;
    default:
        return out;
    }
    /* BsdfSample::isValid (bsdfSample.cpp:39-41) */
    out.valid = !v3_is_zero(out.f) && out.pdf > 0.0f;
    return out;
}

/* The uniforms of one evaluateSample: the sequential stream in draw order, or the
 * Philox slots (x, y, lobe choice) of the GPU layout. */
static or_bsdf_sample bsdf_sample_rng(const or_material* m, const or_si* si, or_rng* rng, int bounce, int slot_x,
                                      int slot_y, int slot_c) {
    float u[3] = {0.0f, 0.0f, 0.0f};
    const int n = bsdf_draws(m);
    if (rng->mode == OR_RNG_SEQ) {
        for (int i = 0; i < n; ++i) u[i] = rng_u(rng, bounce, 0);
    } else if (n == 1) {
        u[0] = rng_u(rng, bounce, slot_c);
    } else if (n == 2) {
        u[0] = rng_u(rng, bounce, slot_y);
        u[1] = rng_u(rng, bounce, slot_x);
    } else if (n == 3) {
        u[0] = rng_u(rng, bounce, slot_c);
        u[1] = rng_u(rng, bounce, slot_y);
        u[2] = rng_u(rng, bounce, slot_x);
    }
    return bsdf_sample_u(m, si, u);
}

void or_bsdf_query_n(const or_scene* s, int mi, const float* in, int64_t n, int nu, float* out) {
    const or_material* m = &s->mats[mi];
    for (int64_t i = 0; i < n; ++i) {
        const float* q = in + (12 + nu) * i;
        float* o = out + 12 * i;
        or_si si;
        memset(&si, 0, sizeof(si));
        si.Ns = si.Ng = v3_load(q);
        si.wi = v3_load(q + 3);
        si.wo = v3_load(q + 6);
        si.uvw = v3_load(q + 9);
        v3_store(o, bsdf_eval(m, &si));
        o[3] = bsdf_pdf(m, &si);
        float u[3] = {0.0f, 0.0f, 0.0f};
        for (int k = 0; k < nu && k < 3; ++k) u[k] = q[12 + k];
        const or_bsdf_sample b = bsdf_sample_u(m, &si, u);
        v3_store(o + 4, b.f);
        v3_store(o + 7, b.L);
        o[10] = b.pdf;
        o[11] = (float)b.valid;
    }
}

void or_bsdf_query(const or_scene* s, int mi, const float* in, int64_t n, float* out) {
    or_bsdf_query_n(s, mi, in, n, 2, out);
}

/* ---- area lights ---------------------------------------------------------------- */

/* AreaLight::_canEmit (areaLight.cpp:24-28) */
static int can_emit(const or_light_state* L, v3 dir, v3 N) { return (v3_dot(dir, N) >= 0.0f) != (L->back_face != 0); }

typedef struct { v3 pos, Ns, radiance; float pdfW; int valid; } or_light_smp;
static or_light_smp env_sample(const or_light_state* L, float sx, float sy);

/* SingleAreaLight::evaluateDirectSample (singleAreaLight.cpp:54-85) with
 * Triangle::evaluatePositionSample (triangle.cpp:162-193) + SampleWarp::uniformTriangleUv
 * (sampleWarp.cpp:10-19), or Rectangle::evaluatePositionSample (rectangle.cpp:134-169).
 * The rectangle's rejection loop draws (s, t) pairs until t <= short/long; in the
 * Philox layout the accepted t is drawn directly as u * short/long (the same
 * distribution, uniform on [0, short/long)). */
static or_light_smp light_sample(const or_scene* s, int li, v3 target, or_rng* rng, int bounce) {
    or_light_smp out;
    memset(&out, 0, sizeof(out));
    const or_light_state* L = &s->lights[li];
    if (L->type == OR_LIGHT_ENVIRONMENT) {
        float sx, sy;
        rng_u2(rng, bounce, SL_LPOS0, SL_LPOS1, &sx, &sy);
        return env_sample(L, sx, sy);
    }
    const or_prim* p = &s->prims[L->prim];
    v3 P, Ns;
    float pdfA;
    if (p->type == OR_PRIM_RECTANGLE) {
        v3 eA = p->e1, eB = p->e2;
        if (v3_len(eA) < v3_len(eB)) { const v3 t = eA; eA = eB; eB = t; }
        const float lw = v3_len(eA), sw = v3_len(eB);
        float su, tu;
        if (rng->mode == OR_RNG_SEQ) {
            do {
                su = rng_u(rng, bounce, SL_LPOS0);
                tu = rng_u(rng, bounce, SL_LPOS1);
            } while (tu > sw / lw);
        } else {
            su = rng_u(rng, bounce, SL_LPOS0);
            tu = rng_u(rng, bounce, SL_LPOS1) * (sw / lw);
        }
        P = v3_add(v3_add(p->vB, v3_scale(eA, su)), v3_scale(v3_normalize(eB), tu * v3_len(eA)));
        v3 N = v3_cross(p->e1, p->e2);
        N = v3_is_zero(N) ? v3_make(0.0f, 1.0f, 0.0f) : v3_normalize(N);
        Ns = N;
        pdfA = 1.0f / prim_area(p);
    } else {
        float u0, u1;
        rng_u2(rng, bounce, SL_LPOS0, SL_LPOS1, &u0, &u1);
        const float sqrtA = sqrtf(u0);
        const float st0 = 1.0f - sqrtA, st1 = sqrtA * u1;
        P = v3_add(v3_add(p->vA, v3_scale(p->e1, st0)), v3_scale(p->e2, st1));
        const v3 b = barycentric(p, P);
        v3 Ng = v3_cross(p->e1, p->e2);
        Ng = v3_is_zero(Ng) ? v3_make(0.0f, 1.0f, 0.0f) : v3_normalize(Ng);
        v3 n = bary_mix(b, p->nA, p->nB, p->nC);
        Ns = v3_is_zero(n) ? Ng : n; /* not renormalized (triangle.cpp:182-183) */
        pdfA = 1.0f / prim_area(p);
    }
    if (!(pdfA > 0.0f)) return out;
    const v3 emitVec = v3_sub(target, P);
    if (!can_emit(L, emitVec, Ns)) return out;
    const float d2 = v3_len2(emitVec);
    const float cosL = v3_absdot(v3_divs(emitVec, sqrtf(d2)), Ns);
    if (cosL <= 0.0f) return out;
    out.pos = P;
    out.Ns = Ns;
    out.radiance = L->radiance;
    out.pdfW = pdfA * d2 / cosL;
    /* DirectLightSample::isValid (directLightSample.cpp:57-59) */
    out.valid = !v3_is_zero(out.radiance) && out.pdfW > 0.0f;
    return out;
}

/* SingleAreaLight::emittance (singleAreaLight.cpp:39-52) */
static v3 light_emittance(const or_scene* s, int li, const or_si* si) {
    const or_light_state* L = &s->lights[li];
    if (!can_emit(L, si->wi, si->Ns)) return v3_make(0, 0, 0);
    return L->radiance;
}
/* SingleAreaLight::evaluateDirectPdfW (singleAreaLight.cpp:87-109) */
static float light_pdfW(const or_scene* s, int li, const or_si* si, v3 target) {
    const or_light_state* L = &s->lights[li];
    const v3 emitVec = v3_sub(target, si->P);
    if (!can_emit(L, si->wi, si->Ns)) return 0.0f;
    const float c = v3_absdot(si->wi, si->Ns);
    if (c <= 0.0f) return 0.0f;
    const float pdfA = 1.0f / prim_area(&s->prims[L->prim]);
    return pdfA * v3_len2(emitVec) / c;
}

/* SphericalMapper::mappingToUvw (sphericalMapper.cpp:14-29), y up */
static v3 spherical_uvw(v3 d) {
    const v3 u = v3_normalize(d);
    const float theta = acosf(m_clamp(u.y, -1.0f, 1.0f));
    float phi = atan2f(u.x, u.z);
    if (phi < 0.0f) phi += OR_TWO_PI;
    return v3_make(phi * (float)0.1591549430918953, (OR_PI - theta) * OR_INV_PI, 0.0f);
}

/* Scene::isIntersecting's background branch (scene.cpp:47-61) with
 * InfiniteSphere::isIntersecting / evaluateSurfaceDetail (infiniteSphere.cpp:36-59):
 * only unbounded rays reach it; the surface position is the ray direction. */
static int background_hit(const or_scene* s, v3 d, float maxT, or_si* si) {
    if (s->env_light < 0 || maxT < FLT_MAX) return 0;
    memset(si, 0, sizeof(*si));
    si->prim = -1;
    si->wi = v3_neg(d);
    si->P = d;
    si->uvw = spherical_uvw(d);
    return 1;
}

/* EnvironmentLight::emittance (environmentLight.cpp:59-65) */
static v3 env_emittance(const or_light_state* L, const or_si* si) { return tex_eval(&L->env, si->uvw); }

/* Distribution2D::pdfContinuous (distribution2D.cpp:46-53) */
static float env_pdf_uv(const or_light_state* L, float u, float v) {
    const int64_t iy = dist1_index(&L->marginal, v);
    const float pdfY = dist1_pdf(&L->marginal, iy);
    const or_dist1* c = &L->conditional[iy];
    const float pdfX = dist1_pdf(c, dist1_index(c, u));
    return pdfX * pdfY;
}

/* EnvironmentLight::evaluateDirectPdfW (environmentLight.cpp:91-105) */
static float env_pdfW(const or_light_state* L, const or_si* si) {
    const float sinTheta = sinf((1.0f - si->uvw.y) * OR_PI);
    if (sinTheta <= 0.0f) return 0.0f;
    return env_pdf_uv(L, si->uvw.x, si->uvw.y) / (2.0f * OR_PI * OR_PI * sinTheta);
}

/* EnvironmentLight::evaluateDirectSample (environmentLight.cpp:67-89) with seed (sx, sy),
 * Distribution2D::sampleContinuous (distribution2D.cpp:30-44) and
 * InfiniteSphere::uvwToPosition (infiniteSphere.cpp:61-76, z up, radius 1e6). */
static or_light_smp env_sample(const or_light_state* L, float sx, float sy) {
    or_light_smp out;
    memset(&out, 0, sizeof(out));
    float pdfY, pdfX;
    int64_t iy, ix;
    const float v = dist1_sample(&L->marginal, sy, &pdfY, &iy);
    const float u = dist1_sample(&L->conditional[iy], sx, &pdfX, &ix);
    const float uvPdf = pdfX * pdfY;
    const float theta = (1.0f - v) * OR_PI;
    const float phi = u * OR_TWO_PI;
    const float cosTheta = cosf(theta);
    const float sinT = sqrtf(1.0f - cosTheta * cosTheta);
    const v3 dir = v3_make(cosf(phi) * sinT, sinf(phi) * sinT, cosTheta);
    const v3 pos = v3_scale(v3_normalize(dir), 1000000.0f);
    const float sinTheta = sinf((1.0f - v) * OR_PI);
    if (sinTheta <= 0.0f) return out;
    out.pos = pos;
    out.radiance = tex_eval(&L->env, v3_make(u, v, 0.0f));
    out.pdfW = uvPdf / (2.0f * OR_PI * OR_PI * sinTheta);
    /* DirectLightSample::isValid (directLightSample.cpp:57-59) */
    out.valid = !v3_is_zero(out.radiance) && out.pdfW > 0.0f;
    return out;
}

void or_light_radiance(const or_scene* s, float* out) {
    for (int32_t i = 0; i < s->nlights; ++i) v3_store(out + 3 * i, s->lights[i].radiance);
}

int or_env_info(const or_scene* s, int li, float* out, float* cdf) {
    if (li < 0 || li >= s->nlights || s->lights[li].type != OR_LIGHT_ENVIRONMENT) return -1;
    const or_light_state* L = &s->lights[li];
    out[0] = L->flux;
    out[1] = L->background_flux;
    if (cdf) {
        const int64_t h = L->marginal.n, w = L->env.width;
        memcpy(cdf, L->marginal.cdf, sizeof(float) * (size_t)(h + 1));
        for (int64_t r = 0; r < h; ++r)
            memcpy(cdf + (h + 1) + r * (w + 1), L->conditional[r].cdf, sizeof(float) * (size_t)(w + 1));
    }
    return 0;
}

void or_power_cdf(const or_scene* s, float* out) {
    memcpy(out, s->power_cdf, sizeof(float) * (size_t)(s->nlights + 1));
}

void or_light_sample(const or_scene* s, int li, const float* in, int64_t n, int nu, float* out) {
    /* Explicit uniforms in the reference's draw order: a triangle draws
     * (y, x) (right-to-left argument evaluation), a rectangle draws (s, t)
     * pairs until t <= short/long. */
    const or_light_state* L = &s->lights[li];
    for (int64_t i = 0; i < n; ++i) {
        const float* q = in + (3 + nu) * i;
        const float* vals = q + 3;
        float* o = out + 11 * i;
        or_light_smp r;
        memset(&r, 0, sizeof(r));
        if (L->type == OR_LIGHT_ENVIRONMENT) {
            /* Vector2R(nextReal(), nextReal()): the first draw is the seed's y */
            r = env_sample(L, vals[1], vals[0]);
            v3_store(o, r.pos);
            v3_store(o + 3, r.Ns);
            v3_store(o + 6, r.radiance);
            o[9] = r.pdfW;
            o[10] = (float)r.valid;
            continue;
        }
        const or_prim* p = &s->prims[L->prim];
        v3 P, Ns;
        if (p->type == OR_PRIM_RECTANGLE) {
            v3 eA = p->e1, eB = p->e2;
            if (v3_len(eA) < v3_len(eB)) { const v3 t = eA; eA = eB; eB = t; }
            const float lw = v3_len(eA), sw = v3_len(eB);
            float su, tu;
            int k = 0;
            do {
                su = vals[k];
                tu = vals[k + 1];
                k += 2;
            } while (tu > sw / lw && k + 1 < nu);
            P = v3_add(v3_add(p->vB, v3_scale(eA, su)), v3_scale(v3_normalize(eB), tu * v3_len(eA)));
            v3 N = v3_cross(p->e1, p->e2);
            Ns = v3_is_zero(N) ? v3_make(0.0f, 1.0f, 0.0f) : v3_normalize(N);
        } else {
            const float u1 = vals[0], u0 = vals[1];
            const float sqrtA = sqrtf(u0);
            P = v3_add(v3_add(p->vA, v3_scale(p->e1, 1.0f - sqrtA)), v3_scale(p->e2, sqrtA * u1));
            const v3 b = barycentric(p, P);
            v3 Ng = v3_cross(p->e1, p->e2);
            Ng = v3_is_zero(Ng) ? v3_make(0.0f, 1.0f, 0.0f) : v3_normalize(Ng);
            const v3 nn = bary_mix(b, p->nA, p->nB, p->nC);
            Ns = v3_is_zero(nn) ? Ng : nn;
        }
        const float pdfA = 1.0f / prim_area(p);
        const v3 emitVec = v3_sub(v3_load(q), P);
        if (pdfA > 0.0f && can_emit(L, emitVec, Ns)) {
            const float d2 = v3_len2(emitVec);
            const float cosL = v3_absdot(v3_divs(emitVec, sqrtf(d2)), Ns);
            if (cosL > 0.0f) {
                r.pos = P; r.Ns = Ns; r.radiance = L->radiance; r.pdfW = pdfA * d2 / cosL;
                r.valid = !v3_is_zero(r.radiance) && r.pdfW > 0.0f;
            }
        }
        v3_store(o, r.pos);
        v3_store(o + 3, r.Ns);
        v3_store(o + 6, r.radiance);
        o[9] = r.pdfW;
        o[10] = (float)r.valid;
    }
}

void or_light_hit(const or_scene* s, const float* rays, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        float* o = out + 5 * i;
        o[0] = -1.0f; o[1] = o[2] = o[3] = o[4] = 0.0f;
        or_si si;
        if (!scene_intersect(s, v3_load(r), v3_load(r + 4), r[3], r[7], &si, NULL)) {
            if (!background_hit(s, v3_load(r + 4), r[7], &si)) continue;
            const or_light_state* L = &s->lights[s->env_light];
            o[0] = (float)s->env_light;
            v3_store(o + 1, env_emittance(L, &si));
            o[4] = env_pdfW(L, &si);
            continue;
        }
        const int li = s->prims[si.prim].light;
        if (li < 0) continue;
        o[0] = (float)li;
        v3_store(o + 1, light_emittance(s, li, &si));
        o[4] = light_pdfW(s, li, &si, v3_load(r));
    }
}

/* ---- camera ------------------------------------------------------------------------ */

/* PerspectivePinholeCamera::spawnPrimaryRay (perspectivePinholeCamera.cpp:52-73),
 * TMatrix4::transformPoint / transformVector (tMatrix4.ipp:180-195), Ray::setDirection (ray.cpp:63-67) */
static v3 camera_dir(const or_scene* s, float fx, float fy) {
    const float* m = s->film_to_camera;
    const float x = fx, y = fy, z = 0.0f;
    const v3 pc = v3_make(m[0] * x + m[1] * y + m[2] * z + m[3], m[4] * x + m[5] * y + m[6] * z + m[7],
                          m[8] * x + m[9] * y + m[10] * z + m[11]);
    const float* c = s->camera_to_world;
    const v3 d = v3_make(c[0] * pc.x + c[1] * pc.y + c[2] * pc.z, c[4] * pc.x + c[5] * pc.y + c[6] * pc.z,
                         c[8] * pc.x + c[9] * pc.y + c[10] * pc.z);
    return v3_normalize(d);
}

void or_camera_matrices(const or_scene* s, float* f2c, float* c2w) {
    memcpy(f2c, s->film_to_camera, sizeof(s->film_to_camera));
    memcpy(c2w, s->camera_to_world, sizeof(s->camera_to_world));
}

void or_camera_rays(const or_scene* s, const float* film_xy, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        v3_store(out + 6 * i, s->cam_pos);
        v3_store(out + 6 * i + 3, camera_dir(s, film_xy[2 * i], film_xy[2 * i + 1]));
    }
}

void or_normalize(const float* v, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) v3_store(out + 3 * i, v3_normalize(v3_load(v + 3 * i)));
}

/* ---- estimator --------------------------------------------------------------------- */

/* TMis<POWER>::weight (tMis.h:21-47) */
static float mis_power(float a, float b) {
    const float a2 = a * a, b2 = b * b;
    return a2 / (a2 + b2);
}

/* UniformLightCluster::sampleOneLight (uniformLightCluster.cpp:15-24) with Random::nextIndex
 * (random.cpp:21-26); PowerLightCluster::sampleOneLight (powerLightCluster.cpp:27-34) with
 * Distribution1D::sampleDiscrete / continuousToDiscrete (distribution1D.cpp:89-125) */
static int pick_light(const or_scene* s, or_rng* rng, int bounce, float* pdf) {
    const float u = rng_u(rng, bounce, SL_PICK);
    const size_t n = (size_t)s->nlights;
    if (s->desc.cluster == OR_CLUSTER_POWER) {
        size_t idx = 0;
        for (size_t i = n - 1 + 1; i-- > 0;) {
            if (s->power_cdf[i] <= u) { idx = i; break; }
        }
        *pdf = s->power_cdf[idx + 1] - s->power_cdf[idx];
        return (int)idx;
    }
    size_t idx = (size_t)(0.0f + u * (float)n);
    idx = idx < n - 1 ? idx : n - 1;
    *pdf = 1.0f / (float)n;
    return (int)idx;
}

typedef struct {
    int64_t closest, anyhit;
} or_counts;

/* DirectLightEvaluator::evaluate (directLightEvaluator.cpp:18-109) */
static v3 direct_light(const or_scene* s, const or_si* si0, const or_material* m, int li, or_rng* rng, int bounce,
                       or_counts* cnt) {
    or_si si = *si0;
    v3 Ld = v3_make(0, 0, 0);
    const v3 P = si.P, Ns = si.Ns;
    {
        const or_light_smp ls = light_sample(s, li, P, rng, bounce);
        if (ls.valid) {
            const v3 LV = v3_sub(ls.pos, P);
            const float dist = v3_len(LV);
            const v3 L = v3_divs(LV, dist);
            si.wo = L;
            const v3 sd = v3_normalize(L); /* Ray(P, L) normalizes again (ray.cpp:19-27) */
            cnt->anyhit++;
            if (!or_occluded(s, P, sd, OR_RAY_EPS, dist - OR_RAY_EPS)) {
                const v3 f = bsdf_eval(m, &si);
                const v3 factor = v3_scale(f, v3_absdot(L, Ns));
                const float bpdf = bsdf_pdf(m, &si);
                const float w = mis_power(ls.pdfW, bpdf);
                Ld = v3_add(Ld, v3_divs(v3_mul(v3_scale(ls.radiance, w), factor), ls.pdfW));
            }
        }
    }
    {
        const or_bsdf_sample bs = bsdf_sample_rng(m, &si, rng, bounce, SL_MIS0, SL_MIS1, SL_MISC);
        if (bs.valid) {
            const v3 d = v3_normalize(bs.L);
            or_si hit;
            cnt->closest++;
            if (scene_intersect(s, P, d, OR_RAY_EPS, FLT_MAX, &hit, NULL)) {
                if (s->prims[hit.prim].light == li) {
                    const v3 factor = v3_scale(bs.f, v3_absdot(bs.L, Ns));
                    const v3 rad = light_emittance(s, li, &hit);
                    const float lpdf = light_pdfW(s, li, &hit, P);
                    const float w = mis_power(bs.pdf, lpdf);
                    Ld = v3_add(Ld, v3_divs(v3_mul(v3_scale(rad, w), factor), bs.pdf));
                }
            } else if (background_hit(s, d, FLT_MAX, &hit) && s->env_light == li) {
                /* the background sphere's area light is the environment light */
                const or_light_state* L = &s->lights[li];
                const v3 factor = v3_scale(bs.f, v3_absdot(bs.L, Ns));
                const v3 rad = env_emittance(L, &hit);
                const float lpdf = env_pdfW(L, &hit);
                const float w = mis_power(bs.pdf, lpdf);
                Ld = v3_add(Ld, v3_divs(v3_mul(v3_scale(rad, w), factor), bs.pdf));
            }
        }
    }
    return Ld;
}

/* PtEstimator::estimate (ptEstimator.cpp:25-123) with RussianRoulette::isSurvivedOnNextRound
 * (russianRoulette.cpp:9-28): next-event estimation at non-specular vertices (an
 * absorber picks a light and returns before sampling it, directLightEvaluator.cpp:33-35);
 * emission counted at the first hit and after specular vertices.
 * VptEstimator::estimate (vptEstimator.cpp:24-88): BSDF sampling only, emission at
 * every hit. */
static v3 estimate(const or_scene* s, v3 o, v3 d, or_rng* rng, or_counts* cnt) {
    v3 L = v3_make(0, 0, 0), beta = v3_make(1, 1, 1);
    const int vpt = s->desc.estimator == OR_EST_VPT;
    int countEmit = 1;
    v3 ro = o, rd = d;
    for (int bounce = 0; bounce < s->desc.max_depth; ++bounce) {
        or_si si;
        cnt->closest++;
        if (!scene_intersect(s, ro, rd, OR_RAY_EPS, FLT_MAX, &si, NULL)) {
            /* the background sphere (scene.cpp:47-61): emission when counted; its
             * absorber BSDF is not specular, so pt still picks one light
             * (ptEstimator.cpp:68-82) but DirectLightEvaluator returns zero for an
             * absorber (directLightEvaluator.cpp:33-35), and the absorber's sample is
             * invalid, which ends the path (ptEstimator.cpp:86-90) */
            if (!background_hit(s, rd, FLT_MAX, &si)) break;
            if (countEmit || vpt) L = v3_add(L, v3_mul(beta, env_emittance(&s->lights[s->env_light], &si)));
            if (!vpt) {
                float pickPdf;
                (void)pick_light(s, rng, bounce, &pickPdf);
            }
            break;
        }
        const or_prim* p = &s->prims[si.prim];
        const or_material* m = &s->mats[p->mat];
        if (p->light >= 0 && (countEmit || vpt)) L = v3_add(L, v3_mul(beta, light_emittance(s, p->light, &si)));
        if (!vpt) {
            if (!mat_specular(m)) {
                countEmit = 0;
                float pickPdf;
                const int li = pick_light(s, rng, bounce, &pickPdf);
                if (m->type != OR_MAT_ABSORB) {
                    const v3 Ld = v3_divs(direct_light(s, &si, m, li, rng, bounce, cnt), pickPdf);
                    L = v3_add(L, v3_mul(beta, Ld));
                }
            } else {
                countEmit = 1;
            }
        }
        const or_bsdf_sample bs = bsdf_sample_rng(m, &si, rng, bounce, SL_CONT0, SL_CONT1, SL_CONTC);
        if (!bs.valid) break;
        const float LdotN = v3_absdot(bs.L, si.Ns);
        beta = v3_mul(beta, v3_divs(v3_scale(bs.f, LdotN), bs.pdf));
        if (bounce > 2) {
            const float q = m_clamp(1.0f - v3_max_comp(beta), 0.05f, 1.0f);
            const float u = rng_u(rng, bounce, SL_RR);
            if (u > q) beta = v3_scale(beta, 1.0f / (1.0f - q));
            else break;
        }
        if (v3_is_zero(beta)) break;
        ro = si.P;
        rd = v3_normalize(bs.L);
    }
    return L;
}

static void rng_init(or_rng* r, int mode, uint64_t seed, uint64_t key, uint32_t pixel, uint32_t sample) {
    memset(r, 0, sizeof(*r));
    r->mode = mode;
    r->state = key;
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->pixel = pixel;
    r->sample = sample;
}

void or_estimate(const or_scene* s, int rng_mode, uint64_t seed, const float* film_xy, const uint64_t* keys,
                 int64_t n, float* out_rgb, int64_t* out_draws) {
    for (int64_t i = 0; i < n; ++i) {
        or_rng rng;
        rng_init(&rng, rng_mode, seed, keys[i], (uint32_t)(keys[i] >> 32), (uint32_t)keys[i]);
        or_counts cnt = {0, 0};
        const v3 d = camera_dir(s, film_xy[2 * i], film_xy[2 * i + 1]);
        v3_store(out_rgb + 3 * i, estimate(s, s->cam_pos, d, &rng, &cnt));
        if (out_draws) out_draws[i] = rng.draws;
    }
}

/* ---- film ------------------------------------------------------------------------------ */

/* filter evaluate: BoxFilter (boxFilter.cpp:9-11), GaussianFilter (gaussianFilter.cpp:9-33),
 * ConeFilter (coneFilter.cpp:13-16), MitchellFilter (mitchellFilter.cpp:16-40) */
static float mitchell1d(float b, float c, float d) {
    const float ad = fabsf(d);
    const float den = 1.0f / 6.0f;
    if (ad < 1.0f)
        return ((12.0f - 9.0f * b - 6.0f * c) * ad * ad * ad + (-18.0f + 12.0f * b + 6.0f * c) * ad * ad +
                (6.0f - 2.0f * b)) * den;
    if (ad < 2.0f)
        return ((-b - 6.0f * c) * ad * ad * ad + (6.0f * b + 30.0f * c) * ad * ad + (-12.0f * b - 48.0f * c) * ad +
                (8.0f * b + 24.0f * c)) * den;
    return 0.0f;
}
static float filter_eval(const or_scene_desc* d, float x, float y) {
    const float hx = d->filter_w / 2.0f, hy = d->filter_h / 2.0f;
    switch (d->filter) {
        case OR_FILTER_GAUSSIAN: {
            const float ex = -1.0f / (2.0f * d->sigma_x * d->sigma_x);
            const float ey = -1.0f / (2.0f * d->sigma_y * d->sigma_y);
            const float edge = d->amplitude * expf(ex * hx * hx + ey * hy * hy);
            const float v = d->amplitude * expf(ex * x * x + ey * y * y);
            return m_max(0.0f, v - edge);
        }
        case OR_FILTER_CONE: return m_max(hx - fabsf(x), 0.0f) * m_max(hy - fabsf(y), 0.0f);
        case OR_FILTER_MITCHELL: return mitchell1d(d->mitchell_b, d->mitchell_c, x) * mitchell1d(d->mitchell_b, d->mitchell_c, y);
        default: return 1.0f;
    }
}

/* FilmTile::addSample (filmTile.cpp:21-66): NaN/inf samples dropped, footprint clamped to the tile */
static void film_add(const or_scene_desc* d, float* film, int tx0, int ty0, int tx1, int ty1, float fx, float fy,
                     v3 rgb) {
    if (v3_has_nan(rgb) || v3_has_inf(rgb)) return;
    const float hx = d->filter_w / 2.0f, hy = d->filter_h / 2.0f;
    float mnx = fx - hx, mny = fy - hy, mxx = fx + hx, mxy = fy + hy;
    mnx = std_max(mnx, (float)tx0); mny = std_max(mny, (float)ty0);
    mxx = std_min(mxx, (float)tx1); mxy = std_min(mxy, (float)ty1);
    const int x0 = (int)ceilf(mnx - 0.5f), y0 = (int)ceilf(mny - 0.5f);
    const int x1 = (int)(floorf(mxx - 0.5f) + 1), y1 = (int)(floorf(mxy - 0.5f) + 1);
    for (int iy = y0; iy < y1; ++iy)
        for (int ix = x0; ix < x1; ++ix) {
            const float x = (float)ix - (fx - 0.5f);
            const float y = (float)iy - (fy - 0.5f);
            const float w = filter_eval(d, x, y);
            float* px = film + 4 * ((size_t)iy * (size_t)d->width + (size_t)ix);
            px[0] += rgb.x * w; px[1] += rgb.y * w; px[2] += rgb.z * w; px[3] += w;
        }
}

/* Kensler's hashed permutation of [0, l) (A. Kensler, "Correlated Multi-Jittered
 * Sampling", Pixar Technical Memo 13-01, 2013): the GPU tracer's per-pixel order of
 * strata, standing in for SampleRecord2D::shuffle's Fisher-Yates draw. */
static uint32_t kensler_permute(uint32_t i, uint32_t l, uint32_t p) {
    uint32_t w = l - 1;
    w |= w >> 1; w |= w >> 2; w |= w >> 4; w |= w >> 8; w |= w >> 16;
    do {
        i ^= p; i *= 0xe170893du;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8; i *= 0x0929eb3fu;
        i ^= p >> 23;
        i ^= (i & w) >> 1; i *= 1u | p >> 27;
        i *= 0x6935fa69u;
        i ^= (i & w) >> 11; i *= 0x74dcb303u;
        i ^= (i & w) >> 2; i *= 0x9e501cc3u;
        i ^= (i & w) >> 2; i *= 0xc860a3dfu;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}

/* Film position of sample k of pixel (ix, iy) in the Philox layout. */
static void philox_film_pos(const or_scene* s, uint64_t seed, int ix, int iy, int k, float* fx, float* fy) {
    const uint32_t pixel = (uint32_t)((size_t)iy * (size_t)s->desc.width + (size_t)ix);
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t c1[4] = {pixel, (uint32_t)k, 0xFFFFFFFFu, 0u};
    uint32_t r1[4];
    or_philox(c1, key, r1);
    const float u0 = u32_to_unit(r1[0]), u1 = u32_to_unit(r1[1]);
    float jx = u0, jy = u1;
    if (s->desc.sampler == OR_SAMPLER_STRATIFIED) {
        const uint32_t c2[4] = {pixel, 0u, 0xFFFFFFFEu, 0u};
        uint32_t r2[4];
        or_philox(c2, key, r2);
        const uint32_t base = (uint32_t)s->strata_base;
        const uint32_t j = kensler_permute((uint32_t)k, (uint32_t)s->desc.spp, r2[0]);
        if (j < base * base) {
            const float dx = 1.0f / (float)base;
            jx = ((float)(j % base) + u0) * dx;
            jy = ((float)(j / base) + u1) * dx;
        }
    }
    *fx = (float)ix + jx;
    *fy = (float)iy + jy;
}

typedef struct {
    const or_scene* s;
    int      mode;
    uint64_t seed;
    int      spp_begin, spp_end;
    float*   film;
    int      ntiles;
    volatile int next;
    pthread_mutex_t lock;
} render_job;

/* One tile: EstimatorTileWork::work (estimatorTileWork.cpp:32-61) with the stratified
 * sampler (stratifiedSampler.cpp:38-73, sampleRecord2D.cpp:16-36) or the random one
 * (randomSampler.cpp:27-38). */
static void render_tile(render_job* J, int t) {
    const or_scene* s = J->s;
    const or_scene_desc* d = &s->desc;
    const int ntx = (d->width + d->tile_x - 1) / d->tile_x;
    const int x0 = (t % ntx) * d->tile_x, y0 = (t / ntx) * d->tile_y;
    const int x1 = x0 + d->tile_x < d->width ? x0 + d->tile_x : d->width;
    const int y1 = y0 + d->tile_y < d->height ? y0 + d->tile_y : d->height;
    const int spp = d->spp;
    float* jit = (float*)malloc(sizeof(float) * 2 * (size_t)(spp > 0 ? spp : 1));
    for (int iy = y0; iy < y1; ++iy)
        for (int ix = x0; ix < x1; ++ix) {
            const uint64_t pixel = (uint64_t)iy * (uint64_t)d->width + (uint64_t)ix;
            if (J->mode == OR_RNG_SEQ) {
                uint64_t st = seqrng_key(J->seed, pixel, 1ull << 40);
                if (d->sampler == OR_SAMPLER_STRATIFIED) {
                    const int base = s->strata_base;
                    const float dx = 1.0f / (float)base;
                    int idx = 0;
                    for (int sy = 0; sy < base; ++sy)
                        for (int sx = 0; sx < base; ++sx) {
                            const float jx = seqrng_next(&st), jy = seqrng_next(&st);
                            jit[2 * idx] = ((float)(size_t)sx + jx) * dx;
                            jit[2 * idx + 1] = ((float)(size_t)sy + jy) * dx;
                            idx++;
                        }
                    for (; idx < spp; ++idx) {
                        jit[2 * idx] = seqrng_next(&st);
                        jit[2 * idx + 1] = seqrng_next(&st);
                    }
                    for (int i = 0; i < spp; ++i) {
                        /* Random::nextIndex(i, spp) */
                        const size_t interval = (size_t)(spp - i);
                        size_t r = (size_t)((float)(size_t)i + seqrng_next(&st) * (float)interval);
                        r = r < (size_t)i ? (size_t)i : (r > (size_t)spp - 1 ? (size_t)spp - 1 : r);
                        const float a0 = jit[2 * i], a1 = jit[2 * i + 1];
                        jit[2 * i] = jit[2 * r]; jit[2 * i + 1] = jit[2 * r + 1];
                        jit[2 * r] = a0; jit[2 * r + 1] = a1;
                    }
                } else {
                    for (int i = 0; i < spp; ++i) {
                        jit[2 * i] = seqrng_next(&st);
                        jit[2 * i + 1] = seqrng_next(&st);
                    }
                }
            }
            for (int k = J->spp_begin; k < J->spp_end; ++k) {
                float fx, fy;
                if (J->mode == OR_RNG_SEQ) {
                    fx = (float)ix + jit[2 * k];
                    fy = (float)iy + jit[2 * k + 1];
                } else {
                    philox_film_pos(s, J->seed, ix, iy, k, &fx, &fy);
                }
                or_rng rng;
                rng_init(&rng, J->mode, J->seed, seqrng_key(J->seed, pixel, (uint64_t)k), (uint32_t)pixel,
                         (uint32_t)k);
                or_counts cnt = {0, 0};
                const v3 rgb = estimate(s, s->cam_pos, camera_dir(s, fx, fy), &rng, &cnt);
                film_add(d, J->film, x0, y0, x1, y1, fx, fy, rgb);
            }
        }
    free(jit);
}

static void* render_worker(void* arg) {
    render_job* J = (render_job*)arg;
    while (1) {
        pthread_mutex_lock(&J->lock);
        const int t = J->next++;
        pthread_mutex_unlock(&J->lock);
        if (t >= J->ntiles) break;
        render_tile(J, t);
    }
    return NULL;
}

void or_render_linear(const or_scene* s, int rng_mode, uint64_t seed, int spp_begin, int spp_end, int nthreads,
                      float* film) {
    const or_scene_desc* d = &s->desc;
    render_job J;
    memset(&J, 0, sizeof(J));
    J.s = s; J.mode = rng_mode; J.seed = seed; J.film = film;
    J.spp_begin = spp_begin; J.spp_end = spp_end;
    J.ntiles = ((d->width + d->tile_x - 1) / d->tile_x) * ((d->height + d->tile_y - 1) / d->tile_y);
    pthread_mutex_init(&J.lock, NULL);
    memset(film, 0, sizeof(float) * 4 * (size_t)d->width * (size_t)d->height);
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, render_worker, &J);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    pthread_mutex_destroy(&J.lock);
}

/* ---- CPU baseline timing -------------------------------------------------------------- */

typedef struct {
    const or_scene* s;
    int64_t begin, end;
    int     spp;
    uint64_t seed;
    int64_t rays[2], samples;
} time_job;

static void* time_worker(void* arg) {
    time_job* T = (time_job*)arg;
    const or_scene* s = T->s;
    for (int64_t p = T->begin; p < T->end; ++p) {
        const int ix = (int)(p % s->desc.width), iy = (int)(p / s->desc.width);
        for (int k = 0; k < T->spp; ++k) {
            float fx, fy;
            philox_film_pos(s, T->seed, ix, iy, k, &fx, &fy);
            or_rng rng;
            rng_init(&rng, OR_RNG_PHILOX, T->seed, 0, (uint32_t)p, (uint32_t)k);
            or_counts cnt = {0, 0};
            (void)estimate(s, s->cam_pos, camera_dir(s, fx, fy), &rng, &cnt);
            T->rays[0] += cnt.closest;
            T->rays[1] += cnt.anyhit;
            T->samples += 1;
        }
    }
    return NULL;
}

double or_time_render(const or_scene* s, int64_t begin, int64_t end, int spp, int nthreads, uint64_t seed,
                      int64_t* out_rays, int64_t* out_samples) {
    if (nthreads < 1) nthreads = 1;
    time_job* jobs = (time_job*)calloc((size_t)nthreads, sizeof(time_job));
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    const int64_t chunk = (end - begin + nthreads - 1) / nthreads;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < nthreads; ++i) {
        jobs[i].s = s; jobs[i].spp = spp; jobs[i].seed = seed;
        jobs[i].begin = begin + i * chunk;
        jobs[i].end = jobs[i].begin + chunk < end ? jobs[i].begin + chunk : end;
        pthread_create(&th[i], NULL, time_worker, &jobs[i]);
    }
    out_rays[0] = out_rays[1] = 0;
    *out_samples = 0;
    for (int i = 0; i < nthreads; ++i) {
        pthread_join(th[i], NULL);
        out_rays[0] += jobs[i].rays[0];
        out_rays[1] += jobs[i].rays[1];
        *out_samples += jobs[i].samples;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(jobs); free(th);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* Film::save (film.cpp:159-196) with math::forward_gamma_correction (math.cpp:92-99)
 * and ImageUtils::hdrToLdr (imageUtils.cpp:23-36). */
static float gamma_fwd(float v) {
    if (v <= 0.0031308f) return 12.92f * v;
    return 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
}
void or_film_to_srgb8(const float* film, int64_t npix, uint8_t* out) {
    for (int64_t i = 0; i < npix; ++i) {
        const float* p = film + 4 * i;
        const float invW = (p[3] == 0.0f) ? 0.0f : 1.0f / p[3];
        const float splat = 0.0f;
        for (int c = 0; c < 3; ++c) {
            const float lin = p[c] * invW + splat;
            const float srgb = gamma_fwd(lin);
            out[3 * i + c] = (uint8_t)m_clamp(srgb * 255.0f + 0.5f, 0.0f, 255.0f);
        }
    }
}

/* The byte of one resolved linear value, as or_film_to_srgb8 forms it. */
static uint8_t srgb8_of(float lin) {
    return (uint8_t)m_clamp(gamma_fwd(lin) * 255.0f + 0.5f, 0.0f, 255.0f);
}

/* Exhaustive scan of the byte function over every float bit pattern in [+0, 1]
 * (0x00000000 .. 0x3F800000), split into nthreads contiguous ranges. Each range
 * records its first byte, its last byte, whether the byte ever decreases inside
 * it, and the first pattern at which each byte value is reached. */
typedef struct {
    uint32_t a, b;
    uint32_t first[256];
    int      seen[256];
    int      monotone;
    uint8_t  b0, b1;
} srgb_scan;

static void* srgb_scan_run(void* p) {
    srgb_scan* J = (srgb_scan*)p;
    J->monotone = 1;
    uint8_t prev = 0;
    for (uint32_t u = J->a; u < J->b; ++u) {
        float v;
        memcpy(&v, &u, 4);
        const uint8_t by = srgb8_of(v);
        if (u == J->a) J->b0 = by;
        else if (by < prev) J->monotone = 0;
        if (!J->seen[by]) { J->seen[by] = 1; J->first[by] = u; }
        prev = by;
    }
    J->b1 = prev;
    return NULL;
}

int or_srgb8_thresholds(uint32_t* out, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    const uint64_t lo = 0, hi = 0x3F800001ull;   /* [+0, 1] inclusive */
    srgb_scan* J = (srgb_scan*)calloc((size_t)nthreads, sizeof(srgb_scan));
    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    for (int i = 0; i < nthreads; ++i) {
        J[i].a = (uint32_t)(lo + (hi - lo) * (uint64_t)i / (uint64_t)nthreads);
        J[i].b = (uint32_t)(lo + (hi - lo) * (uint64_t)(i + 1) / (uint64_t)nthreads);
        pthread_create(&th[i], NULL, srgb_scan_run, &J[i]);
    }
    int ok = 1;
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    for (int k = 0; k < 256; ++k) out[k] = 0xFFFFFFFFu;
    for (int i = 0; i < nthreads; ++i) {
        ok &= J[i].monotone;
        if (i > 0 && J[i].b0 < J[i - 1].b1) ok = 0;
        for (int k = 0; k < 256; ++k)
            if (J[i].seen[k] && J[i].first[k] < out[k]) out[k] = J[i].first[k];
    }
    /* a byte value the scan never produced (skipped by a jump of two) starts
     * where the next one does */
    for (int k = 254; k >= 0; --k)
        if (out[k] == 0xFFFFFFFFu) out[k] = out[k + 1];
    out[0] = 0;
    free(J); free(th);
    return ok ? 0 : -1;
}

/* ---- path-replay gradient, FP64 host reference -------------------------------------------
 * Not a reference function: the reference has no derivatives. This is the host check of
 * the GPU adjoint. Each path of the Philox layout is re-run exactly as estimate() runs it
 * (same floats, same control flow); every factor of its contribution is recorded per
 * vertex; the derivative of the sample's radiance is then formed by a backward pass in
 * double precision over the suffix sums S_j = E_j + N_j + c_j * S_{j+1}:
 *     dL/dtheta_j = beta_j * (dN_j/dtheta + dc_j/dtheta * S_{j+1})
 * with sampling pdfs, MIS weights and the Russian-roulette factor held fixed, so only the
 * BSDF value f (albedo of a Lambertian vertex, roughness of a GGX vertex) and the emitted
 * radiance Le are differentiated. The loss enters through A = sum_q dLoss/dI_q w(q, s) / W_q
 * over the sample's (tile-clamped) filter footprint, W_q = the film's sum of weights. */

typedef struct { double x, y, z; } v3d;
static v3d d3(v3 a) { v3d r = {a.x, a.y, a.z}; return r; }
static v3d d3_make(double x, double y, double z) { v3d r = {x, y, z}; return r; }
static double d3_dot(v3d a, v3d b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3d d3_norm(v3d a) { const double l = sqrt(d3_dot(a, a)); return d3_make(a.x / l, a.y / l, a.z / l); }

/* d ln f / d alpha of f = F G D / (4 |V.N| |L.N|) (isotropicGgx.cpp:11-69): D and both G1
 * factors depend on alpha, F does not. In double from the float vectors. */
static double ggx_dlnf_dalpha(double a, v3 Vf, v3 Lf, v3 Nf) {
    const v3d V = d3(Vf), L = d3(Lf), N = d3(Nf);
    const double vn = d3_dot(V, N), ln = d3_dot(L, N);
    const double sg = vn > 0 ? 1.0 : (vn < 0 ? -1.0 : 0.0);
    const v3d H = d3_norm(d3_make((V.x + L.x) * sg, (V.y + L.y) * sg, (V.z + L.z) * sg));
    const double c = d3_dot(N, H), c2 = c * c;
    const double tan2 = (1.0 - c2) / c2;
    double r = 2.0 / a - 4.0 * a / (a * a + tan2);
    const double xs[2] = {vn, ln};
    for (int i = 0; i < 2; ++i) {
        const double t2 = (1.0 - xs[i] * xs[i]) / (xs[i] * xs[i]);
        const double sq = sqrt(1.0 + a * a * t2);
        r -= a * t2 / (sq * (1.0 + sq));
    }
    return r;
}
/* d alpha / d r of TRoughnessMapper<SQUARE> (alpha = max(max(r, 1e-3)^2, 1e-3)) */
static double dalpha_dr(float r) {
    const float safe = m_max(r, 0.001f);
    return (r > 0.001f && safe * safe > 0.001f) ? 2.0 * (double)r : 0.0;
}

/* The texel a vertex reads; checkerboards carry no gradient in the oracle. */
static int64_t tex_texel(const or_texture* t, v3 uvw) {
    if (!t->data || t->odd) return -1;
    float u, v;
    if (t->wrap_clamp) { u = m_clamp(uvw.x, 0.0f, 1.0f); v = m_clamp(uvw.y, 0.0f, 1.0f); }
    else { u = uvw.x - floorf(uvw.x); v = uvw.y - floorf(uvw.y); }
    int32_t x = (int32_t)(u * (float)t->width);
    int32_t y = (int32_t)(v * (float)t->height);
    x = (x < t->width) ? x : x - 1;
    y = (y < t->height) ? y : y - 1;
    return (int64_t)y * t->width + x;
}

/* df/dtheta of the vertex's differentiable parameter at direction L (wo in si):
 * Lambertian: per channel, df_c/da_c; GGX: df_c/dr. Zero where f vanishes identically. */
static v3d bsdf_dfdtheta(const or_material* m, const or_si* si, v3 f) {
    const float vn = v3_dot(si->wi, si->Ns), ln = v3_dot(si->wo, si->Ns);
    if (vn * ln <= 0.0f) return d3_make(0, 0, 0);
    if (m->type == OR_MAT_LAMBERT) return d3_make(OR_INV_PI, OR_INV_PI, OR_INV_PI);
    if (v3_is_zero(f)) return d3_make(0, 0, 0);
    const float r = tex_eval(&m->roughness, si->uvw).x;
    const double k = ggx_dlnf_dalpha((double)rough_to_alpha(r), si->wi, si->wo, si->Ns) * dalpha_dr(r);
    return d3_make(f.x * k, f.y * k, f.z * k);
}

typedef struct {
    int     mat;
    int64_t texel;          /* texel of the vertex's differentiable texture, -1 none */
    int     emit_li;        /* emission counted here (bounce 0), light index or -1 */
    int     li;             /* picked light */
    double  pick;
    int     lt_on, mt_on, cont;
    v3d     Le;             /* radiance of the picked light */
    v3d     lt_f, lt_df, mt_f, mt_df, c_f, c_df;
    double  lt_k, mt_k, c_k; /* w |cos| / pdf (light, MIS); |cos| / pdf * rr (continuation) */
} adj_vertex;

#define OR_ADJ_MAXV 64

typedef struct {
    const or_scene* s;
    uint64_t seed;
    int spp_begin, spp_end;
    const float* film;
    const float* dimg;
    int64_t ntiles;
    volatile int64_t next;
    pthread_mutex_t lock;
    int64_t nparams;
    const int64_t* alb_off;   /* per material: offset in the flat parameter array or -1 */
    const int64_t* rough_off;
    int64_t light_off;
} adj_job;

typedef struct { adj_job* J; double* acc; } adj_worker_arg;

/* One path: estimate() with every factor recorded (ptEstimator.cpp:25-123,
 * directLightEvaluator.cpp:18-109). Returns the number of vertices. */
static int adj_record(const or_scene* s, v3 o, v3 d, or_rng* rng, adj_vertex* vs) {
    v3 ro = o, rd = d;
    v3 beta = v3_make(1, 1, 1);
    int n = 0;
    for (int bounce = 0; bounce < s->desc.max_depth && n < OR_ADJ_MAXV; ++bounce) {
        or_si si;
        if (!scene_intersect(s, ro, rd, OR_RAY_EPS, FLT_MAX, &si, NULL)) break;
        const or_prim* p = &s->prims[si.prim];
        const or_material* m = &s->mats[p->mat];
        adj_vertex* V = &vs[n++];
        memset(V, 0, sizeof(*V));
        V->mat = p->mat;
        V->texel = tex_texel(m->type == OR_MAT_LAMBERT ? &m->albedo : &m->roughness, si.uvw);
        V->emit_li = -1;
        if (p->light >= 0 && bounce == 0 && can_emit(&s->lights[p->light], si.wi, si.Ns)) V->emit_li = p->light;
        float pickPdf;
        const int li = pick_light(s, rng, bounce, &pickPdf);
        V->li = li;
        V->pick = pickPdf;
        V->Le = d3(s->lights[li].radiance);
        or_si sl = si;
        {
            const or_light_smp ls = light_sample(s, li, si.P, rng, bounce);
            if (ls.valid) {
                const v3 LV = v3_sub(ls.pos, si.P);
                const float dist = v3_len(LV);
                const v3 L = v3_divs(LV, dist);
                sl.wo = L;
                const v3 sd = v3_normalize(L);
                if (!or_occluded(s, si.P, sd, OR_RAY_EPS, dist - OR_RAY_EPS)) {
                    const v3 f = bsdf_eval(m, &sl);
                    const float bpdf = bsdf_pdf(m, &sl);
                    const float w = mis_power(ls.pdfW, bpdf);
                    V->lt_on = 1;
                    V->lt_f = d3(f);
                    V->lt_df = bsdf_dfdtheta(m, &sl, f);
                    V->lt_k = (double)w * (double)v3_absdot(L, si.Ns) / (double)ls.pdfW;
                }
            }
        }
        {
            const or_bsdf_sample bs = bsdf_sample_rng(m, &si, rng, bounce, SL_MIS0, SL_MIS1, SL_MISC);
            if (bs.valid) {
                or_si hit;
                if (scene_intersect(s, si.P, v3_normalize(bs.L), OR_RAY_EPS, FLT_MAX, &hit, NULL) &&
                    s->prims[hit.prim].light == li && can_emit(&s->lights[li], hit.wi, hit.Ns)) {
                    const float lpdf = light_pdfW(s, li, &hit, si.P);
                    const float w = mis_power(bs.pdf, lpdf);
                    or_si sb = si;
                    sb.wo = bs.L;
                    V->mt_on = 1;
                    V->mt_f = d3(bs.f);
                    V->mt_df = bsdf_dfdtheta(m, &sb, bs.f);
                    V->mt_k = (double)w * (double)v3_absdot(bs.L, si.Ns) / (double)bs.pdf;
                }
            }
        }
        const or_bsdf_sample bs = bsdf_sample_rng(m, &si, rng, bounce, SL_CONT0, SL_CONT1, SL_CONTC);
        if (!bs.valid) break;
        const float LdotN = v3_absdot(bs.L, si.Ns);
        beta = v3_mul(beta, v3_divs(v3_scale(bs.f, LdotN), bs.pdf));
        double rrw = 1.0;
        if (bounce > 2) {
            const float q = m_clamp(1.0f - v3_max_comp(beta), 0.05f, 1.0f);
            const float u = rng_u(rng, bounce, SL_RR);
            if (u > q) { beta = v3_scale(beta, 1.0f / (1.0f - q)); rrw = (double)(1.0f / (1.0f - q)); }
            else break;
        }
        if (v3_is_zero(beta)) break;
        if (bounce + 1 >= s->desc.max_depth) break;
        or_si sc = si;
        sc.wo = bs.L;
        V->cont = 1;
        V->c_f = d3(bs.f);
        V->c_df = bsdf_dfdtheta(m, &sc, bs.f);
        V->c_k = (double)LdotN / (double)bs.pdf * rrw;
        ro = si.P;
        rd = v3_normalize(bs.L);
    }
    return n;
}

static void adj_path(adj_job* J, double* acc, uint32_t pixel, int k, float fx, float fy, v3d A) {
    const or_scene* s = J->s;
    adj_vertex vs[OR_ADJ_MAXV];
    or_rng rng;
    rng_init(&rng, OR_RNG_PHILOX, J->seed, 0, pixel, (uint32_t)k);
    const int n = adj_record(s, s->cam_pos, camera_dir(s, fx, fy), &rng, vs);
    /* suffix sums S_j (per channel) */
    double S[OR_ADJ_MAXV + 1][3];
    S[n][0] = S[n][1] = S[n][2] = 0.0;
    for (int j = n - 1; j >= 0; --j) {
        const adj_vertex* V = &vs[j];
        const double Lec[3] = {V->Le.x, V->Le.y, V->Le.z};
        const double ltf[3] = {V->lt_f.x, V->lt_f.y, V->lt_f.z}, mtf[3] = {V->mt_f.x, V->mt_f.y, V->mt_f.z};
        const double cf[3] = {V->c_f.x, V->c_f.y, V->c_f.z};
        for (int c = 0; c < 3; ++c) {
            double e = 0.0;
            if (V->emit_li >= 0) {
                const v3 le = s->lights[V->emit_li].radiance;
                e = c == 0 ? le.x : (c == 1 ? le.y : le.z);
            }
            const double nee = (V->lt_on * Lec[c] * ltf[c] * V->lt_k + V->mt_on * Lec[c] * mtf[c] * V->mt_k) / V->pick;
            S[j][c] = e + nee + (V->cont ? cf[c] * V->c_k * S[j + 1][c] : 0.0);
        }
    }
    /* forward over the vertices with beta_j */
    double beta[3] = {1.0, 1.0, 1.0};
    const double Ad[3] = {A.x, A.y, A.z};
    for (int j = 0; j < n; ++j) {
        const adj_vertex* V = &vs[j];
        const double Lec[3] = {V->Le.x, V->Le.y, V->Le.z};
        const double ltf[3] = {V->lt_f.x, V->lt_f.y, V->lt_f.z}, mtf[3] = {V->mt_f.x, V->mt_f.y, V->mt_f.z};
        const double ltd[3] = {V->lt_df.x, V->lt_df.y, V->lt_df.z}, mtd[3] = {V->mt_df.x, V->mt_df.y, V->mt_df.z};
        const double cd[3] = {V->c_df.x, V->c_df.y, V->c_df.z}, cf[3] = {V->c_f.x, V->c_f.y, V->c_f.z};
        const or_material* m = &s->mats[V->mat];
        const int64_t off = m->type == OR_MAT_LAMBERT ? J->alb_off[V->mat] : J->rough_off[V->mat];
        for (int c = 0; c < 3; ++c) {
            /* texel of this vertex */
            const double dth = (V->lt_on * Lec[c] * ltd[c] * V->lt_k + V->mt_on * Lec[c] * mtd[c] * V->mt_k) / V->pick +
                               (V->cont ? cd[c] * V->c_k * S[j + 1][c] : 0.0);
            if (off >= 0 && V->texel >= 0) {
                if (m->type == OR_MAT_LAMBERT) acc[off + 3 * V->texel + c] += Ad[c] * beta[c] * dth;
                else acc[off + V->texel] += Ad[c] * beta[c] * dth;
            }
            /* emitted radiance */
            if (V->emit_li >= 0) acc[J->light_off + 3 * V->emit_li + c] += Ad[c] * beta[c];
            const double dle = (V->lt_on * ltf[c] * V->lt_k + V->mt_on * mtf[c] * V->mt_k) / V->pick;
            acc[J->light_off + 3 * V->li + c] += Ad[c] * beta[c] * dle;
        }
        if (V->cont)
            for (int c = 0; c < 3; ++c) beta[c] *= cf[c] * V->c_k;
    }
}

/* The sample's adjoint weight A: its filter footprint (FilmTile::addSample's, clamped to
 * the tile; NaN / inf samples never reach the film and get none) against dLoss/dI / W. */
static v3d adj_weight(const adj_job* J, int tx0, int ty0, int tx1, int ty1, float fx, float fy) {
    const or_scene_desc* d = &J->s->desc;
    const float hx = d->filter_w / 2.0f, hy = d->filter_h / 2.0f;
    float mnx = std_max(fx - hx, (float)tx0), mny = std_max(fy - hy, (float)ty0);
    float mxx = std_min(fx + hx, (float)tx1), mxy = std_min(fy + hy, (float)ty1);
    const int x0 = (int)ceilf(mnx - 0.5f), y0 = (int)ceilf(mny - 0.5f);
    const int x1 = (int)(floorf(mxx - 0.5f) + 1), y1 = (int)(floorf(mxy - 0.5f) + 1);
    v3d A = d3_make(0, 0, 0);
    for (int iy = y0; iy < y1; ++iy)
        for (int ix = x0; ix < x1; ++ix) {
            const size_t q = (size_t)iy * (size_t)d->width + (size_t)ix;
            const double W = J->film[4 * q + 3];
            if (W == 0.0) continue;
            const double w = filter_eval(d, (float)ix - (fx - 0.5f), (float)iy - (fy - 0.5f)) / W;
            A.x += J->dimg[3 * q] * w;
            A.y += J->dimg[3 * q + 1] * w;
            A.z += J->dimg[3 * q + 2] * w;
        }
    return A;
}

static void* adj_worker(void* arg) {
    adj_worker_arg* W = (adj_worker_arg*)arg;
    adj_job* J = W->J;
    const or_scene_desc* d = &J->s->desc;
    const int ntx = (d->width + d->tile_x - 1) / d->tile_x;
    while (1) {
        pthread_mutex_lock(&J->lock);
        const int64_t t = J->next++;
        pthread_mutex_unlock(&J->lock);
        if (t >= J->ntiles) break;
        const int x0 = (int)(t % ntx) * d->tile_x, y0 = (int)(t / ntx) * d->tile_y;
        const int x1 = x0 + d->tile_x < d->width ? x0 + d->tile_x : d->width;
        const int y1 = y0 + d->tile_y < d->height ? y0 + d->tile_y : d->height;
        for (int iy = y0; iy < y1; ++iy)
            for (int ix = x0; ix < x1; ++ix)
                for (int k = J->spp_begin; k < J->spp_end; ++k) {
                    float fx, fy;
                    philox_film_pos(J->s, J->seed, ix, iy, k, &fx, &fy);
                    const v3d A = adj_weight(J, x0, y0, x1, y1, fx, fy);
                    if (A.x == 0.0 && A.y == 0.0 && A.z == 0.0) continue;
                    adj_path(J, W->acc, (uint32_t)((size_t)iy * (size_t)d->width + (size_t)ix), k, fx, fy, A);
                }
    }
    return NULL;
}

void or_adjoint(const or_scene* s, uint64_t seed, int spp_begin, int spp_end, int nthreads, const float* film,
                const float* d_image, double* const* d_albedo, double* const* d_rough, double* d_light) {
    adj_job J;
    memset(&J, 0, sizeof(J));
    J.s = s; J.seed = seed; J.spp_begin = spp_begin; J.spp_end = spp_end; J.film = film; J.dimg = d_image;
    const or_scene_desc* d = &s->desc;
    J.ntiles = (int64_t)((d->width + d->tile_x - 1) / d->tile_x) * ((d->height + d->tile_y - 1) / d->tile_y);
    int64_t* alb = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->nmats + 1));
    int64_t* rough = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->nmats + 1));
    int64_t np = 0;
    for (int m = 0; m < s->nmats; ++m) {
        const or_texture* a = &s->mats[m].albedo;
        const or_texture* r = &s->mats[m].roughness;
        alb[m] = a->data ? np : -1;
        if (a->data) np += (int64_t)a->width * a->height * 3;
        rough[m] = r->data ? np : -1;
        if (r->data) np += (int64_t)r->width * r->height;
    }
    J.alb_off = alb; J.rough_off = rough; J.light_off = np;
    np += 3 * (int64_t)s->nlights;
    J.nparams = np;
    pthread_mutex_init(&J.lock, NULL);
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    adj_worker_arg* args = (adj_worker_arg*)malloc(sizeof(adj_worker_arg) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) {
        args[i].J = &J;
        args[i].acc = (double*)calloc((size_t)np, sizeof(double));
        pthread_create(&th[i], NULL, adj_worker, &args[i]);
    }
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    for (int m = 0; m < s->nmats; ++m) {
        const or_texture* a = &s->mats[m].albedo;
        const or_texture* r = &s->mats[m].roughness;
        if (d_albedo && d_albedo[m] && alb[m] >= 0)
            for (int64_t e = 0; e < (int64_t)a->width * a->height * 3; ++e) {
                double v = 0.0;
                for (int i = 0; i < nthreads; ++i) v += args[i].acc[alb[m] + e];
                d_albedo[m][e] = v;
            }
        if (d_rough && d_rough[m] && rough[m] >= 0)
            for (int64_t e = 0; e < (int64_t)r->width * r->height; ++e) {
                double v = 0.0;
                for (int i = 0; i < nthreads; ++i) v += args[i].acc[rough[m] + e];
                d_rough[m][e] = v;
            }
    }
    if (d_light)
        for (int64_t e = 0; e < 3 * (int64_t)s->nlights; ++e) {
            double v = 0.0;
            for (int i = 0; i < nthreads; ++i) v += args[i].acc[J.light_off + e];
            d_light[e] = v;
        }
    for (int i = 0; i < nthreads; ++i) free(args[i].acc);
    free(args); free(th); free(alb); free(rough);
    pthread_mutex_destroy(&J.lock);
}

void or_set_light_radiance(or_scene* s, int li, const float* rgb) {
    if (li >= 0 && li < s->nlights) s->lights[li].radiance = v3_make(rgb[0], rgb[1], rgb[2]);
}

/* df/dtheta of material mi at the or_bsdf_query inputs (Ns, wi, wo, uvw): Lambertian
 * df_c/da_c, GGX df_c/dr, double (the derivative the path-replay gradient uses). */
void or_bsdf_dfdtheta_query(const or_scene* s, int mi, const float* in, int64_t n, double* out) {
    const or_material* m = &s->mats[mi];
    for (int64_t i = 0; i < n; ++i) {
        const float* q = in + 14 * i;
        or_si si;
        memset(&si, 0, sizeof(si));
        si.Ns = si.Ng = v3_load(q);
        si.wi = v3_load(q + 3);
        si.wo = v3_load(q + 6);
        si.uvw = v3_load(q + 9);
        const v3d g = bsdf_dfdtheta(m, &si, bsdf_eval(m, &si));
        out[3 * i] = g.x; out[3 * i + 1] = g.y; out[3 * i + 2] = g.z;
    }
}
