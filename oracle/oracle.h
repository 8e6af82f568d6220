/*
 * Oracle: CPU restatement of the reference's unidirectional path tracer
 * (renderer "sampling" + estimator "pt" + accelerator "bvh"). TEST
 * INFRASTRUCTURE ONLY: tests/, bench.py's CPU-baseline leg and
 * __graft_entry__.smoke() use it as the checker. The product never links,
 * loads or calls it.
 *
 * Every function restates one reference function, cited file:line
 * (paths relative to the reference's source/). Floating-point operations
 * follow the reference's operation order, IEEE binary32, no contraction
 * (built with -ffp-contract=off), and use the same libm entry points the
 * reference calls (acosf, cosf, sinf, atanf, sqrtf, expf, powf, tanf), so that
 * with the same random stream the oracle reproduces the compiled reference
 * bit for bit (checked against oracle/_ref by tests/test_oracle_vs_ref.py).
 *
 * Two random-number layouts:
 *   OR_RNG_SEQ    one SplitMix64 stream per (seed, pixel, sample), consumed in
 *                 the reference's call order (oracle/seqrng.h);
 *   OR_RNG_PHILOX Philox4x32-10 keyed per (pixel, sample, bounce, slot), the
 *                 layout the GPU wavefront tracer uses, so that a GPU path and
 *                 an oracle path draw the same numbers.
 */
#ifndef MPT_ORACLE_H
#define MPT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_PRIM_TRIANGLE = 0, OR_PRIM_RECTANGLE = 1, OR_PRIM_SPHERE = 2 };
enum {
    OR_MAT_LAMBERT = 0, OR_MAT_GGX = 1, OR_MAT_PHONG = 2, OR_MAT_PLASTIC = 3, OR_MAT_MIRROR = 4, OR_MAT_GLASS = 5,
    OR_MAT_TRANSMISSION = 6, OR_MAT_ABSORB = 7
};
enum { OR_EST_PT = 0, OR_EST_VPT = 1 };
enum { OR_FILTER_BOX = 0, OR_FILTER_GAUSSIAN = 1, OR_FILTER_CONE = 2, OR_FILTER_MITCHELL = 3 };
enum { OR_SAMPLER_RANDOM = 0, OR_SAMPLER_STRATIFIED = 1 };
enum { OR_CLUSTER_UNIFORM = 0, OR_CLUSTER_POWER = 1 };
enum { OR_RNG_SEQ = 0, OR_RNG_PHILOX = 1 };
enum { OR_LIGHT_AREA = 0, OR_LIGHT_ENVIRONMENT = 1 };

/* A texture: constant when data == NULL; a checkerboard of odd / even when
 * both are set (odd_size = 1 / odd-number, even_size = 1 / even-number). */
typedef struct or_texture_s {
    const float* data;     /* width*height*channels, row-major, row 0 = v 0 */
    int32_t      width, height, channels;
    int32_t      wrap_clamp; /* 0 repeat, 1 clamp */
    float        value[3];   /* constant value */
    const struct or_texture_s* odd;
    const struct or_texture_s* even;
    float        odd_size, even_size;
} or_texture;

typedef struct {
    int32_t    type;       /* OR_MAT_* */
    or_texture albedo;     /* lambert albedo; plastic diffuse-albedo; mirror / glass / transmission albedo */
    float      f0[3];      /* ggx / conductor-mirror schlick f0 */
    or_texture roughness;  /* ggx roughness (1 channel) */
    float      exponent;   /* phong / plastic specular exponent */
    float      ratio;      /* plastic diffuse ratio */
    int32_t    fresnel;    /* mirror: 0 schlick conductor, 1 dielectric */
    float      ior_outer, ior_inner;
} or_material;

typedef struct {
    int32_t prim;          /* index of the emitting primitive (area lights) */
    float   color[3];
    float   watt;
    int32_t back_face;
    int32_t    type;       /* OR_LIGHT_* */
    or_texture env;        /* environment radiance (image, 3 channels, repeat) */
} or_light;

typedef struct {
    int64_t        nprims;
    const int32_t* prim_type;   /* OR_PRIM_* */
    const float*   prim_v;      /* nprims x 9: v1 v2 v3 (rectangle: vA vB vC; sphere: centre, radius) */
    const float*   prim_n;      /* nprims x 9 vertex normals (triangles) or NULL */
    const float*   prim_uvw;    /* nprims x 9 vertex uvw (triangles) or NULL */
    const int32_t* prim_mat;    /* material per primitive */
    int32_t            nmats;
    const or_material* mats;
    int32_t            nlights;
    const or_light*    lights;
    int32_t cluster;            /* OR_CLUSTER_* */
    /* camera */
    float cam_pos[3], cam_dir[3], cam_up[3];
    float fov_deg, sensor_width_mm;
    /* film */
    int32_t width, height, tile_x, tile_y;
    int32_t filter;             /* OR_FILTER_* */
    float   filter_w, filter_h, sigma_x, sigma_y, amplitude, mitchell_b, mitchell_c;
    /* renderer */
    int32_t sampler, spp, max_depth;
    int32_t split_sah;          /* 1 sah, 0 equal */
    int32_t estimator;          /* OR_EST_* */
} or_scene_desc;

typedef struct or_scene or_scene;

or_scene* or_scene_create(const or_scene_desc* desc);
void      or_scene_destroy(or_scene* s);

/* Binary BVH exactly as BvhAccelerator builds it. Returns the node count and
 * fills the arrays when non-null: bounds (n x 6), is_leaf, a (leaf: first
 * ordered primitive; interior: second child), b (leaf: count; interior: axis),
 * ordered (nprims, creation indices). */
int64_t or_bvh_export(const or_scene* s, float* bounds, int32_t* is_leaf, int64_t* a, int64_t* b,
                      int32_t* ordered);

/* rays: n x 8 (ox oy oz minT dx dy dz maxT), directions already normalized.
 * Closest hit: prim (-1 on miss) and t. Any hit: prim = 1/0.
 * use_bvh 0 = brute force over all primitives in creation order. */
void or_trace(const or_scene* s, const float* rays, int64_t n, int any_hit, int use_bvh, int32_t* out_prim,
              float* out_t);
/* For closest hits: the number of primitives whose hit distance equals the
 * closest one exactly (ties), and whether any such hit lies on an edge. */
void or_trace_ties(const or_scene* s, const float* rays, int64_t n, int32_t* out_ntie);

/* Scene::isIntersecting + SurfaceDetail: out n x 14 (prim, t, P, Ng, Ns, uvw). */
void or_intersect_detail(const or_scene* s, const float* rays, int64_t n, float* out);

/* Camera: the reference's film->camera and camera->world matrices as float
 * (row-major 4x4 each) and the primary ray for film positions. */
void or_camera_matrices(const or_scene* s, float* film_to_camera, float* camera_to_world);
void or_camera_rays(const or_scene* s, const float* film_xy, int64_t n, float* out /* n x 6 */);
void or_normalize(const float* v, int64_t n, float* out);

/* BSDF of material m at one shading point; in/out layout as ref_bsdf_query. */
void or_bsdf_query(const or_scene* s, int m, const float* in, int64_t n, float* out);
/* The same with nu uniforms per query in draw order: in n x (12 + nu). */
void or_bsdf_query_n(const or_scene* s, int m, const float* in, int64_t n, int nu, float* out);
/* Light li direct sample towards target, explicit uniforms: in n x (3 + nu). */
void or_light_sample(const or_scene* s, int li, const float* in, int64_t n, int nu, float* out);
void or_light_hit(const or_scene* s, const float* rays, int64_t n, float* out);
/* Emitted radiance Le of every light (n x 3; environment lights: 0). */
void or_light_radiance(const or_scene* s, float* out);
/* Environment light li: the power cluster's flux estimate and background flux
 * (out[0], out[1]), and its Distribution2D tables: marginal CDF (h + 1) then
 * the h conditional CDFs (w + 1 each) into cdf when non-null. Returns 0, or -1
 * when li is not an environment light. */
int or_env_info(const or_scene* s, int li, float* out, float* cdf);
/* Power-cluster CDF over the lights (nlights + 1). */
void or_power_cdf(const or_scene* s, float* out);

/* PtEstimator::estimate from given film positions. Sequential layout: stream
 * keys[i]. Philox layout: keys[i] = (pixel << 32) | sample, seed = seed. */
void or_estimate(const or_scene* s, int rng_mode, uint64_t seed, const float* film_xy, const uint64_t* keys,
                 int64_t n, float* out_rgb, int64_t* out_draws);

/* Sampling renderer, linear film (n = w*h*4: sum w*rgb, sum w).
 * Sequential layout: same keying as the reference harness.
 * Philox layout: the GPU tracer's sample placement and random layout. */
void or_render_linear(const or_scene* s, int rng_mode, uint64_t seed, int spp_begin, int spp_end, int nthreads,
                      float* film);

/* CPU baseline: pixels [begin, end), spp each, nthreads; returns seconds. */
double or_time_render(const or_scene* s, int64_t begin, int64_t end, int spp, int nthreads, uint64_t seed,
                      int64_t* out_rays, int64_t* out_samples);

/* Gradient of an image loss by path replay, FP64 accumulation (host check of the GPU
 * adjoint; not a reference function). Paths of the Philox layout, passes
 * [spp_begin, spp_end); film = or_render_linear's film of the same passes (its sum of
 * weights normalizes dLoss/dI); d_image = dLoss/dI (w x h x 3). Outputs per material
 * (arrays of nmats pointers, entries NULL to skip): d_albedo[m] (albedo texels x 3,
 * image textures only), d_rough[m] (roughness texels); d_light = dLoss/dLe (nlights x 3).
 * Sampling pdfs, MIS weights and Russian roulette are held fixed. */
void or_adjoint(const or_scene* s, uint64_t seed, int spp_begin, int spp_end, int nthreads, const float* film,
                const float* d_image, double* const* d_albedo, double* const* d_rough, double* d_light);
/* df/dtheta (Lambertian: d f_c / d albedo_c; GGX: d f_c / d roughness) at the
 * or_bsdf_query inputs, n x 3 doubles. */
void or_bsdf_dfdtheta_query(const or_scene* s, int m, const float* in, int64_t n, double* out);
/* Set the emitted radiance of light li (3 floats), as the GPU path's parameter update. */
void or_set_light_radiance(or_scene* s, int li, const float* rgb);

/* Philox4x32-10 block: ctr[4], key[2] -> out[4]. */
void or_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out);

/* Film resolve to 8-bit sRGB as Film::save + ImageUtils::hdrToLdr. */
void or_film_to_srgb8(const float* film, int64_t npix, uint8_t* out);
/* The 8-bit byte of a resolved linear value v as a threshold table: out[k] = bit
 * pattern of the smallest float in [+0, 1] whose byte is >= k (out[0] = 0), from an
 * exhaustive scan of [+0, 1] over nthreads threads; returns -1 if the byte is not
 * monotone there. */
int or_srgb8_thresholds(uint32_t* out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
