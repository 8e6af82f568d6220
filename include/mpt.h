/*
 * mpt.h - C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * The reference (xh5a5n6k6/cadise) has no FFI; its plugin points are C++
 * abstract classes selected by CRSD keys. Each entry point below replaces one
 * of them (paths relative to the reference's source/):
 *
 *   mpt_scene_load_crsd      SdParser::parseSd + RenderDatabase::setUpData /
 *                            prepareRender (file-io/scene-description/sdParser.cpp:19-35,
 *                            core/renderDatabase.cpp:39-131)
 *   mpt_trace / _device      Accelerator::isIntersecting / isOccluded for a batch of
 *                            rays (core/intersector/accelerator/accelerator.h:9-12,
 *                            bvh/bvhAccelerator.cpp:40-156)
 *   mpt_render               Renderer::render for renderer "sampling" + estimator "pt"
 *                            (core/renderer/sampling/samplingRenderer.cpp:33-68,
 *                            core/estimator/ptEstimator.cpp:25-123)
 *   mpt_film_*               Film::save's resolve (core/film/film.cpp:159-196)
 *
 * Conventions: every call returns MPT_OK (0) or a negative MPT_ERR_* code and
 * never aborts; mpt_last_error gives the message. Calls are synchronous unless
 * they take a stream. A context is used from one host thread at a time.
 * Device memory is owned by the library. Host buffers stay owned by the caller.
 */
#ifndef MPT_H
#define MPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MPT_OK = 0,
    MPT_ERR_INVALID = -1,     /* bad argument */
    MPT_ERR_PARSE = -2,       /* malformed CRSD or missing block */
    MPT_ERR_UNSUPPORTED = -3, /* a CRSD type this path does not implement */
    MPT_ERR_DEVICE = -4,      /* HIP error */
    MPT_ERR_IO = -5,          /* file read/write */
    MPT_ERR_NOMEM = -6
};

typedef struct mpt_ctx mpt_ctx;
typedef struct mpt_scene mpt_scene;

/* One ray of a batch query: origin, minT, unit direction, maxT (32 bytes).
 * The reference's Ray (core/ray.h) normalizes its direction on construction;
 * here the caller passes it normalized. */
typedef struct {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} mpt_ray;

/* Closest hit: prim = creation index of the primitive (the position in the
 * reference's RenderDatabase::_intersectors), -1 on a miss; t = hit distance
 * (ray tmax on a miss). Any hit: prim = 1 if occluded, 0 otherwise. */
typedef struct {
    int32_t prim;
    float   t;
} mpt_hit;

/* Primitive description for mpt_scene_create (creation order). */
/* A sphere's 9 values are (centre, radius, 0, 0, 0, 0, 0). */
enum { MPT_PRIM_TRIANGLE = 0, MPT_PRIM_RECTANGLE = 1, MPT_PRIM_SPHERE = 2 };
/* bsdfInstantiator.cpp:177-221: matte-lambertian, conductor-microfacet (GGX),
 * phong, plastic (MixedBsdf(Lambertian, Phong)), mirror, glass, transmission, absorb */
enum {
    MPT_MAT_LAMBERT = 0, MPT_MAT_GGX = 1, MPT_MAT_PHONG = 2, MPT_MAT_PLASTIC = 3, MPT_MAT_MIRROR = 4,
    MPT_MAT_GLASS = 5, MPT_MAT_TRANSMISSION = 6, MPT_MAT_ABSORB = 7
};
enum { MPT_ESTIMATOR_PT = 0, MPT_ESTIMATOR_VPT = 1 };

/* A texture: an image (data), a constant (data NULL), or, when odd and even
 * are both set, a checkerboard (TCheckerboardTexture, tCheckerboardTexture.ipp:11-40):
 * even where floor(u / (1 / odd_number)) + floor(v / (1 / even_number)) is
 * even, else odd (nested up to 8 deep). */
typedef struct mpt_texture_desc_s {
    const float* data;         /* H x W x C float32 (row 0 = v in [0, 1/H)), NULL = constant */
    int32_t      width, height, channels;
    int32_t      wrap_clamp;   /* 0 repeat, 1 clamp */
    float        value[3];     /* constant value */
    const struct mpt_texture_desc_s* odd;
    const struct mpt_texture_desc_s* even;
    float        odd_number, even_number;
} mpt_texture_desc;

typedef struct {
    int32_t          type;     /* MPT_MAT_* */
    mpt_texture_desc albedo;   /* matte-lambertian albedo; plastic diffuse-albedo; mirror / glass /
                                  transmission albedo */
    float            f0[3];    /* schlick f0: conductor-microfacet, mirror with a conductor fresnel */
    mpt_texture_desc roughness;
    float            exponent;  /* phong exponent, plastic specular-exponent */
    float            ratio;     /* plastic diffuse-ratio */
    int32_t          fresnel;   /* mirror: 0 conductor (schlick f0), 1 dielectric */
    float            ior_outer, ior_inner;   /* dielectric fresnel (mirror, glass, transmission) */
} mpt_material_desc;

enum { MPT_LIGHT_AREA = 0, MPT_LIGHT_ENVIRONMENT = 1 };

typedef struct {
    int32_t prim;              /* area: emitting primitive (creation index) */
    float   color[3];
    float   watt;
    int32_t back_face;
    int32_t type;              /* MPT_LIGHT_* */
    mpt_texture_desc env;      /* environment: radiance image, H x W x 3 in texture orientation
                                  (row 0 = v in [0, 1/H): the .hdr file's bottom row, columns
                                  mirrored, as the reference's loader leaves it) */
} mpt_light_desc;

typedef struct {
    int64_t        nprims;
    const int32_t* prim_type;  /* MPT_PRIM_* */
    const float*   prim_v;     /* nprims x 9: v1 v2 v3 */
    const float*   prim_n;     /* nprims x 9 vertex normals or NULL (zero = unset) */
    const float*   prim_uvw;   /* nprims x 9 vertex uvw or NULL (NaN first = unset) */
    const int32_t* prim_mat;
    int32_t                  nmats;
    const mpt_material_desc* mats;
    int32_t                  nlights;
    const mpt_light_desc*    lights;
    int32_t cluster;           /* 0 uniform, 1 power */
    float   cam_pos[3], cam_dir[3], cam_up[3];
    float   fov_deg, sensor_width_mm;
    int32_t width, height, tile_x, tile_y;
    int32_t filter;            /* 0 box, 1 gaussian, 2 cone, 3 mitchell */
    float   filter_w, filter_h, sigma_x, sigma_y, amplitude, mitchell_b, mitchell_c;
    int32_t sampler;           /* 0 random, 1 stratified */
    int32_t spp, max_depth;
    int32_t split_sah;         /* 1 sah, 0 equal */
    int32_t estimator;         /* MPT_ESTIMATOR_*: pt (NEE + MIS) or vpt (BSDF sampling only) */
    const char* output_filename; /* the film's output-filename; NULL: "cadise.jpg" */
} mpt_scene_desc;

/* Accelerator options (the reference's accelerator block plus the wide-BVH
 * layout chosen for gfx950). Zero fields take the defaults. */
typedef struct {
    int32_t width;             /* children per node: 4 or 8 (default 4) */
    int32_t leaf_max;          /* largest subtree folded into one leaf (default 2) */
    int32_t block;             /* traversal threads per block: 128 (default) or 256 */
    int32_t stack_lds;         /* per-lane stack entries kept in LDS: 16, 20, 24 or 32 (default 20
                                  for width 4, 16 for width 8) */
    int32_t blocks_per_cu;     /* resident traversal blocks per CU of the persistent grid (default 8
                                  for width 4, 7 for width 8) */
} mpt_bvh_opts;

typedef struct {
    int64_t nprims, nlights;
    int64_t binary_nodes;      /* nodes of the binary SAH tree the wide BVH is collapsed from */
    int64_t wide_nodes;        /* nodes of the wide BVH uploaded to HBM */
    int32_t binary_depth, wide_depth, width, leaf_max;
    double  build_ms;          /* host SAH build (the reference's "Finished building accelerator") */
    double  collapse_ms;
    int64_t device_bytes;
    int32_t image_width, image_height, spp, max_depth;
    int64_t ntextures;         /* image and constant textures, in creation order */
    int64_t tex_elements;      /* floats of all textures together (the texture-gradient length) */
    int32_t stack_lds, trace_block, blocks_per_cu;   /* the traversal launch shape in use */
} mpt_scene_stats;

typedef struct {
    uint64_t seed;
    int32_t  spp_begin, spp_end; /* sample passes [begin, end); end <= 0 takes the scene's spp */
    int32_t  max_depth;          /* <= 0 takes the scene's max-depth */
    int32_t  accumulate;         /* 1 = add into the film, 0 = clear it first */
    int32_t  flags;              /* MPT_RENDER_KEEP_SAMPLES: keep each sample's radiance for mpt_render_adjoint */
} mpt_render_params;

/* MPT_RENDER_PASSWISE: render sample pass by sample pass (each pass's paths
 * bounce in lock step; the film is gathered per tile) instead of the default
 * streaming wavefront (new paths start as paths finish, samples splatted
 * as they finish). Both draw the same paths; the films differ only in the order
 * of float additions. */
enum { MPT_RENDER_KEEP_SAMPLES = 1, MPT_RENDER_PASSWISE = 2 };

/* Counters of the last render or trace call. */
typedef struct {
    int64_t camera_rays, continuation_rays, shadow_rays, mis_rays;
    int64_t samples;
    int64_t box_tests, tri_tests; /* closest-hit traversal work (slab tests, primitive tests) */
    int64_t box_tests_any, tri_tests_any; /* any-hit traversal work */
    int64_t closest_launches, anyhit_launches;  /* the streaming renderer traces its shadow rays in
                                                   the closest-hit launches: anyhit_launches 0 */
    double  ms_total;             /* wall time of the call */
    double  ms_closest, ms_anyhit, ms_shade, ms_camera, ms_film; /* summed kernel time (events); the
                                     streaming renderer's ms_closest includes its shadow rays */
    double  ms_adjoint;           /* mpt_render_adjoint: GPU time of the replay (events); the other
                                     fields then describe the forward render it replays */
    double  ms_gpu;               /* GPU time of the call's kernels (events; mpt_batch_render: the
                                     streaming loop) */
    int64_t path_slots;           /* streaming renderer: capacity for paths in flight (0 pass by pass) */
    int64_t mis_resolved;         /* MIS-BSDF queries answered without traversal: the ray misses the
                                     picked area light's own primitive, so its closest hit cannot be
                                     that light (mis_rays counts the traced ones) */
    int64_t shadow_resolved;      /* shadow queries answered without traversal: the NEE term they
                                     would gate is exactly zero (shadow_rays counts the traced ones) */
} mpt_stats;

/* device >= 0: a GPU context. device == -1: a host-only context, in which
 * scenes are parsed and their BVHs built but nothing is uploaded; trace and
 * render calls on its scenes return MPT_ERR_DEVICE (for inspection on
 * machines without a GPU). */
int         mpt_create(int device, mpt_ctx** out);
void        mpt_destroy(mpt_ctx* ctx);
const char* mpt_last_error(const mpt_ctx* ctx);
const char* mpt_version(void);

/* At most MPT_MAX_PRIMITIVES primitives per scene, and per batch in all (the
 * device addresses a primitive's three 16-byte records with 32-bit index
 * arithmetic, 3 * index + 2 < 2^31); a larger scene returns MPT_ERR_UNSUPPORTED
 * before any device work. */
#define MPT_MAX_PRIMITIVES ((int64_t)715827882)

/* Scene from CRSD (same grammar, keys and defaults as the reference reader). */
int  mpt_scene_load_crsd(mpt_ctx* ctx, const char* path, const mpt_bvh_opts* opts, mpt_scene** out);
int  mpt_scene_parse_crsd(mpt_ctx* ctx, const char* text, size_t len, const mpt_bvh_opts* opts, mpt_scene** out);
/* Scene from arrays (generated scenes). */
int  mpt_scene_create(mpt_ctx* ctx, const mpt_scene_desc* desc, const mpt_bvh_opts* opts, mpt_scene** out);
void mpt_scene_destroy(mpt_scene* scene);
int  mpt_scene_stats_get(const mpt_scene* scene, mpt_scene_stats* out);
/* The reference's binary BVH, node for node as BvhBuilder builds it (built on
 * the first call; the device traverses a tree built by the same algorithm on
 * true empty boxes, see mpt_scene_stats.binary_nodes): node bounds
 * (n x 6), is_leaf, a (leaf: first ordered primitive; interior: second child),
 * b (leaf: count; interior: split axis) and the ordered primitive list. Pass
 * NULL arrays to query the node count in *count. */
int  mpt_scene_binary_bvh(mpt_scene* scene, int64_t* count, float* bounds, int32_t* is_leaf, int64_t* a,
                          int64_t* b, int32_t* ordered);

/* Camera matrices as the reference builds them (perspectivePinholeCamera.cpp:32-50):
 * film->camera and camera->world, row-major 4x4 float. */
int  mpt_scene_camera(const mpt_scene* scene, float* film_to_camera, float* camera_to_world);
/* Emitted radiance of each light (nlights x 3) as SingleAreaLight's constructor
 * computes it (singleAreaLight.cpp:20-37). */
int  mpt_scene_light_radiance(const mpt_scene* scene, float* out);
/* Probability of picking each light (nlights floats) as the scene's light cluster
 * gives it: LightCluster::evaluatePickLightPdf (powerLightCluster.cpp:38-46,
 * uniformLightCluster.cpp:26-29); the power cluster weighs area lights by
 * SingleAreaLight::approximateFlux and the environment light by
 * EnvironmentLight::approximateFlux (environmentLight.cpp:125-139). */
int  mpt_scene_light_pick_pdf(const mpt_scene* scene, float* out);

/* Batch ray queries; host buffers (copied) or device buffers on a stream.
 * At most MPT_TRACE_MAX_RAYS rays per call (2^31 - 2^24: ray indices stay below
 * 2^31 and the traversal's unsigned 32-bit work counter, which each wave may
 * advance past n by one chunk of 512 rays, below 2^32); more returns
 * MPT_ERR_INVALID before any device work. */
#define MPT_TRACE_MAX_RAYS ((size_t)2147483648u - (size_t)16777216u)
int  mpt_trace(mpt_scene* scene, const mpt_ray* rays, size_t n, int any_hit, mpt_hit* hits);
int  mpt_trace_device(mpt_scene* scene, const mpt_ray* rays_dev, size_t n, int any_hit, mpt_hit* hits_dev,
                      void* stream);
/* Closest hits with their surface detail, as Scene::isIntersecting fills a
 * SurfaceIntersection (core/scene.cpp:30-48, triangle.cpp:125-160,
 * rectangle.cpp:102-132, sphere.cpp:109-130): per ray 14 floats (primitive
 * creation index, t, P, Ng, Ns, uvw); a miss gives -1 and zeros. Host buffers;
 * n <= MPT_TRACE_MAX_RAYS. */
int  mpt_intersect_detail(mpt_scene* scene, const mpt_ray* rays, size_t n, float* out);
/* mpt_trace_device that allocates nothing. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The same traversal and
 * results, queued on stream (NULL: the default stream) without waiting; its
 * work counter, statistics and traversal-stack spill live in one fixed array of
 * MPT_TRACE_SCRATCH_BYTES per device, held until the process ends (statistics
 * are not read back: mpt_stats_get is not updated). Limits: n <= MPT_TRACE_MAX_RAYS;
 * a tree whose stack spill for one block exceeds the array: MPT_ERR_INVALID (a
 * deep tree runs with fewer persistent blocks). Every call on one GPU shares its
 * array, so calls that overlap on one GPU (any streams) race; serialize them. */
#define MPT_TRACE_SCRATCH_BYTES ((size_t)32u << 20)
int  mpt_trace_device_noalloc(mpt_scene* scene, const mpt_ray* rays_dev, size_t n, int any_hit, mpt_hit* hits_dev,
                              void* stream);

/* Sampling renderer + pt estimator. film_host (W x H x 4: sum w*rgb, sum w)
 * receives the linear film when non-NULL. */
int  mpt_render(mpt_scene* scene, const mpt_render_params* params, float* film_host);
/* Device pointer of the scene's linear film (W x H x float4). */
int  mpt_film_device(mpt_scene* scene, void** film_dev);
/* Film::save's resolve (film.cpp:159-196, imageUtils.cpp:23-36): sum w*rgb / sum w,
 * sRGB OETF (math.cpp:92-99), 8-bit (W x H x 3, top row first). */
int  mpt_film_srgb8(mpt_scene* scene, uint8_t* out);
/* The same resolve of a film already in device memory (npix float4: sum w*rgb,
 * sum w) into npix x 3 device bytes, queued on stream (NULL: the default stream);
 * for callers that keep their films on the GPU. */
int  mpt_film_srgb8_device(mpt_ctx* ctx, const float* film_dev, int64_t npix, uint8_t* out_dev, void* stream);
/* Film::save (film.cpp:159-196) + PictureSaver::save (pictureSaver.cpp:12-82):
 * writes the film by the file extension: .png, .jpg / .jpeg (baseline JPEG,
 * quality 92), .ppm (the reference's "P6 w h 255" header), and .pfm (linear,
 * not in the reference). path NULL: the film block's output-filename (default
 * "cadise.jpg", filmInstantiator.cpp:24). Any other extension:
 * MPT_ERR_UNSUPPORTED (the reference then writes nothing). */
int  mpt_film_save(mpt_scene* scene, const char* path);
/* The film block's output-filename. *needed (if non-NULL) receives its length + 1;
 * buf (if non-NULL) receives it NUL-terminated when n >= that length. */
int  mpt_scene_output_filename(const mpt_scene* scene, char* buf, size_t n, size_t* needed);
/* PictureSaver::saveLdrImage (pictureSaver.cpp:30-84) of 8-bit RGB (height rows of
 * width x 3 bytes, top row first) by extension (.png, .jpg / .jpeg, .ppm). Host
 * only: needs no context or device. */
int  mpt_image_write(const char* path, const uint8_t* rgb, int32_t width, int32_t height);

int  mpt_stats_get(const mpt_scene* scene, mpt_stats* out);
/* Frees the context's streaming path state (the vertex queues that carry the
 * paths in flight: up to 2^24 paths, ~8 GB, for single-scene renders of more
 * than 2^23 samples; the adjoint keeps at most 2^22, a batch its own
 * 2^16..2^24). The next streaming render allocates it again. */
int  mpt_release_stream_buffers(mpt_ctx* ctx);

/* ---- differentiable rendering (no counterpart in the reference) -------------------
 * Parameters: the texels of every texture (albedo of matte-lambertian, roughness
 * of conductor-microfacet; constant textures are 1x1) and the emitted radiance
 * Le of every light (SingleAreaLight's radiance, singleAreaLight.cpp:32-36).
 * Texture i occupies elements [offset, offset + W*H*C) of the flat texture
 * array (row-major H x W x C). */
int  mpt_scene_texture_info(const mpt_scene* scene, int32_t index, int64_t* width, int64_t* height,
                            int64_t* channels, int64_t* offset);
/* Replace texture `index` (W x H x C host floats) / one light's radiance (3 floats).
 * An environment light's texture and radiance are not parameters
 * (MPT_ERR_UNSUPPORTED). */
int  mpt_scene_set_texture(mpt_scene* scene, int32_t index, const float* data);
/* Copy texture `index` (W x H x C floats, as the renderer samples it: an
 * environment image after the loader's flips) to host memory. */
int  mpt_scene_get_texture(const mpt_scene* scene, int32_t index, float* out);
int  mpt_scene_set_light_radiance(mpt_scene* scene, int32_t light, const float* rgb);
/* Path-replay adjoint of an image loss: given d_image = dLoss/dI (W x H x 3,
 * host; I = the resolved film sum w*rgb / sum w), accumulate dLoss/d(texel)
 * into d_tex (tex_elements floats, host, overwritten) and dLoss/dLe into
 * d_light (nlights x 3, overwritten); either may be NULL. Replays exactly the
 * paths of mpt_render with the same seed, passes and depth (it renders them
 * first, keeping samples, unless the last mpt_render call did so with the same
 * parameters). Sampling, MIS weights and Russian roulette are held fixed;
 * only BSDF values and emitted radiance are differentiated. */
int  mpt_render_adjoint(mpt_scene* scene, const mpt_render_params* params, const float* d_image, float* d_tex,
                        float* d_light);

/* ---- batched rendering (no counterpart in the reference) -----------------------------
 * Renders many scenes (each with its own camera, film, spp and max-depth) in
 * one streaming wavefront: the scenes' BVHs and tables are merged into one
 * device scene, new paths are drawn from a global work counter over
 * (scene, sample, pixel), and the bounce loop runs as a captured HIP graph.
 * Each sample is the same path that mpt_render draws for that scene with the
 * same seed; films differ from mpt_render's only by floating-point summation
 * order. The scenes must stay alive while the batch exists. */
typedef struct mpt_batch mpt_batch;

typedef struct {
    int32_t width;             /* BVH width, 4 or 8 */
    int32_t stack_lds;         /* per-lane LDS stack entries: 16, 20, 24 or 32 */
    int32_t block;             /* traversal threads per block: 128 or 256 */
    int32_t capacity_log2;     /* paths in flight = 2^capacity_log2 (16..24) */
    int32_t graph_steps;       /* wavefront steps per graph launch (even, 2..64) */
    int32_t blocks_per_cu;     /* resident traversal blocks per CU (2..32) */
} mpt_batch_opts;

typedef struct {
    mpt_batch_opts opts;
    double ms;                 /* timed trial (HIP events) */
} mpt_batch_trial;

int  mpt_batch_create(mpt_ctx* ctx, mpt_scene* const* scenes, int32_t n, const mpt_batch_opts* opts, mpt_batch** out);
void mpt_batch_destroy(mpt_batch* batch);
/* seeds: one per scene (NULL = all 0); spp_override > 0 renders that many
 * samples per pixel of every scene instead of each scene's own. */
int  mpt_batch_render(mpt_batch* batch, const uint64_t* seeds, int32_t spp_override, mpt_stats* stats);
/* Scene i's linear film (W x H x 4: sum w*rgb, sum w). */
int  mpt_batch_film(mpt_batch* batch, int32_t i, float* film_host);
/* Autotuner: times trial renders (spp_override = trial_spp) over launch shapes
 * and path-slot counts by coordinate descent, keeps the fastest in the batch
 * and returns it; trials (up to max_trials) receives every timing. */
int  mpt_batch_autotune(mpt_batch* batch, const uint64_t* seeds, int32_t trial_spp, mpt_batch_opts* best,
                        mpt_batch_trial* trials, int32_t max_trials, int32_t* ntrials);
int  mpt_batch_get_opts(const mpt_batch* batch, mpt_batch_opts* out);
int  mpt_batch_set_opts(mpt_batch* batch, const mpt_batch_opts* opts);

#ifdef __cplusplus
}
#endif
#endif
