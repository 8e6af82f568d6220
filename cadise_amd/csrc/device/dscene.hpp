// Device-resident scene layout (HBM). All arrays are owned by DeviceScene
// (render.hip); kernels receive this POD by value.
//
//   nodes     wide BVH, QNode<W> (64 B for W = 4, 128 B for W = 8), root at 0
//   tris      3 x float4 per primitive in BVH leaf order:
//               triangle  (vA, id) (eAB, 0) (eAC, 0)          as Triangle stores them
//               rectangle (vB, id|1<<31) (eA, 0) (eB, 0)       as Rectangle stores them
//             id = creation index; bit 31 of the first w marks a rectangle
//   prim_info per primitive (creation order): material, light, has-normals, has-uvw
//   prim_n / prim_uvw  9 floats per primitive (creation order), when present
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mpt {

// Material types (host MatType): 0 lambert, 1 ggx conductor, 2 phong, 3 plastic
// (mixed lambert + phong), 4 mirror, 5 glass, 6 transmission, 7 absorber.
enum : int32_t {
    DM_LAMBERT = 0, DM_GGX = 1, DM_PHONG = 2, DM_PLASTIC = 3, DM_MIRROR = 4, DM_GLASS = 5, DM_TRANSMISSION = 6,
    DM_ABSORB = 7
};

struct DevMaterial {
    int32_t type;        // DM_*
    int32_t albedo_tex;  // -1 constant
    int32_t rough_tex;   // -1 constant
    int32_t fresnel;     // mirror: 0 schlick conductor (f0), 1 dielectric (iors)
    float   albedo[3];
    float   roughness;
    float   f0[3];
    float   exponent;    // phong / plastic
    float   ratio;       // plastic diffuse ratio
    float   ior_outer, ior_inner;
    float   pad;
};

// Kinds of traversal launch: every ray closest hit, every ray any hit, or a
// mixed launch whose lanes carry their own kind.
enum { TRACE_CLOSEST = 0, TRACE_ANY = 1, TRACE_MIXED = 2 };

// An image (a constant is a 1 x 1 image) or, when odd >= 0, a checkerboard
// of textures odd / even (its one texel in tex_data is unused).
struct DevTexture {
    int64_t offset;      // into tex_data
    int32_t width, height, channels, wrap_clamp;
    int32_t odd, even;
    float   odd_size, even_size;
};

// An area light with the geometry of its primitive (lights are few, so the
// light-sampling path reads this instead of the primitive arrays).
struct DevLight {
    int32_t prim;        // creation index
    int32_t type;        // primitive type
    int32_t back_face;
    int32_t has_normals;
    float   radiance[3];
    float   pdfA;        // 1 / area, as Primitive::evaluatePositionPdfA computes it
    float   a[3];        // triangle vA | rectangle vB
    float   e1[3];       // eAB | eA
    float   e2[3];       // eAC | eB
    float   n[9];        // triangle vertex normals nA nB nC (face normal unless set)
    // environment light (type 2): radiance texture and sampling tables at
    // env_cdf: marginal CDF (env_h + 1), then env_h conditional CDFs (env_w + 1)
    int32_t env_tex, env_w, env_h, env_cdf;
    int32_t rec;         // leaf position of the primitive's record in DevScene::tris (-1: environment)
};

constexpr int32_t LIGHT_TYPE_ENV = 2;

struct DevScene {
    const void*    nodes;
    int32_t        width;       // BVH width (4 or 8)
    int32_t        max_stack;   // stack entries a traversal can need
    int64_t        nnodes;
    int64_t        nprims;
    const float4*  tris;
    const int4*    prim_info;
    const float*   prim_n;
    const float*   prim_uvw;
    const DevMaterial* mats;
    const DevTexture*  texs;
    const float*   tex_data;
    const DevLight* lights;
    const float*   power_cdf;   // nlights + 1
    int32_t        nlights;
    int32_t        env_light;   // index of the environment light, -1 none
    const float*   env_cdf;     // environment lights' sampling tables
    int32_t        cluster;     // 0 uniform, 1 power
    int32_t        estimator;   // 0 pt, 1 vpt
    // camera
    float f2c[12];              // film -> camera rows 0..2 (row-major)
    float c2w[12];              // camera -> world rows 0..2
    float cam_pos[3];
    // film
    int32_t img_w, img_h, tile_x, tile_y;
    int32_t filter;
    float   filter_hw, filter_hh;   // half width / height
    float   g_expx, g_expy, g_edge, amplitude, mitchell_b, mitchell_c;
    int32_t sampler, spp, strata;
};

}  // namespace mpt
