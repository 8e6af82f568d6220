/* Oracle (test infrastructure): internal scene state shared by the oracle's .c files. */
#ifndef MPT_OR_INTERNAL_H
#define MPT_OR_INTERNAL_H

#include "oracle.h"
#include "or_math.h"

typedef struct {
    int32_t type;
    v3      vA, vB, vC;      /* rectangle: vA vB vC as given, vD derived */
    v3      e1, e2;          /* triangle: eAB, eAC (triangle.cpp:29-30); rectangle: eA, eB (rectangle.cpp:29-30) */
    v3      vD;
    v3      nA, nB, nC;      /* triangle vertex normals (face normal unless set) */
    v3      uvwA, uvwB, uvwC, uvwD;
    float   radius;          /* sphere: centre in vA */
    int32_t mat;
    int32_t light;           /* area light bound to this primitive, or -1 */
} or_prim;

typedef struct {
    or_aabb bound;
    int64_t a, b;            /* leaf: first ordered primitive, count; interior: second child, axis */
    int32_t leaf;
} or_node;

/* Distribution1D (distribution1D.cpp:10-41): cdf has n + 1 entries. */
typedef struct {
    float*  cdf;
    int64_t n;
    float   delta;
} or_dist1;

typedef struct {
    int32_t prim;
    v3      radiance;        /* emit radiance (singleAreaLight.cpp:20-37) */
    int32_t back_face;
    /* environment light (environmentLight.cpp:16-57) */
    int32_t    type;
    or_texture env;
    or_dist1   marginal;     /* over rows (v) */
    or_dist1*  conditional;  /* per row, over columns (u) */
    float      background_flux, radius, flux;
} or_light_state;

struct or_scene {
    or_scene_desc desc;
    int64_t       nprims;
    or_prim*      prims;
    int32_t       nmats;
    or_material*  mats;
    int32_t       nlights;
    or_light_state* lights;
    float*        power_cdf;  /* nlights + 1 (power cluster) */
    int32_t       env_light;  /* the environment light (background sphere), -1 none */
    or_node*      nodes;
    int64_t       nnodes;
    int32_t*      ordered;    /* ordered primitive indices (BVH leaf order) */
    float         film_to_camera[16];
    float         camera_to_world[16];
    v3            cam_pos;
    int           strata_base;
};

/* geometry (or_bvh.c) */
or_aabb or_prim_bound(const or_prim* p);
int  or_prim_intersect(const or_prim* p, v3 o, v3 d, float minT, float* maxT);
void or_build_bvh(or_scene* s);
int  or_closest(const or_scene* s, v3 o, v3 d, float minT, float* maxT, int32_t* prim);
int  or_occluded(const or_scene* s, v3 o, v3 d, float minT, float maxT);
int  or_closest_brute(const or_scene* s, v3 o, v3 d, float minT, float* maxT, int32_t* prim);
int  or_occluded_brute(const or_scene* s, v3 o, v3 d, float minT, float maxT);

#endif
