/*
 * Oracle (test infrastructure): binary32 vector algebra in the reference's
 * operation order. Restates source/math/tVector.ipp and math/math.h.
 */
#ifndef MPT_OR_MATH_H
#define MPT_OR_MATH_H

#include <float.h>
#include <math.h>
#include <stdint.h>

typedef struct { float x, y, z; } v3;

static inline v3 v3_make(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 v3_load(const float* p) { return v3_make(p[0], p[1], p[2]); }
static inline void v3_store(float* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline float v3_get(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* tVector.ipp:53-151 component-wise operators */
static inline v3 v3_add(v3 a, v3 b) { return v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3_sub(v3 a, v3 b) { return v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_mul(v3 a, v3 b) { return v3_make(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 v3_divv(v3 a, v3 b) { return v3_make(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 v3_scale(v3 a, float s) { return v3_make(a.x * s, a.y * s, a.z * s); }
static inline v3 v3_adds(v3 a, float s) { return v3_make(a.x + s, a.y + s, a.z + s); }
static inline v3 v3_subs(v3 a, float s) { return v3_make(a.x - s, a.y - s, a.z - s); }
/* tVector.ipp:103-111: division by a scalar multiplies by its reciprocal */
static inline v3 v3_divs(v3 a, float s) { const float inv = 1.0f / s; return v3_scale(a, inv); }
static inline v3 v3_neg(v3 a) { return v3_make(-a.x, -a.y, -a.z); }

/* tVector.ipp:427-434: accumulation starts from 0 */
static inline float v3_dot(v3 a, v3 b) {
    float r = 0.0f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
static inline float v3_absdot(v3 a, v3 b) { return fabsf(v3_dot(a, b)); }
static inline float v3_len2(v3 a) {
    float r = 0.0f;
    r += a.x * a.x;
    r += a.y * a.y;
    r += a.z * a.z;
    return r;
}
static inline float v3_len(v3 a) { return sqrtf(v3_len2(a)); }
/* tVector.ipp:346-351 */
static inline v3 v3_normalize(v3 a) { const float inv = 1.0f / v3_len(a); return v3_scale(a, inv); }
/* tVector.ipp:453-459 */
static inline v3 v3_cross(v3 a, v3 b) {
    return v3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* tVector.ipp:256-263 */
static inline int v3_is_zero(v3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
/* tVector.ipp:301-308 */
static inline float v3_sum(v3 a) { float r = 0.0f; r += a.x; r += a.y; r += a.z; return r; }
/* tVector.ipp:395-406 */
static inline int v3_max_dim(v3 a) {
    float m = a.x; int d = 0;
    if (a.y > m) { m = a.y; d = 1; }
    if (a.z > m) { m = a.z; d = 2; }
    return d;
}
static inline float v3_max_comp(v3 a) { return v3_get(a, v3_max_dim(a)); }
/* tVector.ipp:462-465 */
static inline v3 v3_reflect(v3 v, v3 n) {
    const v3 r = v3_scale(n, 2.0f * v3_absdot(v, n));
    return v3_sub(r, v);
}
/* tVector.ipp:369-372 */
static inline v3 v3_complement(v3 a) { return v3_sub(v3_make(1.0f, 1.0f, 1.0f), a); }
/* tVector.ipp:442-450 */
static inline v3 v3_lerp(v3 a, v3 b, float t) {
    const float c = 1.0f - t;
    return v3_make(c * a.x + t * b.x, c * a.y + t * b.y, c * a.z + t * b.z);
}
static inline int v3_has_nan(v3 a) { return isnan(a.x) || isnan(a.y) || isnan(a.z); }
static inline int v3_has_inf(v3 a) { return isinf(a.x) || isinf(a.y) || isinf(a.z); }

/* std::min / std::max as TVector::min / max use them (tVector.ipp:14-31) */
static inline float std_min(float a, float b) { return (b < a) ? b : a; }
static inline float std_max(float a, float b) { return (a < b) ? b : a; }
static inline v3 v3_min(v3 a, v3 b) { return v3_make(std_min(a.x, b.x), std_min(a.y, b.y), std_min(a.z, b.z)); }
static inline v3 v3_max(v3 a, v3 b) { return v3_make(std_max(a.x, b.x), std_max(a.y, b.y), std_max(a.z, b.z)); }

/* math::min / max / clamp (math.h:116-129) */
static inline float m_min(float a, float b) { return (a < b) ? a : b; }
static inline float m_max(float a, float b) { return (a > b) ? a : b; }
static inline float m_clamp(float v, float lo, float hi) { return m_min(m_max(v, lo), hi); }
/* math::sign (math.h:110-114) */
static inline float m_sign(float v) { return (float)((0.0f < v) - (v < 0.0f)); }

/* constants (math/constant.h), converted to binary32 as static_cast<real> does */
#define OR_PI        ((float)3.1415926535897932)
#define OR_TWO_PI    ((float)6.2831853071795864)
#define OR_INV_PI    ((float)0.3183098861837906)
#define OR_RAY_EPS   ((float)0.001)
#define OR_DEG2RAD   ((float)0.0174532925199432)

/* AABB3R (math/tAabb.ipp). An empty box is (FLT_MAX, FLT_MIN): the reference
 * default-constructs max with numeric_limits<float>::min() (tAabb.ipp:13-17). */
typedef struct { v3 mn, mx; } or_aabb;
static inline or_aabb aabb_empty(void) {
    or_aabb b = {{FLT_MAX, FLT_MAX, FLT_MAX}, {FLT_MIN, FLT_MIN, FLT_MIN}};
    return b;
}
static inline or_aabb aabb_point(v3 p) { or_aabb b = {p, p}; return b; }
static inline or_aabb aabb_union(or_aabb a, or_aabb b) {
    or_aabb r = {v3_min(a.mn, b.mn), v3_max(a.mx, b.mx)};
    return r;
}
static inline or_aabb aabb_union_p(or_aabb a, v3 p) {
    or_aabb r = {v3_min(a.mn, p), v3_max(a.mx, p)};
    return r;
}
static inline or_aabb aabb_expand(or_aabb a, float s) {
    or_aabb r = {v3_subs(a.mn, s), v3_adds(a.mx, s)};
    return r;
}
static inline v3 aabb_extent(or_aabb a) { return v3_sub(a.mx, a.mn); }
static inline v3 aabb_centroid(or_aabb a) { return v3_scale(v3_add(a.mn, a.mx), 0.5f); }
static inline float aabb_area(or_aabb a) {
    const v3 e = aabb_extent(a);
    const float half = e.x * e.y + e.x * e.z + e.y * e.z;
    return 2.0f * half;
}

#endif
