/*
 * Oracle (test infrastructure): primitives, the reference's binary BVH build
 * and its traversal.
 *
 * The build depends on two libstdc++ algorithms (GCC 11.4, the toolchain the
 * reference is compiled with here): std::partition (bvhBuilder.cpp:336-347)
 * and std::nth_element (bvhBuilder.cpp:231-238). Both are restated below
 * from libstdc++'s bits/stl_algo.h / bits/stl_heap.h (GCC 11): Hoare-style
 * two-sided partition; introselect with median-of-three pivot, heap-select
 * fallback and insertion sort on ranges of at most three.
 */
#include "or_internal.h"

#include <stdlib.h>
#include <string.h>

/* ---- primitives ------------------------------------------------------------ */

/* Triangle::evaluateBound (triangle.cpp:45-52), Rectangle::evaluateBound (rectangle.cpp:39-47),
 * Sphere::evaluateBound (sphere.cpp:33-40) */
or_aabb or_prim_bound(const or_prim* p) {
    if (p->type == OR_PRIM_SPHERE) {
        const or_aabb s = {v3_subs(p->vA, p->radius), v3_adds(p->vA, p->radius)};
        return aabb_expand(s, 0.0001f);
    }
    or_aabb b = aabb_point(p->vA);
    b = aabb_union_p(b, p->vB);
    b = aabb_union_p(b, p->vC);
    if (p->type == OR_PRIM_RECTANGLE) b = aabb_union_p(b, p->vD);
    return aabb_expand(b, 0.0001f);
}

/* Triangle::isIntersecting (triangle.cpp:54-89): double-sided Moller-Trumbore
 * with the edge swap for back-facing rays; t == maxT is accepted. */
static int tri_intersect(const or_prim* p, v3 o, v3 D, float minT, float* maxT) {
    v3 eAB = p->e1, eAC = p->e2;
    if (v3_dot(D, v3_cross(eAB, eAC)) > 0.0f) {
        const v3 tmp = eAB; eAB = eAC; eAC = tmp;
    }
    const v3 T = v3_sub(o, p->vA);
    const v3 Q = v3_cross(T, eAB);
    const v3 P = v3_cross(D, eAC);
    const float den = v3_dot(P, eAB);
    if (den == 0.0f) return 0;
    const float inv = 1.0f / den;
    const float t = v3_dot(Q, eAC) * inv;
    const float u = v3_dot(P, T) * inv;
    const float v = v3_dot(Q, D) * inv;
    if (u < 0.0f || v < 0.0f || u + v > 1.0f) return 0;
    if (t < minT || t > *maxT) return 0;
    *maxT = t;
    return 1;
}

/* Rectangle::isIntersecting (rectangle.cpp:49-75) */
static int rect_intersect(const or_prim* p, v3 o, v3 D, float minT, float* maxT) {
    v3 eA = p->e1, eB = p->e2;
    if (v3_dot(D, v3_cross(eA, eB)) > 0.0f) {
        const v3 tmp = eA; eA = eB; eB = tmp;
    }
    const v3 normal = v3_normalize(v3_cross(eA, eB));
    const float t = v3_dot(normal, v3_sub(p->vB, o)) / v3_dot(normal, D);
    if (t < minT || t > *maxT) return 0;
    const v3 onPlane = v3_sub(v3_add(o, v3_scale(D, t)), p->vB);
    const float p1 = v3_dot(onPlane, v3_normalize(p->e1));
    const float p2 = v3_dot(onPlane, v3_normalize(p->e2));
    if (p1 < 0.0f || p1 > v3_len(p->e1) || p2 < 0.0f || p2 > v3_len(p->e2)) return 0;
    *maxT = t;
    return 1;
}

/* Sphere::isIntersecting / isOccluded (sphere.cpp:42-107) with _isSolutionValid (sphere.cpp:146-171) */
static int sphere_intersect(const or_prim* p, v3 o, v3 d, float minT, float* maxT) {
    const v3 G = p->vA;
    const v3 f = v3_sub(o, G);
    const float b = -v3_dot(f, d);
    const float r = p->radius;
    const v3 l = v3_add(f, v3_scale(d, b));
    const float disc = r * r - v3_absdot(l, l);
    if (disc < 0.0f) return 0;
    const float signB = (b > 0.0f) ? 1.0f : -1.0f;
    const float q = b + signB * sqrtf(disc);
    const float c = v3_absdot(f, f) - r * r;
    const float t0 = c / q, t1 = q;
    if (t0 > *maxT || t1 < minT) return 0;
    float t = t0;
    if (t < minT) {
        t = t1;
        if (t > *maxT) return 0;
    }
    *maxT = t;
    return 1;
}

int or_prim_intersect(const or_prim* p, v3 o, v3 d, float minT, float* maxT) {
    if (p->type == OR_PRIM_SPHERE) return sphere_intersect(p, o, d, minT, maxT);
    return p->type == OR_PRIM_RECTANGLE ? rect_intersect(p, o, d, minT, maxT) : tri_intersect(p, o, d, minT, maxT);
}

/* ---- libstdc++ algorithms on the build's bound-info records ---------------- */

typedef struct {
    or_aabb bound;
    v3      centroid;
    int64_t index;
} bound_info;

static inline void bi_swap(bound_info* a, bound_info* b) { bound_info t = *a; *a = *b; *b = t; }

typedef struct { int axis; } cent_less;  /* comparator: centroid[axis] < */
static inline int cl(const cent_less* c, const bound_info* a, const bound_info* b) {
    return v3_get(a->centroid, c->axis) < v3_get(b->centroid, c->axis);
}

/* std::__unguarded_linear_insert / std::__insertion_sort */
static void insertion_sort(bound_info* first, bound_info* last, const cent_less* c) {
    if (first == last) return;
    for (bound_info* i = first + 1; i != last; ++i) {
        if (cl(c, i, first)) {
            bound_info val = *i;
            memmove(first + 1, first, (size_t)(i - first) * sizeof(bound_info));
            *first = val;
        } else {
            bound_info val = *i;
            bound_info* l = i;
            bound_info* next = i - 1;
            while (cl(c, &val, next)) {
                *l = *next;
                l = next;
                --next;
            }
            *l = val;
        }
    }
}

/* std::__push_heap, std::__adjust_heap, std::__make_heap, std::__pop_heap */
static void push_heap(bound_info* first, long hole, long top, bound_info value, const cent_less* c) {
    long parent = (hole - 1) / 2;
    while (hole > top && cl(c, first + parent, &value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}
static void adjust_heap(bound_info* first, long hole, long len, bound_info value, const cent_less* c) {
    const long top = hole;
    long second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (cl(c, first + second, first + (second - 1))) second--;
        first[hole] = first[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        first[hole] = first[second - 1];
        hole = second - 1;
    }
    push_heap(first, hole, top, value, c);
}
static void make_heap(bound_info* first, bound_info* last, const cent_less* c) {
    const long len = last - first;
    if (len < 2) return;
    long parent = (len - 2) / 2;
    while (1) {
        bound_info value = first[parent];
        adjust_heap(first, parent, len, value, c);
        if (parent == 0) return;
        parent--;
    }
}
static void pop_heap(bound_info* first, bound_info* last, bound_info* result, const cent_less* c) {
    bound_info value = *result;
    *result = *first;
    adjust_heap(first, 0, last - first, value, c);
}
static void heap_select(bound_info* first, bound_info* middle, bound_info* last, const cent_less* c) {
    make_heap(first, middle, c);
    for (bound_info* i = middle; i < last; ++i)
        if (cl(c, i, first)) pop_heap(first, middle, i, c);
}

/* std::__move_median_to_first, std::__unguarded_partition(_pivot) */
static void move_median_to_first(bound_info* result, bound_info* a, bound_info* b, bound_info* cc,
                                 const cent_less* c) {
    if (cl(c, a, b)) {
        if (cl(c, b, cc)) bi_swap(result, b);
        else if (cl(c, a, cc)) bi_swap(result, cc);
        else bi_swap(result, a);
    } else if (cl(c, a, cc)) bi_swap(result, a);
    else if (cl(c, b, cc)) bi_swap(result, cc);
    else bi_swap(result, b);
}
static bound_info* unguarded_partition(bound_info* first, bound_info* last, bound_info* pivot, const cent_less* c) {
    while (1) {
        while (cl(c, first, pivot)) ++first;
        --last;
        while (cl(c, pivot, last)) --last;
        if (!(first < last)) return first;
        bi_swap(first, last);
        ++first;
    }
}
static int lg(long n) { int k = 0; while (n > 1) { n >>= 1; ++k; } return k; }

/* std::nth_element -> std::__introselect */
static void nth_element(bound_info* first, bound_info* nth, bound_info* last, const cent_less* c) {
    if (first == last || nth == last) return;
    long depth = 2 * lg(last - first);
    while (last - first > 3) {
        if (depth == 0) {
            heap_select(first, nth + 1, last, c);
            bi_swap(first, nth);
            return;
        }
        --depth;
        bound_info* mid = first + (last - first) / 2;
        move_median_to_first(first, first + 1, mid, last - 1, c);
        bound_info* cut = unguarded_partition(first + 1, last, first, c);
        if (cut <= nth) first = cut;
        else last = cut;
    }
    insertion_sort(first, last, c);
}

/* ---- BvhBuilder (bvhBuilder.cpp) ------------------------------------------ */

typedef struct {
    or_scene*   s;
    bound_info* infos;
    int64_t     nordered;
    or_node*    nodes;     /* pre-order, filled during the recursion */
    int64_t     nnodes;
    int         sah;
} builder;

/* bucket index of a centroid (bvhBuilder.cpp:281-287, 340-344) */
static inline size_t sah_bucket(float c, float mn, float invExtent) {
    const float offset = (c - mn) * invExtent;
    const size_t idx = (size_t)(offset * (float)12);
    return idx == 12 ? 11 : idx;
}

/* _canSplitWithEqual (bvhBuilder.cpp:212-249): returns the split position */
static int64_t split_equal(bound_info* b, int64_t n, int axis) {
    const int64_t size = (n + 1) / 2;
    cent_less c = {axis};
    nth_element(b, b + size, b + n, &c);
    return size;
}

/* libstdc++ std::__partition (bidirectional form) with the bucket predicate */
static int64_t partition_bucket(bound_info* first0, int64_t n, int axis, float mn, float inv, size_t best) {
    bound_info* first = first0;
    bound_info* last  = first0 + n;
    while (1) {
        while (1) {
            if (first == last) return first - first0;
            else if (sah_bucket(v3_get(first->centroid, axis), mn, inv) <= best) ++first;
            else break;
        }
        --last;
        while (1) {
            if (first == last) return first - first0;
            else if (!(sah_bucket(v3_get(last->centroid, axis), mn, inv) <= best)) --last;
            else break;
        }
        bi_swap(first, last);
        ++first;
    }
}

/* _canSplitWithSah (bvhBuilder.cpp:251-358): returns the split position, or -1
 * when no split beats the leaf cost */
static int64_t split_sah(bound_info* b, int64_t n, int axis, or_aabb nodeBound, or_aabb centroidBound) {
    const float mn  = v3_get(centroidBound.mn, axis);
    const float inv = 1.0f / v3_get(aabb_extent(centroidBound), axis);
    if (n <= 2) return split_equal(b, n, axis);

    or_aabb bucketBound[12];
    size_t  bucketCount[12];
    for (int i = 0; i < 12; ++i) { bucketBound[i] = aabb_empty(); bucketCount[i] = 0; }
    for (int64_t i = 0; i < n; ++i) {
        const size_t k = sah_bucket(v3_get(b[i].centroid, axis), mn, inv);
        bucketBound[k] = aabb_union(bucketBound[k], b[i].bound);
        bucketCount[k] += 1;
    }
    const float noSplitCost = 1.0f * (float)n;
    const float invArea = 1.0f / aabb_area(nodeBound);
    size_t best = (size_t)-1;
    float  bestCost = FLT_MAX;
    for (size_t split = 0; split < 11; ++split) {
        size_t  ca = 0, cb = 0;
        or_aabb ba = aabb_empty(), bb = aabb_empty();
        for (size_t i = 0; i <= split; ++i) { ba = aabb_union(ba, bucketBound[i]); ca += bucketCount[i]; }
        for (size_t i = split + 1; i < 12; ++i) { bb = aabb_union(bb, bucketBound[i]); cb += bucketCount[i]; }
        const float pa = aabb_area(ba) * invArea;
        const float pb = aabb_area(bb) * invArea;
        const float cost = 0.125f + 1.0f * (pa * (float)ca + pb * (float)cb);
        if (cost < bestCost) { bestCost = cost; best = split; }
    }
    if (bestCost >= noSplitCost) return -1;
    return partition_bucket(b, n, axis, mn, inv, best);
}

static void make_leaf(builder* B, int64_t node, int64_t begin, int64_t n, or_aabb bound) {
    B->nodes[node].bound = bound;
    B->nodes[node].leaf = 1;
    B->nodes[node].a = B->nordered;
    B->nodes[node].b = n;
    for (int64_t i = 0; i < n; ++i) B->s->ordered[B->nordered++] = (int32_t)B->infos[begin + i].index;
}

/* _buildBinaryNodesRecursively (bvhBuilder.cpp:59-169) fused with the
 * pre-order flattening of _buildLinearNodesRecursively (bvhBuilder.cpp:171-210):
 * first child at index + 1, second child index stored with the split axis. */
static void build_rec(builder* B, int64_t begin, int64_t n) {
    const int64_t node = B->nnodes++;
    bound_info* b = B->infos + begin;
    or_aabb nodeBound = aabb_empty();
    for (int64_t i = 0; i < n; ++i) nodeBound = aabb_union(nodeBound, b[i].bound);
    if (n <= 1) { make_leaf(B, node, begin, n, nodeBound); return; }
    or_aabb cb = aabb_empty();
    for (int64_t i = 0; i < n; ++i) cb = aabb_union_p(cb, b[i].centroid);
    const int axis = v3_max_dim(aabb_extent(cb));
    if (v3_get(cb.mn, axis) == v3_get(cb.mx, axis)) { make_leaf(B, node, begin, n, nodeBound); return; }
    const int64_t mid = B->sah ? split_sah(b, n, axis, nodeBound, cb) : split_equal(b, n, axis);
    if (mid < 0) { make_leaf(B, node, begin, n, nodeBound); return; }
    B->nodes[node].bound = nodeBound;
    B->nodes[node].leaf = 0;
    B->nodes[node].b = axis;
    build_rec(B, begin, mid);
    B->nodes[node].a = B->nnodes;
    build_rec(B, begin + mid, n - mid);
}

void or_build_bvh(or_scene* s) {
    builder B;
    memset(&B, 0, sizeof(B));
    B.s = s;
    B.sah = s->desc.split_sah;
    B.infos = (bound_info*)malloc(sizeof(bound_info) * (size_t)(s->nprims > 0 ? s->nprims : 1));
    for (int64_t i = 0; i < s->nprims; ++i) {
        B.infos[i].bound = or_prim_bound(&s->prims[i]);
        B.infos[i].centroid = aabb_centroid(B.infos[i].bound);
        B.infos[i].index = i;
    }
    s->ordered = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s->nprims > 0 ? s->nprims : 1));
    s->nodes = (or_node*)malloc(sizeof(or_node) * (size_t)(2 * s->nprims + 1));
    B.nodes = s->nodes;
    build_rec(&B, 0, s->nprims);
    s->nnodes = B.nnodes;
    free(B.infos);
}

/* ---- traversal (bvhAccelerator.cpp:40-156, tAabb.ipp:68-127) --------------- */

static inline int slab(const or_aabb* bb, v3 o, v3 inv, float rMin, float rMax) {
    float minT = rMin, maxT = rMax;
    const v3 nearT = v3_mul(v3_sub(bb->mn, o), inv);
    const v3 farT  = v3_mul(v3_sub(bb->mx, o), inv);
    if (inv.x > 0.0f) { minT = m_max(minT, nearT.x); maxT = m_min(maxT, farT.x); }
    else              { minT = m_max(minT, farT.x);  maxT = m_min(maxT, nearT.x); }
    if (inv.y > 0.0f) { minT = m_max(minT, nearT.y); maxT = m_min(maxT, farT.y); }
    else              { minT = m_max(minT, farT.y);  maxT = m_min(maxT, nearT.y); }
    if (inv.z > 0.0f) { minT = m_max(minT, nearT.z); maxT = m_min(maxT, farT.z); }
    else              { minT = m_max(minT, farT.z);  maxT = m_min(maxT, nearT.z); }
    return !(minT > maxT);
}

static int traverse(const or_scene* s, v3 o, v3 d, float minT, float* maxT, int32_t* prim, int any) {
    if (s->nnodes == 0) return 0;
    const v3 inv = v3_make(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    const int neg[3] = {inv.x < 0.0f, inv.y < 0.0f, inv.z < 0.0f};
    int64_t stack[64];
    int sp = 0, result = 0;
    int64_t cur = 0;
    while (1) {
        const or_node* n = &s->nodes[cur];
        if (slab(&n->bound, o, inv, minT, *maxT)) {
            if (n->leaf) {
                for (int64_t i = 0; i < n->b; ++i) {
                    const int32_t pi = s->ordered[n->a + i];
                    if (or_prim_intersect(&s->prims[pi], o, d, minT, maxT)) {
                        if (any) return 1;
                        *prim = pi;
                        result = 1;
                    }
                }
                if (sp == 0) break;
                cur = stack[--sp];
            } else {
                if (neg[n->b]) { stack[sp++] = cur + 1; cur = n->a; }
                else           { stack[sp++] = n->a;    cur = cur + 1; }
            }
        } else {
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    return result;
}

int or_closest(const or_scene* s, v3 o, v3 d, float minT, float* maxT, int32_t* prim) {
    return traverse(s, o, d, minT, maxT, prim, 0);
}
int or_occluded(const or_scene* s, v3 o, v3 d, float minT, float maxT) {
    int32_t dummy;
    return traverse(s, o, d, minT, &maxT, &dummy, 1);
}

/* BruteForceAccelerator (bruteForceAccelerator.cpp:26-45): all primitives in order */
int or_closest_brute(const or_scene* s, v3 o, v3 d, float minT, float* maxT, int32_t* prim) {
    int result = 0;
    for (int64_t i = 0; i < s->nprims; ++i)
        if (or_prim_intersect(&s->prims[i], o, d, minT, maxT)) { *prim = (int32_t)i; result = 1; }
    return result;
}
int or_occluded_brute(const or_scene* s, v3 o, v3 d, float minT, float maxT) {
    for (int64_t i = 0; i < s->nprims; ++i)
        if (or_prim_intersect(&s->prims[i], o, d, minT, &maxT)) return 1;
    return 0;
}

int64_t or_bvh_export(const or_scene* s, float* bounds, int32_t* is_leaf, int64_t* a, int64_t* b,
                      int32_t* ordered) {
    if (bounds) {
        for (int64_t i = 0; i < s->nnodes; ++i) {
            v3_store(bounds + 6 * i, s->nodes[i].bound.mn);
            v3_store(bounds + 6 * i + 3, s->nodes[i].bound.mx);
            is_leaf[i] = s->nodes[i].leaf;
            a[i] = s->nodes[i].a;
            b[i] = s->nodes[i].b;
        }
        memcpy(ordered, s->ordered, sizeof(int32_t) * (size_t)s->nprims);
    }
    return s->nnodes;
}

void or_trace(const or_scene* s, const float* rays, int64_t n, int any_hit, int use_bvh, int32_t* out_prim,
              float* out_t) {
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        const v3 o = v3_load(r), d = v3_load(r + 4);
        float maxT = r[7];
        if (any_hit) {
            out_prim[i] = use_bvh ? or_occluded(s, o, d, r[3], maxT) : or_occluded_brute(s, o, d, r[3], maxT);
            out_t[i] = 0.0f;
        } else {
            int32_t prim = -1;
            if (use_bvh) or_closest(s, o, d, r[3], &maxT, &prim);
            else or_closest_brute(s, o, d, r[3], &maxT, &prim);
            out_prim[i] = prim;
            out_t[i] = maxT;
        }
    }
}

/* Number of primitives whose hit distance equals the closest hit's exactly.
 * With the reference's "t == maxT accepted" rule (triangle.cpp:81) the winner
 * among such primitives depends on the test order, so the parity check counts
 * these rays separately. Walks every node whose box the ray overlaps within
 * [minT, t_closest] and counts the primitives hit at exactly t_closest. */
static int64_t count_at(const or_scene* s, v3 o, v3 d, float minT, float tbest) {
    if (s->nnodes == 0) return 0;
    const v3 inv = v3_make(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int64_t stack[128];
    int sp = 0;
    int64_t cur = 0, count = 0;
    while (1) {
        const or_node* n = &s->nodes[cur];
        if (slab(&n->bound, o, inv, minT, tbest)) {
            if (n->leaf) {
                for (int64_t i = 0; i < n->b; ++i) {
                    float t = tbest;
                    if (or_prim_intersect(&s->prims[s->ordered[n->a + i]], o, d, minT, &t) && t == tbest) count++;
                }
            } else if (sp < 127) {
                stack[sp++] = n->a;
                cur = cur + 1;
                continue;
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return count;
}

void or_trace_ties(const or_scene* s, const float* rays, int64_t n, int32_t* out_ntie) {
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        const v3 o = v3_load(r), d = v3_load(r + 4);
        float best = r[7];
        int32_t prim = -1;
        or_closest(s, o, d, r[3], &best, &prim);
        out_ntie[i] = prim >= 0 ? (int32_t)count_at(s, o, d, r[3], best) : 0;
    }
}
