// Wide-BVH traversal for gfx950: one ray per lane, no ray-tracing hardware,
// so every slab and primitive test is VALU work.
//
// Primitive tests restate the reference's exactly (operation order, edge
// swap, t == maxT accepted), so the closest-hit primitive and t agree bit for
// bit with the reference's binary BVH whenever the closest hit is unique:
//   Triangle::isIntersecting   (triangle.cpp:54-89)
//   Rectangle::isIntersecting  (rectangle.cpp:49-75)
//   Sphere::isIntersecting     (sphere.cpp:42-78, 146-171)
// Box culling follows the reference's slab test (tAabb.ipp:68-127) in form,
// (plane - o) * invDir per plane, with plane - o formed as q * step + (node
// origin - o) so that its rounding scales with the distances near the node, as
// the reference's (min - o) does, and not with the size of the coordinates; the
// far distance is widened by 4 ulps; the boxes are 8-bit quantized boxes that
// contain the tight boxes (which contain each primitive's 1e-4-padded bound),
// so it does not cull a primitive the reference would hit.
#pragma once

#include "dmath.hpp"
#include "dscene.hpp"

namespace mpt {

// Slab and primitive tests of closest-hit and of any-hit queries.
struct TravCounters {
    uint32_t boxes, prims, boxes_any, prims_any;
};


// Sphere::isIntersecting / isOccluded with _isSolutionValid: the Haines et al.
// form (q = b + sign(b) sqrt(disc), roots c / q and q, taken in that order).
__device__ __forceinline__ bool sphere_test(V3 G, float r, V3 o, V3 D, float tmin, float& tmax) {
    const V3 f = o - G;
    const float b = -dot(f, D);
    const V3 l = f + D * b;
    const float disc = r * r - absdot(l, l);
    if (disc < 0.0f) return false;
    const float sb = (b > 0.0f) ? 1.0f : -1.0f;
    const float q = b + sb * sqrtf(disc);
    const float c = absdot(f, f) - r * r;
    const float t0 = c / q, t1 = q;
    if (t0 > tmax || t1 < tmin) return false;
    float t = t0;
    if (t < tmin) {
        t = t1;
        if (t > tmax) return false;
    }
    tmax = t;
    return true;
}

// Primitive record at leaf position li: r0 = (A, id bits), r1 = e1, r2 = e2
// (r2.w: the primitive's area light, -1 none, read by shading only);
// the id's top bit marks a rectangle, or a sphere when r1.w != 0 (then
// r0 = centre, r1.x = radius).
struct PrimRec {
    float4 r0, r1, r2;
};
__device__ __forceinline__ PrimRec prim_load(const float4* __restrict__ tris, int32_t li) {
    return PrimRec{tris[3 * li + 0], tris[3 * li + 1], tris[3 * li + 2]};
}
__device__ __forceinline__ bool prim_hit(const PrimRec& pr, V3 o, V3 D, float tmin, float& tmax) {
    const float4 r0 = pr.r0, r1 = pr.r1, r2 = pr.r2;
    V3 ea = v3f4(r1), eb = v3f4(r2);
    const bool rect = (__float_as_uint(r0.w) & 0x80000000u) != 0;
    if (rect && r1.w != 0.0f) return sphere_test(v3f4(r0), r1.x, o, D, tmin, tmax);
    if (dot(D, cross(ea, eb)) > 0.0f) {
        const V3 t = ea;
        ea = eb;
        eb = t;
    }
    const V3 A = v3f4(r0);
    if (!rect) {
        const V3 T = o - A;
        const V3 Q = cross(T, ea);
        const V3 P = cross(D, eb);
        const float den = dot(P, ea);
        if (den == 0.0f) return false;
        const float inv = 1.0f / den;
        const float t = dot(Q, eb) * inv;
        const float u = dot(P, T) * inv;
        const float v = dot(Q, D) * inv;
        if (u < 0.0f || v < 0.0f || u + v > 1.0f) return false;
        if (t < tmin || t > tmax) return false;
        tmax = t;
        return true;
    }
    // rectangle: plane through vB, then projections on the unswapped edges
    const V3 N = normalize(cross(ea, eb));
    const float t = dot(N, A - o) / dot(N, D);
    if (t < tmin || t > tmax) return false;
    const V3 e1 = v3f4(r1), e2 = v3f4(r2);
    const V3 onp = (o + D * t) - A;
    const float p1 = dot(onp, normalize(e1));
    const float p2 = dot(onp, normalize(e2));
    if (p1 < 0.0f || p1 > len(e1) || p2 < 0.0f || p2 > len(e2)) return false;
    tmax = t;
    return true;
}

// Per-lane traversal stack: the first S entries live in LDS (entry-major, so
// lanes of a wave hit distinct banks), the rest spill to a per-lane global
// slice. A scene whose worst-case depth fits S never touches the spill.
template <int S>
struct LaneStack {
    __attribute__((address_space(3))) int32_t* lds;   // this block's LDS stack (LDS-typed: ds_read / ds_write)
    int32_t* spill;    // this lane's spill slice (may be null when max <= S)
    int      lane;     // thread index in block
    int      shift;    // log2(threads per block): entry e of this lane is lds[(e << shift) + lane]
    int      sp;
    __device__ __forceinline__ void push(int32_t v) {
        if (sp < S) lds[(sp << shift) + lane] = v;
        else spill[sp - S] = v;
        ++sp;
    }
    // a push the caller knows to fit in the LDS entries (sp < S)
    __device__ __forceinline__ void push_lds(int32_t v) {
        lds[(sp << shift) + lane] = v;
        ++sp;
    }
    __device__ __forceinline__ int32_t pop() {
        --sp;
        // the LDS read always (clamped index), the spill read only past S: two
        // loads of their own address spaces instead of one generic (flat) load
        int32_t v = lds[(min(sp, S - 1) << shift) + lane];
        if (sp >= S) v = spill[sp - S];
        return v;
    }
};

// Direction reciprocal for culling: zero components become +-1e-30 so that
// (lo - o) * inv stays finite (0 when the ray lies in the plane).
__device__ __forceinline__ V3 safe_inv(V3 d) {
    auto one = [](float x) {
        const float ax = fabsf(x) < 1e-30f ? copysignf(1e-30f, x) : x;
        return 1.0f / ax;
    };
    return v3(one(d.x), one(d.y), one(d.z));
}

// Traversal state of one lane: the ray's direction reciprocal, its running
// closest distance and hit, the node to visit next (-1: pop the stack). Its
// origin, tmin and direction stay in the wave's LDS ray slots (CoopLds::ray).
struct LaneRay {
    V3      inv;
    float   tmax;
    int32_t hit;
    int32_t node;
    bool    any;      // any hit (occlusion) query
    bool    pend;     // leaves of this ray wait in the wave's leaf queue
    bool    parked;   // traversal has ended; waiting for the queued leaves' round
};

// Compare-exchange of a bitonic sorting network on (distance, child) pairs.
__device__ __forceinline__ void cmpx(float& da, int32_t& ia, float& db, int32_t& ib, bool asc) {
    const bool sw = (da > db) == asc;
    const float td = sw ? db : da;
    const int32_t ti = sw ? ib : ia;
    db = sw ? da : db;
    ib = sw ? ia : ib;
    da = td;
    ia = ti;
}

// LDS of the wave-cooperative leaf phase, per wave: every lane's ray
// (origin, tmin) (direction, tmax) and best-hit key, and a queue of the hit
// leaves of all lanes (first primitive, count | owner lane << 16), up to W
// per lane.
typedef float lds_f4 __attribute__((ext_vector_type(4)));
typedef int32_t lds_i2 __attribute__((ext_vector_type(2)));
typedef int32_t lds_i4 __attribute__((ext_vector_type(4)));
struct CoopLds {
    __attribute__((address_space(3))) lds_f4* ray;
    __attribute__((address_space(3))) unsigned long long* key;
    __attribute__((address_space(3))) lds_i2* queue;
};

// Best-hit keys ordered as unsigned integers: distance first (the float's
// order-preserving integer form), then the larger leaf position, so that a
// wave-wide minimum selects what testing the leaf primitives one by one in
// leaf order selects (t == tmax is accepted, so the later of equal
// distances wins). The low word 0xFFFFFFFF means "no hit".
__device__ __forceinline__ uint32_t ord_dist(float t) {
    const uint32_t u = __float_as_uint(t);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float dist_of(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}
__device__ __forceinline__ unsigned long long hit_key(float t, int32_t li) {
    return (static_cast<unsigned long long>(ord_dist(t)) << 32) | static_cast<uint32_t>(0x7FFFFFFF - li);
}

// Exclusive prefix sum over the wave of v in [0, 2^BITS), and the wave total,
// from BITS ballots (every lane of the wave must be active).
template <int BITS>
__device__ __forceinline__ int wave_excl_scan(int v, int& total) {
    int excl = 0;
    total = 0;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const unsigned long long m = __ballot((v >> b) & 1);
        const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
        excl += static_cast<int>(below) << b;
        total += __popcll(m) << b;
    }
    return excl;
}

// One traversal step of a wave. Every lane that is still traversing pops its
// next node, slab-tests the W children, appends its hit leaves to the wave's
// leaf queue in LDS, sorts its hit interior children by entry distance,
// continues with the nearest and pushes the rest. The queued leaves are tested
// by the whole wave, one leaf per lane, in rounds of 64 (the per-ray minimum
// taken by LDS atomics on hit keys), once at least LEAF_DEFER leaves are
// queued, LEAF_PARK lanes wait on them or no lane is left traversing: a round
// then serves the leaves of several steps instead of the ~10 of one. A lane
// whose traversal has ended while leaves of its ray are still queued waits
// (parked) until that round; interior children are culled by the distance of
// the last round, which can only be larger than the current best. Returns
// true when the lane's ray has finished (any hit found for an any-hit query,
// or traversal ended and every leaf tested). qn (wave-uniform) carries the
// number of queued leaves from step to step. Must be called by all lanes of
// the wave together. All register arrays are indexed with compile-time
// indices.
template <int W, int S, int MODE>
__device__ __forceinline__ bool visit(const DevScene& sc, LaneStack<S>& st, LaneRay& r, TravCounters& cnt, bool act,
                                      const CoopLds& cl, int lane, int& qn) {
    const bool any = MODE == TRACE_ANY || (MODE == TRACE_MIXED && r.any);
    bool trav = act && !r.parked;
    bool ended = false;   // traversal of the lane's ray ends with this step
    if (trav && r.node < 0) {
        if (st.sp == 0) {
            ended = true;
            trav = false;
        } else {
            r.node = st.pop();
        }
    }
    int32_t child[W], count[W];
    float tn[W];
    bool hitc[W];
    // every lane of a wave with a traversing lane reads a node (the others read
    // node 0, which is cached, and discard it), so the slab tests need no
    // per-lane defaults and no divergent branch
    if (__ballot(trav) != 0ull) {
        // QNode<W>: origin + exponents, W children, 6 x W box bytes, W 16-bit counts
        constexpr int NWORD = W == 4 ? 16 : 28;
        constexpr int QW = 4 + W;                 // first word of the box bytes
        constexpr int CW = 4 + W + 6 * W / 4;     // first word of the counts
        const int32_t nidx = trav ? r.node : 0;
        const uint4* p = reinterpret_cast<const uint4*>(sc.nodes) + static_cast<int64_t>(nidx) * (W == 4 ? 4 : 8);
        uint32_t wd[NWORD];
#pragma unroll
        for (int k = 0; k < NWORD / 4; ++k) {
            const uint4 v = p[k];
            wd[4 * k + 0] = v.x;
            wd[4 * k + 1] = v.y;
            wd[4 * k + 2] = v.z;
            wd[4 * k + 3] = v.w;
        }
        const lds_f4 ro = cl.ray[2 * lane];   // the lane's origin and tmin (kept in LDS, not registers)
        const float dox = __uint_as_float(wd[0]) - ro.x, doy = __uint_as_float(wd[1]) - ro.y,
                    doz = __uint_as_float(wd[2]) - ro.z;
        const float sx = __uint_as_float((wd[3] & 0xFFu) << 23), sy = __uint_as_float(((wd[3] >> 8) & 0xFFu) << 23),
                    sz = __uint_as_float(((wd[3] >> 16) & 0xFFu) << 23);
        auto qb = [&](int plane, int c) {
            return static_cast<float>((wd[QW + (plane * W + c) / 4] >> (8 * (c & 3))) & 0xFFu);
        };
        cnt.boxes_any += (trav && any) ? W : 0;
        cnt.boxes += (trav && !any) ? W : 0;
#pragma unroll
        for (int c = 0; c < W; ++c) {
            child[c] = static_cast<int32_t>(wd[4 + c]);
            count[c] = static_cast<int32_t>((wd[CW + c / 2] >> (16 * (c & 1))) & 0xFFFFu);
            // plane - origin as q * step + (node origin - ray origin): the rounding is
            // relative to the distances near the node, as in the reference's
            // (min - o) * invDir, not to the size of the coordinates
            const float lox = __builtin_fmaf(qb(0, c), sx, dox), loy = __builtin_fmaf(qb(1, c), sy, doy),
                        loz = __builtin_fmaf(qb(2, c), sz, doz), hix = __builtin_fmaf(qb(3, c), sx, dox),
                        hiy = __builtin_fmaf(qb(4, c), sy, doy), hiz = __builtin_fmaf(qb(5, c), sz, doz);
            const float tx0 = lox * r.inv.x, tx1 = hix * r.inv.x;
            const float ty0 = loy * r.inv.y, ty1 = hiy * r.inv.y;
            const float tz0 = loz * r.inv.z, tz1 = hiz * r.inv.z;
            const float n = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), ro.w));
            const float f = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
            hitc[c] = trav & (child[c] >= 0) & (n <= fminf(f * 1.0000005f, r.tmax));
            tn[c] = n;
        }
    } else {
#pragma unroll
        for (int c = 0; c < W; ++c) {
            child[c] = -1;
            count[c] = 0;
            tn[c] = INFINITY;
            hitc[c] = false;
        }
    }
    // -- queue this step's hit leaves, in child order, behind the pending ones (at
    // most LEAF_DEFER - 1 are pending here, so 64 * W + LEAF_DEFER entries hold them)
    bool isl[W];
    int nl = 0;
#pragma unroll
    for (int c = 0; c < W; ++c) {
        isl[c] = hitc[c] && count[c] > 0;
        nl += isl[c] ? 1 : 0;
    }
    int ntot;
    const int excl = wave_excl_scan<W == 4 ? 3 : 4>(nl, ntot);
    {
        const int32_t tag = (any ? 0x8000 : 0) | (lane << 16);
        int k = qn + excl;
#pragma unroll
        for (int c = 0; c < W; ++c) {
            if (isl[c]) cl.queue[k] = lds_i2{child[c], count[c] | tag};
            k += isl[c] ? 1 : 0;
        }
        if (nl > 0) r.pend = true;
    }
    qn += ntot;
    // -- interior children, nearest first
    if (trav) {
        float key[W];
        int32_t kid[W];
        int nk = 0;
#pragma unroll
        for (int c = 0; c < W; ++c) {
            const bool inner = hitc[c] && count[c] == 0 && tn[c] <= r.tmax;
            key[c] = inner ? tn[c] : INFINITY;
            kid[c] = child[c];
            nk += inner ? 1 : 0;
        }
        // ascending distance: the least comparator networks for 4 and 8 keys (5 and
        // 19 compare-exchanges; Knuth, TAOCP vol. 3, 5.3.4)
        if constexpr (W == 4) {
            constexpr int net[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
#pragma unroll
            for (int e = 0; e < 5; ++e) cmpx(key[net[e][0]], kid[net[e][0]], key[net[e][1]], kid[net[e][1]], true);
        } else {
            constexpr int net[19][2] = {{0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 1}, {2, 3},
                                        {4, 5}, {6, 7}, {2, 4}, {3, 5}, {1, 4}, {3, 6}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
            for (int e = 0; e < 19; ++e) cmpx(key[net[e][0]], kid[net[e][0]], key[net[e][1]], kid[net[e][1]], true);
        }
        if (st.sp + nk - 1 > sc.max_stack) {   // bounded by the tree depth; never taken for a built tree
            r.hit = -2;
            ended = true;
        } else {
            if (st.sp + W - 1 <= S) {   // every push fits in the LDS entries: no spill test per push
#pragma unroll
                for (int k = W - 1; k >= 1; --k) {
                    if (k < nk) st.push_lds(kid[k]);
                }
            } else {
#pragma unroll
                for (int k = W - 1; k >= 1; --k) {
                    if (k < nk) st.push(kid[k]);
                }
            }
            r.node = nk > 0 ? kid[0] : -1;
            if (nk == 0 && st.sp == 0) ended = true;
        }
    }
    if (ended) r.parked = true;
    bool done = r.parked && !r.pend;   // ended with no leaf of its ray queued
    // -- the queued leaves: one leaf per lane, rounds of 64
    if (qn > 0) {
        const int waiting = __popcll(__ballot(r.parked && r.pend));
        if (qn >= LEAF_DEFER || waiting >= LEAF_PARK || __ballot(act && !r.parked) == 0ull) {
            __builtin_amdgcn_wave_barrier();
            for (int base = 0; base < qn; base += 64) {
                const int e = base + lane;
                if (e < qn) {
                    const lds_i2 q = cl.queue[e];
                    const int own = q.y >> 16, n = q.y & 0x7FFF;
                    const bool qany = MODE == TRACE_ANY || (MODE == TRACE_MIXED && (q.y & 0x8000) != 0);
                    const lds_f4 ro = cl.ray[2 * own], rd = cl.ray[2 * own + 1];
                    const V3 o = v3(ro.x, ro.y, ro.z), d = v3(rd.x, rd.y, rd.z);
                    float tmax = rd.w;
                    int32_t best = -1;
                    // two records per round: both loads in flight together
                    for (int j = 0; j < n; j += 2) {
                        const int32_t la = q.x + j;
                        const bool two = j + 1 < n;
                        const int32_t lb = two ? la + 1 : la;
                        const PrimRec A = prim_load(sc.tris, la);
                        const PrimRec B = prim_load(sc.tris, lb);
                        const uint32_t np = two ? 2u : 1u;
                        cnt.prims_any += qany ? np : 0u;
                        cnt.prims += qany ? 0u : np;
                        if (prim_hit(A, o, d, ro.w, tmax)) best = la;
                        if (two && prim_hit(B, o, d, ro.w, tmax)) best = lb;
                    }
                    if (best >= 0)
                        __hip_atomic_fetch_min(cl.key + own, hit_key(tmax, best), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WAVEFRONT);
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (r.pend) {
                const unsigned long long k = cl.key[lane];
                const uint32_t lo = static_cast<uint32_t>(k);
                if (lo != 0xFFFFFFFFu && r.hit != -2) {
                    r.hit = 0x7FFFFFFF - static_cast<int32_t>(lo);
                    r.tmax = dist_of(static_cast<uint32_t>(k >> 32));
                    cl.ray[2 * lane + 1].w = r.tmax;
                    cl.key[lane] = (static_cast<unsigned long long>(ord_dist(r.tmax)) << 32) | 0xFFFFFFFFull;
                    if (any) done = true;
                }
                r.pend = false;
            }
            if (r.parked) done = true;
            qn = 0;
        }
    }
    if (done) r.parked = false;
    return done;
}

}  // namespace mpt
