// gfx950 kernels of the wavefront path tracer:
//   k_trace     persistent wide-BVH traversal (closest or any hit)
//   k_camera    primary rays of one sample pass
//   k_shade     resolve the previous vertex's NEE, shade the new hit, emit the
//               next vertex's shadow / MIS-BSDF / continuation rays
//   k_film      per-tile filter splat of one pass into the linear film
// The bounce loop of PtEstimator::estimate (reference core/estimator/
// ptEstimator.cpp:25-123) is unrolled across launches: each launch moves every
// live path by one step, and live paths are compacted by wave ballots.
#include "adjoint.hpp"
#include "kernels.hpp"
#include "shade.hpp"
#include "traverse.hpp"

namespace mpt {

constexpr int WAVE = 64;

// Philox slots within one bounce b: counter (pixel, sample, 2b + slot / 4, 0);
// the lobe choices of two-lobe samples (plastic, glass) at slots 8 and 9:
// counter (pixel, sample, 2b, 1).
enum {
    SL_PICK = 0, SL_LPOS0 = 1, SL_LPOS1 = 2, SL_RR = 3, SL_MIS0 = 4, SL_MIS1 = 5, SL_CONT0 = 6, SL_CONT1 = 7,
    SL_MISC = 8, SL_CONTC = 9
};

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// ---- traversal -------------------------------------------------------------------

// Persistent traversal with per-lane prefetch: every lane holds the next ray it
// will trace (reserved with one atomic per wave, its loads issued long before
// they are used), and a lane whose ray has finished takes it at once instead of
// idling until the whole wave's 64 rays are done, so lanes stay busy despite the
// very uneven cost of incoherent rays. Reservations happen between blocks of up
// to STEPS node visits. A mixed launch (MODE TRACE_MIXED) traces the closest-hit
// sets and the shadow (any-hit) set from one work counter, so one ramp and one
// tail of stragglers serve both kinds of query.
template <int W, int S, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) k_trace(TraceArgs a) {
    // node visits between hand-outs of new rays (1080p frame, ms per launch, steps
    // 2 / 3 / 4 / 6 / 8: BVH4 2.04 / 1.97 / 1.95 / 1.93 / 1.96, BVH8 2.25 / 2.21 /
    // 2.22 / 2.31 / 2.44)
    constexpr int STEPS = W == 4 ? 6 : 3;
    extern __shared__ int32_t lds_stack[];
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int64_t gslot = static_cast<int64_t>(blockIdx.x) * blockDim.x + tid;
    LaneStack<S> st;
    st.lds = (__attribute__((address_space(3))) int32_t*)lds_stack;
    st.lane = tid;
    st.shift = 31 - __builtin_clz(blockDim.x);   // blocks of 128 or 256 threads
    st.spill = a.spill ? a.spill + gslot * (a.sc.max_stack - S) : nullptr;
    st.sp = 0;
    // after the stacks: per thread two float4 (ray) and a key, then per wave a
    // queue of 64 * W + LEAF_DEFER leaves
    CoopLds cl;
    {
        const int wv = tid / WAVE, nt = static_cast<int>(blockDim.x);
        char* base = reinterpret_cast<char*>(lds_stack + S * nt);
        cl.ray = (__attribute__((address_space(3))) lds_f4*)(reinterpret_cast<lds_f4*>(base) + 2 * WAVE * wv);
        cl.key = (__attribute__((address_space(3))) unsigned long long*)(
                     reinterpret_cast<unsigned long long*>(base + 32 * nt) + WAVE * wv);
        cl.queue = (__attribute__((address_space(3))) lds_i2*)(reinterpret_cast<lds_i2*>(base + 40 * nt) +
                                                               (WAVE * W + LEAF_DEFER) * wv);
    }
    const int64_t n0 = a.count ? static_cast<int64_t>(*a.count) : a.n_host;
    const int64_t n = n0 + (a.lcount ? static_cast<int64_t>(*a.lcount) : 0);
    TravCounters cnt{0, 0, 0, 0};
    uint32_t rays = 0, rays2 = 0, rays3 = 0, errors = 0;
    // ray i: for i < n0 the first set's ray i (set 0: dir -> hit), then list ray i - n0,
    // whose lq entry names its set (1: MIS-BSDF -> hit2, 2: shadow -> hit3) and slot.
    // Ray indices are 32-bit and the work counter unsigned 32-bit: the host caps n at
    // 2^31 - 2^24, so the counter (n plus at most one chunk of TRACE_CHUNK <= 512 per
    // wave, at most 2^15 waves) stays below 2^32.
    // rays per reservation: TRACE_CHUNK for large launches; a smaller launch is cut
    // so that every wave of the grid gets about TRACE_SPREAD reservations (at least
    // TRACE_CHUNK_MIN rays each), else a few waves would take all of it
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / WAVE;
    const int32_t chunk = static_cast<int32_t>(
        min<int64_t>(TRACE_CHUNK, max<int64_t>(TRACE_CHUNK_MIN, n / (TRACE_SPREAD * nwaves))));
    // a ray's set from its index and its destination word (i for the first set, else lq)
    auto set_of = [&](int32_t i, int32_t dst) { return i < n0 ? 0 : (dst < 0 ? 1 : 2); };
    auto is_any = [&](int set) { return MODE == TRACE_ANY || (MODE == TRACE_MIXED && set == 2); };
    auto put = [&](int set, int32_t q, int2 out) {
        if (set == 0) a.hit[q] = out;
        else if (set == 1) a.hit2[q] = out;
        else a.hit3[q] = out;
    };
    const bool no_tree = a.sc.nnodes == 0;
    LaneRay r;
    r.pend = false;
    r.parked = false;
    int qn = 0;                  // wave-uniform: leaves in the wave's queue
    int32_t rid = -1;            // ray of this lane, -1 = idle
    int32_t pid = -1;            // prefetched ray (loads in flight or landed), -1 = none
    int32_t rdst = 0, pdst = 0;  // their destination words: i, or the list ray's lq entry
    float4 po = make_float4(0.f, 0.f, 0.f, 0.f), pd = po;
    int32_t proot = 0;           // its BVH root
    bool exhausted = false;      // wave-uniform: the work counter has run out and the pool is empty
    bool counter_done = false;   // wave-uniform: the work counter has run out
    int32_t pool = 0, pool_end = 0;   // wave-uniform: rays reserved by this wave and not yet handed out
    // a finished lane takes its prefetched ray at once; invalid rays are answered here
    auto install = [&]() {
        while (rid < 0 && pid >= 0) {
            const int32_t i = pid;
            pid = -1;
            const int set = set_of(i, pdst);
            if (pd.w < 0.0f || no_tree) {
                put(set, pdst & 0x7FFFFFFF, make_int2(is_any(set) ? 0 : -1, __float_as_int(pd.w)));
            } else {
                rays += set == 0 ? 1 : 0;
                rays2 += set == 1 ? 1 : 0;
                rays3 += set == 2 ? 1 : 0;
                rid = i;
                rdst = pdst;
                r.any = is_any(set);
                r.inv = safe_inv(v3f4(pd));
                r.tmax = pd.w;
                cl.ray[2 * lane] = lds_f4{po.x, po.y, po.z, po.w};
                cl.ray[2 * lane + 1] = lds_f4{pd.x, pd.y, pd.z, pd.w};
                cl.key[lane] = (static_cast<unsigned long long>(ord_dist(pd.w)) << 32) | 0xFFFFFFFFull;
                r.pend = false;
                r.parked = false;
                r.hit = -1;
                r.node = proot;
                st.sp = 0;
            }
        }
    };
    while (true) {
        // -- hand out rays to lanes without a prefetched one and start loading them:
        // from the wave's pool, refilled `chunk` rays at a time by one atomic (one
        // atomic per refill of a few lanes serializes every wave of the GPU on the
        // counter: 3.29 -> 1.93 ms per launch of the 1080p frame)
        if (!exhausted && __ballot(pid < 0) != 0ull) {
          for (int pass = 0; pass < 2; ++pass) {
            if (pool == pool_end && !counter_done) {
                uint32_t b = 0;
                if (lane == 0) b = atomicAdd(reinterpret_cast<uint32_t*>(a.work), static_cast<uint32_t>(chunk));
                b = __shfl(b, 0, WAVE);
                if (static_cast<int64_t>(b) >= n) {
                    counter_done = true;
                } else {
                    pool = static_cast<int32_t>(b);
                    pool_end = static_cast<int32_t>(min(static_cast<int64_t>(b) + chunk, n));
                }
            }
            const unsigned long long need = __ballot(pid < 0);
            const int take = min(__popcll(need), pool_end - pool);
            const int rank = static_cast<int>(__popcll(need & ((1ull << lane) - 1ull)));
            const int32_t base = pool;
            pool += take;
            if (pid < 0 && rank < take) {
                const int32_t i = base + rank;
                pid = i;
                if (i < n0) {
                    const int64_t o = static_cast<int64_t>(i) * a.stride;
                    po = a.org[o];
                    pd = a.dir[o];
                    proot = a.root ? a.root[i] : 0;
                    pdst = i;
                } else {   // the list's loads do not depend on each other
                    const int32_t k = i - static_cast<int32_t>(n0);
                    po = a.lorg[k];
                    pd = a.ldir[k];
                    proot = a.lroot ? a.lroot[k] : 0;
                    pdst = a.lq[k];
                }
            }
            if (__ballot(pid < 0) == 0ull || (counter_done && pool == pool_end)) break;
          }
          if (counter_done && pool == pool_end) exhausted = true;
        }
        install();
        if (__ballot(rid >= 0) == 0ull) {
            if (exhausted && __ballot(pid >= 0) == 0ull) break;
            continue;
        }
        // -- traversal steps
        for (int s = 0; s < STEPS; ++s) {
            const bool fin = visit<W, S, MODE>(a.sc, st, r, cnt, rid >= 0, cl, lane, qn);
            if (rid >= 0 && fin) {
                if (r.hit == -2) {
                    errors += 1;
                    r.hit = -1;
                }
                const int set = set_of(rid, rdst);
                const int32_t q = rdst & 0x7FFFFFFF;
                int2 out;
                if (is_any(set)) {
                    out = make_int2(r.hit >= 0 ? 1 : 0, 0);
                } else {
                    int32_t idx = r.hit;
                    if (r.hit >= 0 && a.want_prim)
                        idx = static_cast<int32_t>(__float_as_uint(a.sc.tris[3 * r.hit].w) & 0x7FFFFFFFu);
                    // a MIS-BSDF ray only counts when it reaches a light: its record was
                    // just read by the leaf round, so the emitter check costs a cache hit
                    // here instead of a gather in shading
                    if (set == 1 && a.mis_emitters && r.hit >= 0 && __float_as_int(a.sc.tris[3 * r.hit + 2].w) < 0)
                        idx = -2;
                    out = make_int2(r.hit >= 0 ? idx : -1, __float_as_int(r.tmax));
                }
                put(set, q, out);
                rid = -1;
                install();
            }
            if (__ballot(rid >= 0) == 0ull) break;
        }
    }
    // one atomic per wave for the counters
    const unsigned long long sr = wave_sum(rays), sr2 = wave_sum(rays2), sr3 = wave_sum(rays3),
                             sb = wave_sum(cnt.boxes), sp = wave_sum(cnt.prims), sba = wave_sum(cnt.boxes_any),
                             spa = wave_sum(cnt.prims_any), se = wave_sum(errors);
    if (lane == 0 && a.stats) {
        atomicAdd(&a.stats[a.stat_slot], sr);
        if (sr2) atomicAdd(&a.stats[ST_MIS_RAYS], sr2);
        if (sr3) atomicAdd(&a.stats[ST_SHADOW_RAYS], sr3);
        if (sb) atomicAdd(&a.stats[ST_BOXES], sb);
        if (sp) atomicAdd(&a.stats[ST_PRIMS], sp);
        if (sba) atomicAdd(&a.stats[ST_BOXES_ANY], sba);
        if (spa) atomicAdd(&a.stats[ST_PRIMS_ANY], spa);
        if (se) atomicAdd(&a.stats[ST_ERRORS], se);
    }
}

template __global__ void k_trace<4, 24, TRACE_CLOSEST>(TraceArgs);
template __global__ void k_trace<4, 24, TRACE_ANY>(TraceArgs);
template __global__ void k_trace<4, 24, TRACE_MIXED>(TraceArgs);
template __global__ void k_trace<8, 24, TRACE_CLOSEST>(TraceArgs);
template __global__ void k_trace<8, 24, TRACE_ANY>(TraceArgs);
template __global__ void k_trace<8, 24, TRACE_MIXED>(TraceArgs);
template __global__ void k_trace<4, 16, TRACE_CLOSEST>(TraceArgs);
template __global__ void k_trace<4, 16, TRACE_ANY>(TraceArgs);
template __global__ void k_trace<4, 16, TRACE_MIXED>(TraceArgs);
template __global__ void k_trace<8, 16, TRACE_CLOSEST>(TraceArgs);
template __global__ void k_trace<8, 16, TRACE_ANY>(TraceArgs);
template __global__ void k_trace<8, 16, TRACE_MIXED>(TraceArgs);
template __global__ void k_trace<4, 20, TRACE_CLOSEST>(TraceArgs);
template __global__ void k_trace<4, 20, TRACE_ANY>(TraceArgs);
template __global__ void k_trace<4, 20, TRACE_MIXED>(TraceArgs);
template __global__ void k_trace<8, 20, TRACE_CLOSEST>(TraceArgs);
template __global__ void k_trace<8, 20, TRACE_ANY>(TraceArgs);
template __global__ void k_trace<8, 20, TRACE_MIXED>(TraceArgs);
template __global__ void k_trace<4, 32, TRACE_CLOSEST>(TraceArgs);
template __global__ void k_trace<4, 32, TRACE_ANY>(TraceArgs);
template __global__ void k_trace<4, 32, TRACE_MIXED>(TraceArgs);
template __global__ void k_trace<8, 32, TRACE_CLOSEST>(TraceArgs);
template __global__ void k_trace<8, 32, TRACE_ANY>(TraceArgs);
template __global__ void k_trace<8, 32, TRACE_MIXED>(TraceArgs);

// ---- camera ----------------------------------------------------------------------

// Kensler's hashed permutation of [0, l) (A. Kensler, "Correlated
// Multi-Jittered Sampling", Pixar Technical Memo 13-01, 2013): the per-pixel
// order in which sample passes visit the strata, in place of
// SampleRecord2D::shuffle's Fisher-Yates draw (sampleRecord2D.cpp:16-28).
__device__ __forceinline__ uint32_t kensler_permute(uint32_t i, uint32_t l, uint32_t p) {
    uint32_t w = l - 1;
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    w |= w >> 8;
    w |= w >> 16;
    do {
        i ^= p;
        i *= 0xe170893du;
        i ^= p >> 16;
        i ^= (i & w) >> 4;
        i ^= p >> 8;
        i *= 0x0929eb3fu;
        i ^= p >> 23;
        i ^= (i & w) >> 1;
        i *= 1u | p >> 27;
        i *= 0x6935fa69u;
        i ^= (i & w) >> 11;
        i *= 0x74dcb303u;
        i ^= (i & w) >> 2;
        i *= 0x9e501cc3u;
        i ^= (i & w) >> 2;
        i *= 0xc860a3dfu;
        i &= w;
        i ^= i >> 5;
    } while (i >= l);
    return (i + p) % l;
}

// Filter::evaluate of the four reference filters (boxFilter.cpp:9-11,
// gaussianFilter.cpp:25-33, coneFilter.cpp:13-16, mitchellFilter.cpp:16-40).
// F is DevScene or BatchView (the same filter fields).
__device__ __forceinline__ float mitchell1d(float b, float c, float d) {
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
template <class F>
__device__ __forceinline__ float filter_eval(const F& sc, float x, float y) {
    switch (sc.filter) {
        case 1: {
            const float v = sc.amplitude * expf(sc.g_expx * x * x + sc.g_expy * y * y);
            return mmax(0.0f, v - sc.g_edge);
        }
        case 2: return mmax(sc.filter_hw - fabsf(x), 0.0f) * mmax(sc.filter_hh - fabsf(y), 0.0f);
        case 3: return mitchell1d(sc.mitchell_b, sc.mitchell_c, x) * mitchell1d(sc.mitchell_b, sc.mitchell_c, y);
        default: return 1.0f;
    }
}

// Pixel range [x0, x1) x [y0, y1) a sample at film position (fx, fy) of the
// tile [tx0, tx1) x [ty0, ty1) reaches: FilmTile::addSample's footprint,
// clamped to the tile (filmTile.cpp:36-50).
template <class F>
__device__ __forceinline__ int4 footprint(const F& sc, float fx, float fy, int tx0, int ty0, int tx1, int ty1) {
    const float mnx = fmaxf(fx - sc.filter_hw, static_cast<float>(tx0));
    const float mny = fmaxf(fy - sc.filter_hh, static_cast<float>(ty0));
    const float mxx = fminf(fx + sc.filter_hw, static_cast<float>(tx1));
    const float mxy = fminf(fy + sc.filter_hh, static_cast<float>(ty1));
    return make_int4(static_cast<int>(ceilf(mnx - 0.5f)), static_cast<int>(ceilf(mny - 0.5f)),
                     static_cast<int>(floorf(mxx - 0.5f) + 1), static_cast<int>(floorf(mxy - 0.5f) + 1));
}

__device__ __forceinline__ bool finite3(float4 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

// The film position of sample `sample` of pixel p: the stratified (or random)
// film jitter of StratifiedSampler::requestSample2D (stratifiedSampler.cpp:38-73).
// F is DevScene or BatchView.
template <class F>
__device__ __forceinline__ void film_position(const F& sc, int32_t p, int32_t sample, uint2 key, float& fx, float& fy) {
    const int32_t ix = p % sc.img_w, iy = p / sc.img_w;
    const uint4 r = philox(make_uint4(static_cast<uint32_t>(p), static_cast<uint32_t>(sample), 0xFFFFFFFFu, 0u), key);
    const float u0 = unit24(r.x), u1 = unit24(r.y);
    float jx = u0, jy = u1;
    if (sc.sampler == 1) {
        const uint4 pk = philox(make_uint4(static_cast<uint32_t>(p), 0u, 0xFFFFFFFEu, 0u), key);
        const uint32_t base = static_cast<uint32_t>(sc.strata);
        const uint32_t j = kensler_permute(static_cast<uint32_t>(sample), static_cast<uint32_t>(sc.spp), pk.x);
        if (j < base * base) {
            const float dx = 1.0f / static_cast<float>(base);
            jx = (static_cast<float>(j % base) + u0) * dx;
            jy = (static_cast<float>(j / base) + u1) * dx;
        }
    }
    fx = static_cast<float>(ix) + jx;
    fy = static_cast<float>(iy) + jy;
}

// The film position and primary ray of sample `sample` of pixel p
// (PerspectivePinholeCamera::spawnPrimaryRay, perspectivePinholeCamera.cpp:52-73).
template <class F>
__device__ __forceinline__ V3 camera_ray(const F& sc, int32_t p, int32_t sample, uint2 key, float& fx, float& fy) {
    film_position(sc, p, sample, key, fx, fy);
    // film -> camera point, camera -> world vector (tMatrix4.ipp:180-195)
    const float* m = sc.f2c;
    const float z = 0.0f;
    const float cx = m[0] * fx + m[1] * fy + m[2] * z + m[3];
    const float cy = m[4] * fx + m[5] * fy + m[6] * z + m[7];
    const float cz = m[8] * fx + m[9] * fy + m[10] * z + m[11];
    const float* c = sc.c2w;
    return normalize(v3(c[0] * cx + c[1] * cy + c[2] * cz, c[4] * cx + c[5] * cy + c[6] * cz,
                        c[8] * cx + c[9] * cy + c[10] * cz));
}

// One sample pass's primary rays (one path per pixel).
// ADJ: also the path's adjoint weight A = sum_q dLoss/dI_q w(q, s) / W_q over the
// sample's footprint (zero for samples the film drops: NaN / inf).
template <bool ADJ>
__global__ void __launch_bounds__(256) k_camera(PassArgs a, AdjArgs g) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (gid < a.nwork) a.work[gid] = 0;
    if (gid < a.max_depth + 2) {   // queue entries and list rays per bounce
        a.qcount[2 * gid] = gid == 0 ? a.npaths : 0;
        a.qcount[2 * gid + 1] = 0;
    }
    if (gid >= a.npaths) return;
    const int32_t p = static_cast<int32_t>(gid);
    const DevScene& sc = a.sc;
    float fx, fy;
    const V3 dir = camera_ray(sc, p, a.sample, a.key, fx, fy);
    PathRec& pr = a.path[p];
    pr.fpos = make_float2(fx, fy);
    pr.beta = make_float4(1.f, 1.f, 1.f, 1.f);   // w: emission counts at the next hit
    pr.rad = make_float4(0.f, 0.f, 0.f, 0.f);
    const Queue& q = a.q[0];
    q.path[p] = p;
    q.org[p] = make_float4(sc.cam_pos[0], sc.cam_pos[1], sc.cam_pos[2], K_RAY_EPS);
    q.dir[p] = make_float4(dir.x, dir.y, dir.z, a.max_depth > 0 ? FLT_MAX : -1.0f);
    if constexpr (ADJ) {
        const int32_t ix = p % sc.img_w, iy = p / sc.img_w;
        V3 A = v3(0.f, 0.f, 0.f);
        if (finite3(g.samples[p])) {
            const int tx0 = (ix / sc.tile_x) * sc.tile_x, ty0 = (iy / sc.tile_y) * sc.tile_y;
            const int tx1 = min(tx0 + sc.tile_x, sc.img_w), ty1 = min(ty0 + sc.tile_y, sc.img_h);
            const int4 fp = footprint(sc, fx, fy, tx0, ty0, tx1, ty1);
            for (int qy = fp.y; qy < fp.w; ++qy)
                for (int qx = fp.x; qx < fp.z; ++qx) {
                    const int64_t qi = static_cast<int64_t>(qy) * sc.img_w + qx;
                    const float W = g.film[qi].w;
                    if (W == 0.0f) continue;
                    const float w = filter_eval(sc, static_cast<float>(qx) - (fx - 0.5f), static_cast<float>(qy) - (fy - 0.5f)) / W;
                    A = A + v3(g.dimg[3 * qi], g.dimg[3 * qi + 1], g.dimg[3 * qi + 2]) * w;
                }
        }
        g.A[p] = make_float4(A.x, A.y, A.z, 0.f);
        g.q[0].par[p] = make_int2(-1, 0);
        g.q[0].sC[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// ---- shading ---------------------------------------------------------------------

__device__ __forceinline__ float u01(uint4 r, int slot) {
    const uint32_t x = (slot & 3) == 0 ? r.x : (slot & 3) == 1 ? r.y : (slot & 3) == 2 ? r.z : r.w;
    return unit24(x);
}

constexpr int TEX_SLOTS = 128;
constexpr int LIGHT_SLOTS = 64;
constexpr int SHADE_WAVES = 256 / WAVE;

// Splat a finished sample into the merged film of a batch: the tile-clamped
// footprint of FilmTile::addSample (filmTile.cpp:21-66) with float atomics;
// NaN / inf samples are dropped.
__device__ __forceinline__ void batch_splat(const BatchArgs& bt, const BatchView& v, int32_t pix, float2 fp4, V3 L) {
    if (!(isfinite(L.x) && isfinite(L.y) && isfinite(L.z))) return;
    const int32_t ix = pix % v.img_w, iy = pix / v.img_w;
    const int tx0 = (ix / v.tile_x) * v.tile_x, ty0 = (iy / v.tile_y) * v.tile_y;
    const int tx1 = min(tx0 + v.tile_x, v.img_w), ty1 = min(ty0 + v.tile_y, v.img_h);
    const int4 fp = footprint(v, fp4.x, fp4.y, tx0, ty0, tx1, ty1);
    for (int qy = fp.y; qy < fp.w; ++qy)
        for (int qx = fp.x; qx < fp.z; ++qx) {
            const float w = filter_eval(v, static_cast<float>(qx) - (fp4.x - 0.5f), static_cast<float>(qy) - (fp4.y - 0.5f));
            float* f = reinterpret_cast<float*>(bt.film + v.film_base + static_cast<int64_t>(qy) * v.img_w + qx);
            unsafeAtomicAdd(f, L.x * w);
            unsafeAtomicAdd(f + 1, L.y * w);
            unsafeAtomicAdd(f + 2, L.z * w);
            unsafeAtomicAdd(f + 3, w);
        }
}

// The state of one queue entry in the shading kernel: its path's settings and
// state, and what shading the vertex produces for the next queue.
struct ShadeVtx {
    int64_t s = 0;           // entry of the current queue
    int32_t p = 0;           // pass by pass: the path (its pixel)
    int32_t pb = 0, smp = 0, pix = 0, md = 0;
    int32_t nl = 0, cluster = 0, lbase = 0, cdf0 = 0, envl = -1, est = 0, sid = 0, root_next = 0;
    uint2   key = make_uint2(0u, 0u);
    int4    meta = make_int4(0, 0, 0, 0);
    int64_t sample_idx = 0;  // the forward pass's sample record of this path (adjoint)
    V3      P0 = {0.f, 0.f, 0.f}, L = {0.f, 0.f, 0.f}, beta = {0.f, 0.f, 0.f}, A = {0.f, 0.f, 0.f};
    float4  A4 = make_float4(0.f, 0.f, 0.f, 0.f);
    bool    count_emit = false, count_next = false, pending = false;
    int2    ch = make_int2(-1, 0);
    float4  org = make_float4(0.f, 0.f, 0.f, 0.f), cdir = org, mdir = org, sdir = org, lt4 = org, fac4 = org;
    int32_t pick_light = 0;
    // adjoint: this lane's contributions and the factors of the vertex shaded now
    int32_t tkey = -1, lkey = -1;
    V3      tval = {0.f, 0.f, 0.f}, lval = {0.f, 0.f, 0.f};
    int2    par = make_int2(-1, 0);
    float4  sL4 = org, eL4 = org, sB4 = org, eB4 = org, sC4 = org;
};

// A streaming entry's path settings from its (scene, pixel, sample, bounce)
// word: the scene's view (depth, random key, lights, estimator, BVH root).
__device__ __forceinline__ void shade_view(ShadeVtx& v, const BatchArgs& bt, int4 meta) {
    v.meta = meta;
    v.sid = meta.x;
    v.pix = meta.y;
    v.smp = meta.z;
    v.pb = meta.w;
    const BatchView& bv = bt.views[v.sid];
    v.md = bv.max_depth;
    v.key = make_uint2(bv.key_lo, bv.key_hi);
    v.nl = bv.nlights;
    v.cluster = bv.cluster;
    v.lbase = bv.light_base;
    v.cdf0 = bv.cdf_base;
    v.envl = bv.env_light;
    v.est = bv.estimator;
    v.root_next = bv.root;
    v.sample_idx = static_cast<int64_t>(v.smp - bv.sample_base) * bv.img_w * bv.img_h + v.pix;
}

// Loads entry s of queue step & 1: the path's step, sample, pixel and scene
// settings, its throughput and (with rad) radiance, the vertex and its
// continuation ray and hit.
template <int MODE>
__device__ __forceinline__ void shade_load(ShadeVtx& v, const PassArgs& a, const AdjArgs& g, const BatchArgs& bt,
                                           int32_t step, int64_t s, bool rad) {
    constexpr bool ADJ = (MODE & SHADE_ADJ) != 0;
    constexpr bool BATCH = (MODE & SHADE_BATCH) != 0;
    const DevScene& sc = a.sc;
    const int qc = step & 1;
    const Queue& cur = a.q[qc];
    v.s = s;
    if constexpr (!BATCH) v.p = cur.path[s];   // pass by pass: the path of pixel p
    v.pb = step;
    v.smp = a.sample;
    v.pix = v.p;
    v.md = a.max_depth;
    v.key = a.key;
    v.nl = sc.nlights;
    v.cluster = sc.cluster;
    v.envl = sc.env_light;
    v.est = sc.estimator;
    if constexpr (BATCH) {
        shade_view(v, bt, cur.pmeta[s]);
    } else {
        v.sample_idx = v.p;
    }
    v.org = cur.org[s];
    v.P0 = v3f4(v.org);
    if (rad) v.L = v3f4(BATCH ? cur.prad[s] : a.path[v.p].rad);
    const float4 beta4 = BATCH ? cur.pbeta[s] : a.path[v.p].beta;
    v.beta = v3f4(beta4);
    // emission counts at a hit that follows the camera or a specular vertex (pt),
    // at every hit (vpt) (ptEstimator.cpp:51-56, 80-82; vptEstimator.cpp:48-51)
    v.count_emit = beta4.w != 0.0f;
    if constexpr (ADJ) {
        v.A4 = BATCH ? g.q[qc].A[s] : g.A[v.p];
        v.A = v3f4(v.A4);
    }
    v.cdir = cur.dir[s];
    v.ch = cur.hit[s];
}

// Resolves the previous vertex's NEE (its terms come pre-scaled by the path
// throughput over the light-pick pdf: L += beta * (Ld / pick),
// ptEstimator.cpp:94-99, distributed over Ld's two terms) and adds the
// background's emission when the continuation ray left the scene.
template <int MODE>
__device__ __forceinline__ void shade_resolve(ShadeVtx& v, const PassArgs& a, const AdjArgs& g, int32_t step) {
    constexpr bool ADJ = (MODE & SHADE_ADJ) != 0;
    const DevScene& sc = a.sc;
    const int qc = step & 1;
    const Queue& cur = a.q[qc];
    const int64_t s = v.s;
    const float4 lt4r = v.pb > 0 ? cur.lt[s] : make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
    const int32_t lw = __float_as_int(lt4r.w);
    if (lw != -1) {
        V3 Ld = v3(0.f, 0.f, 0.f);
        const int32_t li = lw >= 0 ? lw : -2 - lw;   // lw < -1: the shadow ray was absent
        if (lw >= 0 && cur.occ[s].x == 0) {
            Ld = Ld + v3f4(lt4r);
            if constexpr (ADJ) {
                v.tval = v.tval + v.A * v3f4(g.q[qc].sL[s]);
                v.lval = v.lval + v.A * v3f4(g.q[qc].eL[s]);
            }
        }
        const float4 f4 = cur.fac[s];
        if (f4.w != -1.0f) {   // the MIS-BSDF ray was traced (a sample pdf is never -1)
            const int2 mh = cur.mis_hit[s];   // x: -1 background, -2 a primitive that emits nothing
            if (mh.x == -1 && li == v.envl) {
                // the MIS-BSDF ray reaches the background, which the picked light lights
                const float4 md4 = cur.mis_dir[s];
                const DevLight& lg = sc.lights[li];
                const V3 Le = env_emittance(sc, lg, v3f4(md4));
                const float lp = env_pdfW(sc, lg, v3f4(md4));
                const float w = mis_power(f4.w, lp);
                Ld = Ld + divs((Le * w) * v3f4(f4), f4.w);
            }
            if (mh.x >= 0) {
                if (__float_as_int(sc.tris[3 * mh.x + 2].w) == li) {   // the record's area light
                    const float4 md4 = cur.mis_dir[s];
                    const SurfaceHit h = surface_detail(sc, mh.x, v.P0, v3f4(md4), __int_as_float(mh.y));
                    const DevLight& lg = sc.lights[li];
                    const V3 Le = light_emittance(lg, h);
                    const float lp = light_pdfW(lg, h, v.P0);
                    const float w = mis_power(f4.w, lp);
                    Ld = Ld + divs((Le * w) * v3f4(f4), f4.w);
                    if constexpr (ADJ) {
                        if (can_emit(lg, h.wi, h.Ns)) {
                            v.tval = v.tval + (v.A * v3f4(g.q[qc].sB[s])) * (load3(lg.radiance) * w);
                            v.lval = v.lval + v.A * (v3f4(g.q[qc].eB[s]) * w);
                        }
                    }
                }
            }
        }
        v.L = v.L + Ld;
        if constexpr (ADJ) {
            const float4 sc4 = g.q[qc].sC[s];
            if (sc4.w > 0.0f) {
                const V3 R = v3f4(g.samples[v.sample_idx]) - v.L;
                v.tval = v.tval + (v.A * R) * v3f4(sc4);
            }
            const int2 pp = g.q[qc].par[s];
            if (pp.x >= 0) v.tkey = 2 * pp.x + (pp.y == 3 ? 1 : 0);
            v.lkey = li;
        }
    }
    if (v.count_emit && v.pb < v.md && v.cdir.w >= 0.0f && v.ch.x < 0 && v.envl >= 0) {
        // an unbounded ray reaches the background: its emission, and the path
        // ends there (the background absorbs)
        v.L = v.L + v.beta * env_emittance(sc, sc.lights[v.envl], v3f4(v.cdir));
    }
}

// Whether the entry's continuation ray hit a surface to shade.
__device__ __forceinline__ bool shade_has_hit(const ShadeVtx& v) { return v.pb < v.md && v.cdir.w >= 0.0f && v.ch.x >= 0; }

// Shades the continuation hit (one iteration of PtEstimator::estimate's loop
// body): emission, light pick (uniformLightCluster.cpp:15-24 /
// powerLightCluster.cpp:27-34), light sample + shadow ray, MIS-BSDF ray,
// continuation sample, Russian roulette (russianRoulette.cpp:9-28). The path
// goes on (pending) with the next vertex at the hit.
template <int MODE>
__device__ __forceinline__ void shade_hit(ShadeVtx& v, const PassArgs& a, uint32_t& mis_resolved,
                                          uint32_t& shadow_resolved) {
    constexpr bool ADJ = (MODE & SHADE_ADJ) != 0;
    const DevScene& sc = a.sc;
    const int32_t pb = v.pb;
    const V3 beta0 = v.beta;
    const uint32_t upix = static_cast<uint32_t>(v.pix);
    const uint4 rA = philox(make_uint4(upix, static_cast<uint32_t>(v.smp), 2u * pb, 0u), v.key);
    const uint4 rB = philox(make_uint4(upix, static_cast<uint32_t>(v.smp), 2u * pb + 1u, 0u), v.key);
    const SurfaceHit h = surface_detail(sc, v.ch.x, v.P0, v3f4(v.cdir), __int_as_float(v.ch.y));
    const int4 info = sc.prim_info[h.prim];
    const DevMaterial mat = sc.mats[info.x];
    if (v.count_emit && info.y >= 0) {
        v.L = v.L + beta0 * light_emittance(sc.lights[info.y], h);
        if constexpr (ADJ) {
            if (can_emit(sc.lights[info.y], h.wi, h.Ns)) {
                v.lkey = info.y;
                v.lval = v.A * beta0;
            }
        }
    }
    V3 alb = v3(0.f, 0.f, 0.f);
    if constexpr (ADJ) {
        v.par = vertex_param(sc, mat, h.uvw);
        if (mat.type == 0) alb = mat_albedo(sc, mat, h.uvw);
    }
    // next-event estimation at non-specular vertices of pt; the absorber's
    // DirectLightEvaluator returns before sampling (directLightEvaluator.cpp:33-35)
    const bool nee = v.est == 0 && !mat_specular(mat.type) && mat.type != DM_ABSORB;
    v.count_next = v.est != 0 || mat_specular(mat.type);
    // lobe choices of the plastic and glass samples: counter (pixel, sample, 2b, 1)
    float uc_mis = 0.0f, uc_cont = 0.0f;
    if (mat_needs_choice(mat.type)) {
        const uint4 rC = philox(make_uint4(upix, static_cast<uint32_t>(v.smp), 2u * pb, 1u), v.key);
        uc_mis = u01(rC, SL_MISC);
        uc_cont = u01(rC, SL_CONTC);
    }
    v.sdir = make_float4(0.f, 0.f, 0.f, -1.f);
    v.lt4 = make_float4(0.f, 0.f, 0.f, 0.f);
    v.mdir = make_float4(0.f, 0.f, 0.f, -1.f);
    v.fac4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float pick = 0.0f;
    if (nee) {
        // light pick
        {
            const float u = u01(rA, SL_PICK);
            int32_t idx = 0;
            if (v.cluster == 1) {
                const float* cdf = sc.power_cdf + v.cdf0;
                for (int32_t i = v.nl - 1; i >= 0; --i)
                    if (cdf[i] <= u) { idx = i; break; }
                pick = cdf[idx + 1] - cdf[idx];
            } else {
                const uint32_t raw = static_cast<uint32_t>(0.0f + u * static_cast<float>(v.nl));
                idx = raw < static_cast<uint32_t>(v.nl - 1) ? static_cast<int32_t>(raw) : v.nl - 1;
                pick = 1.0f / static_cast<float>(v.nl);
            }
            v.pick_light = v.lbase + idx;
        }
        // light sample + shadow ray
        const DevLight& lg = sc.lights[v.pick_light];
        const LightSample ls = any_light_sample(sc, lg, h.P, u01(rA, SL_LPOS0), u01(rA, SL_LPOS1));
        if (ls.valid) {
            const V3 LV = ls.pos - h.P;
            const float dist = len(LV);
            const V3 Ldir = divs(LV, dist);
            const V3 sd = normalize(Ldir);
            v.sdir = make_float4(sd.x, sd.y, sd.z, dist - K_RAY_EPS);
            V3 f;
            float bpdf;
            bsdf_eval_pdf(sc, mat, h, Ldir, f, bpdf);
            const V3 factor = f * absdot(Ldir, h.Ns);
            const float w = mis_power(ls.pdfW, bpdf);
            const V3 lt = divs((ls.radiance * w) * factor, ls.pdfW);
            v.lt4 = make_float4(lt.x, lt.y, lt.z, 0.f);
            if constexpr (ADJ) {
                const float k = w * absdot(Ldir, h.Ns) / ls.pdfW / pick;
                const V3 e = (beta0 * f) * k;
                v.eL4 = make_float4(e.x, e.y, e.z, 0.f);
                V3 sl = v3(0.f, 0.f, 0.f);
                if (mat.type == 0) {
                    if (dot(h.wi, h.Ns) * dot(Ldir, h.Ns) > 0.0f) sl = (beta0 * ls.radiance) * (k * K_INV_PI);
                } else if (!is_zero(f)) {
                    sl = (e * ls.radiance) * ggx_dlnf_dr(sc, mat, h, Ldir);
                }
                v.sL4 = make_float4(sl.x, sl.y, sl.z, 0.f);
            }
        }
        // MIS-BSDF ray
        {
            const BsdfSample bs = bsdf_sample(sc, mat, h, u01(rB, SL_MIS0), u01(rB, SL_MIS1), uc_mis);
            if (bs.valid) {
                const V3 mdv = normalize(bs.L);
                // the term counts only when the ray's closest hit is the picked
                // area light's primitive, so the ray must hit that primitive: one
                // that misses it (the test the traversal would run on its record,
                // same origin, tmin and direction) is answered here, untraced
                bool traced = true;
                if (lg.rec >= 0) {
                    float tl = FLT_MAX;
                    traced = prim_hit(prim_load(sc.tris, lg.rec), h.P, mdv, K_RAY_EPS, tl);
                    mis_resolved += traced ? 0u : 1u;
                }
                v.mdir = make_float4(mdv.x, mdv.y, mdv.z, traced ? FLT_MAX : -1.0f);
                const V3 fc = bs.f * absdot(bs.L, h.Ns);
                v.fac4 = make_float4(fc.x, fc.y, fc.z, bs.pdf);
                if constexpr (ADJ) {
                    const float k = absdot(bs.L, h.Ns) / bs.pdf / pick;
                    const V3 e = (beta0 * bs.f) * k;
                    v.eB4 = make_float4(e.x, e.y, e.z, 0.f);
                    const V3 sb = mat.type == 0 ? beta0 * (k * K_INV_PI) : e * ggx_dlnf_dr(sc, mat, h, bs.L);
                    v.sB4 = make_float4(sb.x, sb.y, sb.z, 0.f);
                }
            }
        }
    }
    // the NEE terms pre-scaled by beta / pick; lt4.w: the picked light (-2 - it
    // when the shadow ray is absent, -1 without NEE); fac4.w: the MIS-BSDF
    // sample's pdf (-1: no MIS-BSDF ray)
    if (nee && pick > 0.0f) {
        const float ip = 1.0f / pick;
        const V3 lts = beta0 * (v3f4(v.lt4) * ip), fcs = beta0 * (v3f4(v.fac4) * ip);
        bool shadow = v.sdir.w >= 0.0f;
        // a shadow ray whose term is exactly zero adds +0 to the radiance
        // whether or not it is occluded: it is answered here (the adjoint
        // still traces it, its gradient factors need not be zero)
        if (!ADJ && shadow && is_zero(lts)) {
            shadow = false;
            v.sdir.w = -1.0f;
            ++shadow_resolved;
        }
        v.lt4 = make_float4(shadow ? lts.x : 0.f, shadow ? lts.y : 0.f, shadow ? lts.z : 0.f,
                            __int_as_float(shadow ? v.pick_light : -2 - v.pick_light));
        v.fac4 = make_float4(fcs.x, fcs.y, fcs.z, v.mdir.w >= 0.0f ? v.fac4.w : -1.0f);
    } else {
        v.lt4 = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
        v.fac4 = make_float4(0.f, 0.f, 0.f, -1.0f);
        if (!ADJ && v.sdir.w >= 0.0f) {   // no NEE term reads its result
            v.sdir.w = -1.0f;
            ++shadow_resolved;
        }
    }
    // continuation
    v.cdir = make_float4(0.f, 0.f, 0.f, -1.f);
    {
        const BsdfSample cs = bsdf_sample(sc, mat, h, u01(rB, SL_CONT0), u01(rB, SL_CONT1), uc_cont);
        if (cs.valid) {
            V3 beta = beta0 * divs(cs.f * absdot(cs.L, h.Ns), cs.pdf);
            bool alive = true;
            if (pb > 2) {
                const float qq = mclamp(1.0f - max_comp(beta), 0.05f, 1.0f);
                if (u01(rA, SL_RR) > qq) beta = beta * (1.0f / (1.0f - qq));
                else alive = false;
            }
            v.beta = beta;
            if (alive && is_zero(beta)) alive = false;
            if (alive && pb + 1 < v.md) {
                const V3 nd = normalize(cs.L);
                v.cdir = make_float4(nd.x, nd.y, nd.z, FLT_MAX);
                if constexpr (ADJ) {
                    if (mat.type == 0) {
                        v.sC4 = make_float4(alb.x > 0.f ? 1.0f / alb.x : 0.f, alb.y > 0.f ? 1.0f / alb.y : 0.f,
                                            alb.z > 0.f ? 1.0f / alb.z : 0.f, 1.f);
                    } else {
                        const float d = ggx_dlnf_dr(sc, mat, h, cs.L);
                        v.sC4 = make_float4(d, d, d, 1.f);
                    }
                }
            }
        }
    }
    v.org = make_float4(h.P.x, h.P.y, h.P.z, K_RAY_EPS);
    v.pending = true;
}

// A path that does not go on: a streaming sample splats its radiance into its
// scene's film (and the kept samples); pass by pass, the pixel's record takes
// the radiance for the film gather. (The adjoint splats nothing.)
template <int MODE>
__device__ __forceinline__ void shade_finish(const ShadeVtx& v, const PassArgs& a, const BatchArgs& bt) {
    constexpr bool ADJ = (MODE & SHADE_ADJ) != 0;
    constexpr bool BATCH = (MODE & SHADE_BATCH) != 0;
    if constexpr (BATCH) {
        if constexpr (!ADJ) {
            const BatchView& bv = bt.views[v.sid];
            // the sample's film position, drawn again from its Philox stream
            float fx, fy;
            film_position(bv, v.pix, v.smp, v.key, fx, fy);
            batch_splat(bt, bv, v.pix, make_float2(fx, fy), v.L);
            if (bt.keep)
                bt.keep[static_cast<int64_t>(v.smp - bv.sample_base) * bv.img_w * bv.img_h + v.pix] =
                    make_float4(v.L.x, v.L.y, v.L.z, 0.f);
        }
    } else {
        a.path[v.p].rad = make_float4(v.L.x, v.L.y, v.L.z, 0.f);
        a.path[v.p].beta = make_float4(v.beta.x, v.beta.y, v.beta.z, v.count_next ? 1.0f : 0.0f);
    }
}

// One step of every queued path: first the deferred half of
// DirectLightEvaluator::evaluate (directLightEvaluator.cpp:18-109) for the
// vertex shaded one step earlier, then, for the continuation hit, one
// iteration of PtEstimator::estimate's loop body (shade_hit). A path at
// bounce == max-depth only resolves.
//
// SHADE_FWD: one sample pass of one scene; every queued path is at bounce
// `step`, queues alternate by step.
//
// SHADE_ADJ replays the same steps with the same numbers (the radiance sum L
// is bit-identical to the forward pass's) and adds the path-replay gradient:
// when vertex b-1's NEE resolves, its light and MIS-BSDF terms give
// dLoss/dtexel and dLoss/dLe directly; its continuation gives
// A * (L_sample - L) * d ln f/dtexel, (L_sample - L) being all radiance the path
// gathers from vertex b on. Sampling pdfs, MIS weights and the roulette factor
// are held fixed. Each lane then has at most one texel and one light
// contribution, which the wave's LDS tables merge before one FP32 atomic per
// distinct key.
//
// SHADE_BATCH: paths of many scenes at mixed bounces (each path carries its
// scene, pixel, sample and bounce in its queue entry); queue `step` (0 / 1) is
// consumed; a path that goes on moves with its state to the other queue, one
// that does not splats its radiance into its scene's film.
//
// The forward modes shade in two phases: every entry of a tile of 256 resolves
// its NEE, and the entries whose continuation hit a surface (about half) join
// a block-wide backlog in LDS, which is shaded 256 at a time, so that the
// shading code runs with full waves instead of waves half of whose lanes have
// nothing to shade. The adjoint shades each tile in one pass.
template <int MODE>
// At least 3 waves per SIMD: the adjoint modes fit 168 VGPRs without spilling
// (left free they took 216, 2 waves; streaming adjoint 7.98 -> 7.84 ms on the
// 512^2 textured workload); the forward modes need 119-121 either way.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8)))
k_shade(PassArgs a, AdjArgs g, BatchArgs bt, int32_t step) {
    constexpr bool ADJ = (MODE & SHADE_ADJ) != 0;
    constexpr bool BATCH = (MODE & SHADE_BATCH) != 0;
    constexpr int TILE = 256;
    const int qc = step & 1;
    const Queue& nxt = a.q[qc ^ 1];
    const int32_t n = BATCH ? a.qcount[2 * qc] : a.qcount[2 * step];
    // the next queue's entry count and list length (one 64-bit word)
    int32_t* ncount = BATCH ? &a.qcount[2 * (qc ^ 1)] : &a.qcount[2 * (step + 1)];
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    __shared__ int32_t s_wcount[SHADE_WAVES], s_wfree[SHADE_WAVES], s_wlist[SHADE_WAVES], s_qbase, s_lbase;
    __shared__ int   s_tkeys[ADJ ? SHADE_WAVES * TEX_SLOTS : 1];
    __shared__ float s_tvals[ADJ ? SHADE_WAVES * TEX_SLOTS * 3 : 1];
    __shared__ int   s_lkeys[ADJ ? SHADE_WAVES * LIGHT_SLOTS : 1];
    __shared__ float s_lvals[ADJ ? SHADE_WAVES * LIGHT_SLOTS * 3 : 1];
    // forward: the backlog of entries to shade. Pass by pass: the entry and its
    // radiance after the NEE (the rest is read again from the queue). Streaming
    // (STASH): everything shading reads of the entry, so that the shading phase
    // reads no queue field: (P0, L.x) (beta, L.y) (direction, L.z), the
    // (scene, pixel, sample, bounce | emission flag << 30) word and the hit.
    constexpr bool STASH = BATCH && !ADJ;
    constexpr int BL = 2 * TILE;
    __shared__ int64_t s_bl_s[(ADJ || STASH) ? 1 : BL];
    __shared__ float   s_bl_L[(ADJ || STASH) ? 1 : BL * 3];
    __shared__ lds_f4  s_bl_a[STASH ? BL : 1], s_bl_b[STASH ? BL : 1], s_bl_c[STASH ? BL : 1];
    __shared__ lds_i4  s_bl_m[STASH ? BL : 1];
    __shared__ lds_i2  s_bl_h[STASH ? BL : 1];
    __shared__ int32_t s_bl_n;
    WaveTable<TEX_SLOTS> ttab{s_tkeys + (ADJ ? wave * TEX_SLOTS : 0), s_tvals + (ADJ ? wave * TEX_SLOTS * 3 : 0)};
    WaveTable<LIGHT_SLOTS> ltab{s_lkeys + (ADJ ? wave * LIGHT_SLOTS : 0), s_lvals + (ADJ ? wave * LIGHT_SLOTS * 3 : 0)};
    if constexpr (ADJ) {
        ttab.clear_lane(lane);
        ltab.clear_lane(lane);
    }
    if (threadIdx.x == 0) s_bl_n = 0;
    __syncthreads();
    uint32_t mis_resolved = 0, shadow_resolved = 0;   // queries answered without tracing

    // pending paths go to the next queue: wave ballots, then one atomic per block
    // (the block's waves' counts prefix-summed in LDS), so that the global counter
    // sees a quarter of the requests a per-wave atomic would send. Called by every
    // thread of the block.
    auto emit = [&](const ShadeVtx& v) __attribute__((always_inline)) {
        const bool pending = v.pending;
        const unsigned long long mask = __ballot(pending);
        // the rays of the next queue's list: the vertex's shadow ray and its traced
        // MIS-BSDF ray (absent ones are not listed, so no traversal lane draws them)
        const bool lshd = pending && v.sdir.w >= 0.0f, lmis = pending && v.mdir.w >= 0.0f;
        int ltot;
        const int lexcl = wave_excl_scan<2>((lshd ? 1 : 0) + (lmis ? 1 : 0), ltot);
        if (lane == 0) {
            s_wcount[wave] = __popcll(mask);
            s_wlist[wave] = ltot;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int32_t tq = 0, tl = 0;
            for (int w = 0; w < SHADE_WAVES; ++w) {
                const int32_t cq = s_wcount[w], cl = s_wlist[w];
                s_wcount[w] = tq;
                s_wlist[w] = tl;
                tq += cq;
                tl += cl;
            }
            // entries and list rays reserved together: (list << 32) | entries
            unsigned long long old = 0ull;
            if (tq | tl)
                old = atomicAdd(reinterpret_cast<unsigned long long*>(ncount),
                                static_cast<unsigned long long>(tq) | (static_cast<unsigned long long>(tl) << 32));
            s_qbase = static_cast<int32_t>(old & 0xFFFFFFFFull);
            s_lbase = static_cast<int32_t>(old >> 32);
        }
        __syncthreads();
        if (pending) {
            const int32_t slot =
                s_qbase + s_wcount[wave] + static_cast<int32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
            if constexpr (BATCH) {
                nxt.prad[slot] = make_float4(v.L.x, v.L.y, v.L.z, 0.f);
                nxt.pbeta[slot] = make_float4(v.beta.x, v.beta.y, v.beta.z, v.count_next ? 1.0f : 0.0f);
                nxt.pmeta[slot] = make_int4(v.meta.x, v.meta.y, v.meta.z, v.pb + 1);
            } else {
                nxt.path[slot] = v.p;
            }
            nxt.org[slot] = v.org;
            nxt.dir[slot] = v.cdir;
            if (lmis) nxt.mis_dir[slot] = v.mdir;   // read back only when its ray was traced
            nxt.lt[slot] = v.lt4;
            nxt.fac[slot] = v.fac4;
            if constexpr (BATCH) bt.root[qc ^ 1][slot] = v.root_next;
            int32_t li = s_lbase + s_wlist[wave] + lexcl;
            if (lshd) {
                nxt.lorg[li] = v.org;
                nxt.ldir[li] = v.sdir;
                nxt.lq[li] = slot;
                if constexpr (BATCH) nxt.lroot[li] = v.root_next;
                ++li;
            }
            if (lmis) {
                nxt.lorg[li] = v.org;
                nxt.ldir[li] = v.mdir;
                nxt.lq[li] = static_cast<int32_t>(static_cast<uint32_t>(slot) | 0x80000000u);
                if constexpr (BATCH) nxt.lroot[li] = v.root_next;
            }
            if constexpr (ADJ) {
                const AdjQueue& an = g.q[qc ^ 1];
                an.par[slot] = v.par;
                an.sL[slot] = v.sL4;
                an.eL[slot] = v.eL4;
                an.sB[slot] = v.sB4;
                an.eB[slot] = v.eB4;
                an.sC[slot] = v.sC4;
                if constexpr (BATCH) an.A[slot] = v.A4;
            }
        }
        __syncthreads();   // the LDS counts are read before anyone writes them again
    };

    if constexpr (ADJ) {
        // block-uniform trip count: every lane reaches the ballot and the barriers
        for (int64_t bbase = static_cast<int64_t>(blockIdx.x) * blockDim.x; bbase < n; bbase += stride) {
            const int64_t s = bbase + threadIdx.x;
            ShadeVtx v;
            if (s < n) {
                shade_load<MODE>(v, a, g, bt, step, s, true);
                shade_resolve<MODE>(v, a, g, step);
                if (shade_has_hit(v)) shade_hit<MODE>(v, a, mis_resolved, shadow_resolved);
                if (!v.pending) shade_finish<MODE>(v, a, bt);
                else if constexpr (!BATCH) shade_finish<MODE>(v, a, bt);   // pass by pass: every step
            }
            // per-wave pre-reduction, then one atomic per distinct texel / light
            if (v.tkey >= 0 && !is_zero(v.tval)) ttab.insert(v.tkey, v.tval);
            if (v.lkey >= 0 && !is_zero(v.lval)) ltab.insert(v.lkey, v.lval);
            __syncthreads();
            for (int i = lane; i < TEX_SLOTS; i += WAVE) {
                const int k = ttab.keys[i];
                if (k < 0) continue;
                const float x = ttab.vals[3 * i], y = ttab.vals[3 * i + 1], z = ttab.vals[3 * i + 2];
                float* dst = g.gtex + (k >> 1);
                if (k & 1) {
                    unsafeAtomicAdd(dst, x);
                    unsafeAtomicAdd(dst + 1, y);
                    unsafeAtomicAdd(dst + 2, z);
                } else {
                    unsafeAtomicAdd(dst, (x + y) + z);
                }
                ttab.keys[i] = -1;
                ttab.vals[3 * i] = ttab.vals[3 * i + 1] = ttab.vals[3 * i + 2] = 0.f;
            }
            for (int i = lane; i < LIGHT_SLOTS; i += WAVE) {
                const int k = ltab.keys[i];
                if (k < 0) continue;
                float* dst = g.glight + 3 * k;
                unsafeAtomicAdd(dst, ltab.vals[3 * i]);
                unsafeAtomicAdd(dst + 1, ltab.vals[3 * i + 1]);
                unsafeAtomicAdd(dst + 2, ltab.vals[3 * i + 2]);
                ltab.keys[i] = -1;
                ltab.vals[3 * i] = ltab.vals[3 * i + 1] = ltab.vals[3 * i + 2] = 0.f;
            }
            __syncthreads();
            emit(v);
        }
    } else {
        // shade the top `count` entries of the backlog (count <= TILE; block-uniform)
        auto shade_backlog = [&](int count) __attribute__((always_inline)) {
            const int top = s_bl_n;
            ShadeVtx v;
            if (static_cast<int>(threadIdx.x) < count) {
                const int i = top - count + static_cast<int>(threadIdx.x);
                if constexpr (STASH) {
                    const lds_f4 A = s_bl_a[i], B = s_bl_b[i], C = s_bl_c[i];
                    const lds_i4 M = s_bl_m[i];
                    const lds_i2 H = s_bl_h[i];
                    shade_view(v, bt, make_int4(M.x, M.y, M.z, M.w & 0x3FFFFFFF));
                    v.count_emit = (M.w & 0x40000000) != 0;
                    v.P0 = v3(A.x, A.y, A.z);
                    v.org = make_float4(A.x, A.y, A.z, 0.f);
                    v.beta = v3(B.x, B.y, B.z);
                    v.cdir = make_float4(C.x, C.y, C.z, FLT_MAX);
                    v.L = v3(A.w, B.w, C.w);
                    v.ch = make_int2(H.x, H.y);
                } else {
                    shade_load<MODE>(v, a, g, bt, step, s_bl_s[i], false);
                    v.L = v3(s_bl_L[3 * i], s_bl_L[3 * i + 1], s_bl_L[3 * i + 2]);
                }
                shade_hit<MODE>(v, a, mis_resolved, shadow_resolved);
                if (!v.pending) shade_finish<MODE>(v, a, bt);
                else if constexpr (!BATCH) shade_finish<MODE>(v, a, bt);   // pass by pass: every step
            }
            __syncthreads();   // every read of the backlog is done
            if (threadIdx.x == 0) s_bl_n = top - count;
            emit(v);           // (its barriers publish s_bl_n)
        };
        // block-uniform trip count: every lane reaches the ballot and the barriers
        for (int64_t bbase = static_cast<int64_t>(blockIdx.x) * blockDim.x; bbase < n; bbase += stride) {
            const int64_t s = bbase + threadIdx.x;
            bool hit = false;
            ShadeVtx v;
            if (s < n) {
                shade_load<MODE>(v, a, g, bt, step, s, true);
                shade_resolve<MODE>(v, a, g, step);
                hit = shade_has_hit(v);
                if (!hit) shade_finish<MODE>(v, a, bt);
            }
            // append this tile's hits to the backlog (below TILE entries are waiting)
            const unsigned long long hm = __ballot(hit);
            if (lane == 0) s_wcount[wave] = __popcll(hm);
            __syncthreads();
            int base = s_bl_n;
            for (int w = 0; w < wave; ++w) base += s_wcount[w];
            if (hit) {
                const int i = base + static_cast<int>(__popcll(hm & ((1ull << lane) - 1ull)));
                if constexpr (STASH) {
                    s_bl_a[i] = lds_f4{v.P0.x, v.P0.y, v.P0.z, v.L.x};
                    s_bl_b[i] = lds_f4{v.beta.x, v.beta.y, v.beta.z, v.L.y};
                    s_bl_c[i] = lds_f4{v.cdir.x, v.cdir.y, v.cdir.z, v.L.z};
                    s_bl_m[i] = lds_i4{v.meta.x, v.meta.y, v.meta.z, v.meta.w | (v.count_emit ? 0x40000000 : 0)};
                    s_bl_h[i] = lds_i2{v.ch.x, v.ch.y};
                } else {
                    s_bl_s[i] = s;
                    s_bl_L[3 * i] = v.L.x;
                    s_bl_L[3 * i + 1] = v.L.y;
                    s_bl_L[3 * i + 2] = v.L.z;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int t = 0;
                for (int w = 0; w < SHADE_WAVES; ++w) t += s_wcount[w];
                s_bl_n += t;
            }
            __syncthreads();
            while (s_bl_n >= TILE) shade_backlog(TILE);
        }
        while (s_bl_n > 0) shade_backlog(min(s_bl_n, TILE));
    }
    // one atomic per block for each count of queries answered here (the waves'
    // sums in LDS)
    const unsigned long long wres = wave_sum(mis_resolved), wsres = wave_sum(shadow_resolved);
    __syncthreads();   // every wave has read the last iteration's LDS counts
    if (lane == 0) {
        s_wcount[wave] = static_cast<int32_t>(wres);
        s_wfree[wave] = static_cast<int32_t>(wsres);
    }
    __syncthreads();
    if (threadIdx.x == 0 && a.stats) {
        unsigned long long t = 0, ts = 0;
        for (int w = 0; w < SHADE_WAVES; ++w) {
            t += static_cast<unsigned long long>(s_wcount[w]);
            ts += static_cast<unsigned long long>(s_wfree[w]);
        }
        if (t) atomicAdd(&a.stats[ST_MIS_RESOLVED], t);
        if (ts) atomicAdd(&a.stats[ST_SHADOW_RESOLVED], ts);
    }
}

// ---- batched renderer: path regeneration ------------------------------------------------

// The adjoint weight of a sample: A = sum_q dLoss/dI_q w(q, s) / W_q over the
// sample's tile-clamped footprint; zero for samples the film drops (NaN / inf).
template <class F>
__device__ __forceinline__ V3 adjoint_weight(const F& sc, const AdjArgs& g, int64_t film0, int32_t pix, float fx,
                                             float fy, float4 Ls) {
    V3 A = v3(0.f, 0.f, 0.f);
    if (!finite3(Ls)) return A;
    const int32_t ix = pix % sc.img_w, iy = pix / sc.img_w;
    const int tx0 = (ix / sc.tile_x) * sc.tile_x, ty0 = (iy / sc.tile_y) * sc.tile_y;
    const int tx1 = min(tx0 + sc.tile_x, sc.img_w), ty1 = min(ty0 + sc.tile_y, sc.img_h);
    const int4 fp = footprint(sc, fx, fy, tx0, ty0, tx1, ty1);
    for (int qy = fp.y; qy < fp.w; ++qy)
        for (int qx = fp.x; qx < fp.z; ++qx) {
            const int64_t qi = film0 + static_cast<int64_t>(qy) * sc.img_w + qx;
            const float W = g.film[qi].w;
            if (W == 0.0f) continue;
            const float w = filter_eval(sc, static_cast<float>(qx) - (fx - 0.5f), static_cast<float>(qy) - (fy - 0.5f)) / W;
            A = A + v3(g.dimg[3 * qi], g.dimg[3 * qi + 1], g.dimg[3 * qi + 2]) * w;
        }
    return A;
}

// New camera paths of a batch, appended to queue qc with their state (a path
// lives in the queues, one entry per path in flight): n_new = min(work left,
// queue room), all threads reading the same counters (k_bcommit applies them
// afterwards). Work item w of scene s (w in
// [work_base_s, work_base_s+1)) is sample (w - base) / pixels of pixel
// (w - base) % pixels, so consecutive lanes trace neighbouring pixels.
template <bool ADJ>
__global__ void __launch_bounds__(256) k_bregen(PassArgs a, BatchArgs bt, AdjArgs g, int32_t qc) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t next = static_cast<int64_t>(*bt.work_next);
    const int32_t qn = a.qcount[2 * qc];
    const int64_t n_new = min(bt.total_work - next, static_cast<int64_t>(bt.capacity - qn));
    if (gid >= n_new) return;
    const int64_t w = next + gid;
    int lo = 0, hi = bt.nviews - 1;    // last view with work_base <= w
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bt.views[mid].work_base <= w) lo = mid;
        else hi = mid - 1;
    }
    const BatchView& v = bt.views[lo];
    const int64_t local = w - v.work_base;
    const int64_t npx = static_cast<int64_t>(v.img_w) * v.img_h;
    const int32_t sample = v.sample_base + static_cast<int32_t>(local / npx), pix = static_cast<int32_t>(local % npx);
    float fx, fy;
    const V3 dir = camera_ray(v, pix, sample, make_uint2(v.key_lo, v.key_hi), fx, fy);
    const int32_t slot = qn + static_cast<int32_t>(gid);
    const Queue& q = a.q[qc];
    q.prad[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    q.pbeta[slot] = make_float4(1.f, 1.f, 1.f, 1.f);   // w: emission counts at the next hit
    q.pmeta[slot] = make_int4(lo, pix, sample, 0);
    q.org[slot] = make_float4(v.cam_pos[0], v.cam_pos[1], v.cam_pos[2], K_RAY_EPS);
    q.dir[slot] = make_float4(dir.x, dir.y, dir.z, v.max_depth > 0 ? FLT_MAX : -1.0f);
    bt.root[qc][slot] = v.root;
    if constexpr (ADJ) {
        const int64_t si = static_cast<int64_t>(sample - v.sample_base) * npx + pix;
        const V3 A = adjoint_weight(v, g, v.film_base, pix, fx, fy, g.samples[si]);
        g.q[qc].A[slot] = make_float4(A.x, A.y, A.z, 0.f);
        g.q[qc].par[slot] = make_int2(-1, 0);
        g.q[qc].sC[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Applies k_bregen's counts (one thread): work consumed, queue qc grown;
// clears the other queue's count and this step's traversal work counters;
// raises the done flag once no work is left and qc is empty.
__global__ void k_bcommit(PassArgs a, BatchArgs bt, int32_t qc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t next = static_cast<int64_t>(*bt.work_next);
    const int32_t qn = a.qcount[2 * qc];
    const int64_t n_new = min(bt.total_work - next, static_cast<int64_t>(bt.capacity - qn));
    *bt.work_next = static_cast<unsigned long long>(next + n_new);
    a.qcount[2 * qc] = qn + static_cast<int32_t>(n_new);
    a.qcount[2 * (qc ^ 1)] = 0;       // the queue this step's shading fills, and its list
    a.qcount[2 * (qc ^ 1) + 1] = 0;
    bt.ctr[2] = 0;
    bt.ctr[3] = 0;
    bt.ctr[1] = (next + n_new >= bt.total_work && qn + n_new == 0) ? 1 : 0;
}

template __global__ void k_camera<false>(PassArgs, AdjArgs);
template __global__ void k_camera<true>(PassArgs, AdjArgs);
template __global__ void k_shade<SHADE_FWD>(PassArgs, AdjArgs, BatchArgs, int32_t);
template __global__ void k_shade<SHADE_ADJ>(PassArgs, AdjArgs, BatchArgs, int32_t);
template __global__ void k_shade<SHADE_BATCH>(PassArgs, AdjArgs, BatchArgs, int32_t);
template __global__ void k_shade<SHADE_BATCH_ADJ>(PassArgs, AdjArgs, BatchArgs, int32_t);
template __global__ void k_bregen<false>(PassArgs, BatchArgs, AdjArgs, int32_t);
template __global__ void k_bregen<true>(PassArgs, BatchArgs, AdjArgs, int32_t);


// ---- film ------------------------------------------------------------------------

// FilmTile::addSample (filmTile.cpp:21-66) as a gather: one block per film
// tile; each pixel sums the pass's samples of its tile whose filter footprint,
// clamped to the tile, covers it. NaN / inf samples are dropped. The film
// holds sum w * rgb and sum w per pixel across passes.
__global__ void __launch_bounds__(256) k_film(PassArgs a) {
    const DevScene& sc = a.sc;
    const int ntx = (sc.img_w + sc.tile_x - 1) / sc.tile_x;
    const int tile = blockIdx.x;
    const int x0 = (tile % ntx) * sc.tile_x, y0 = (tile / ntx) * sc.tile_y;
    const int x1 = min(x0 + sc.tile_x, sc.img_w), y1 = min(y0 + sc.tile_y, sc.img_h);
    const int tw = x1 - x0, th = y1 - y0;
    const int rx = static_cast<int>(ceilf(sc.filter_hw)) + 1, ry = static_cast<int>(ceilf(sc.filter_hh)) + 1;
    for (int k = threadIdx.x; k < tw * th; k += blockDim.x) {
        const int qx = x0 + k % tw, qy = y0 + k / tw;
        float4 acc = a.film[static_cast<int64_t>(qy) * sc.img_w + qx];
        for (int sy = max(y0, qy - ry); sy <= min(y1 - 1, qy + ry); ++sy) {
            for (int sx = max(x0, qx - rx); sx <= min(x1 - 1, qx + rx); ++sx) {
                const int sp = sy * sc.img_w + sx;
                const float4 L = a.path[sp].rad;
                if (isnan(L.x) || isnan(L.y) || isnan(L.z) || isinf(L.x) || isinf(L.y) || isinf(L.z)) continue;
                const float2 fp = a.path[sp].fpos;
                const float mnx = fmaxf(fp.x - sc.filter_hw, static_cast<float>(x0));
                const float mny = fmaxf(fp.y - sc.filter_hh, static_cast<float>(y0));
                const float mxx = fminf(fp.x + sc.filter_hw, static_cast<float>(x1));
                const float mxy = fminf(fp.y + sc.filter_hh, static_cast<float>(y1));
                const int fx0 = static_cast<int>(ceilf(mnx - 0.5f)), fy0 = static_cast<int>(ceilf(mny - 0.5f));
                const int fx1 = static_cast<int>(floorf(mxx - 0.5f) + 1), fy1 = static_cast<int>(floorf(mxy - 0.5f) + 1);
                if (qx < fx0 || qx >= fx1 || qy < fy0 || qy >= fy1) continue;
                const float w = filter_eval(sc, static_cast<float>(qx) - (fp.x - 0.5f), static_cast<float>(qy) - (fp.y - 0.5f));
                acc.x += L.x * w;
                acc.y += L.y * w;
                acc.z += L.z * w;
                acc.w += w;
            }
        }
        a.film[static_cast<int64_t>(qy) * sc.img_w + qx] = acc;
    }
}

// Film::save's resolve (film.cpp:159-196), math::forward_gamma_correction
// (math.cpp:92-99) and ImageUtils::hdrToLdr (imageUtils.cpp:23-36). The power
// is pow in double rounded once to float: the correctly rounded x^(1/2.4f). Over
// every float in [0, 1] that gives the same byte as the reference's glibc powf
// (which differs from the correctly rounded value in its last bit for some
// inputs, none of them at a byte boundary); the device's single-precision powf
// is not used because its last-bit differences are not so placed.
__device__ __forceinline__ uint8_t srgb8_of(float lin) {
    const float s = lin <= 0.0031308f
                        ? 12.92f * lin
                        : 1.055f * static_cast<float>(pow(static_cast<double>(lin), static_cast<double>(1.0f / 2.4f))) -
                              0.055f;
    return static_cast<uint8_t>(mclamp(s * 255.0f + 0.5f, 0.0f, 255.0f));
}

__global__ void __launch_bounds__(256) k_srgb8(const float4* film, int64_t npix, uint8_t* out) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < npix; i += stride) {
        const float4 p = film[i];
        const float invw = (p.w == 0.0f) ? 0.0f : 1.0f / p.w;
        out[3 * i + 0] = srgb8_of(p.x * invw + 0.0f);
        out[3 * i + 1] = srgb8_of(p.y * invw + 0.0f);
        out[3 * i + 2] = srgb8_of(p.z * invw + 0.0f);
    }
}

// Scene::isIntersecting's surface intersection for rays whose closest hit is
// known (leaf position, t): per ray (creation id, t, P, Ng, Ns, uvw) as
// surface_detail computes it for shading (scene.cpp:30-48); -1 and zeros for a
// miss.
__global__ void __launch_bounds__(256) k_detail(DevScene sc, const float4* rays, const int2* hits, int64_t n,
                                                float* out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* o = out + 14 * i;
    const int2 h = hits[i];
    if (h.x < 0) {
        o[0] = -1.0f;
        for (int k = 1; k < 14; ++k) o[k] = 0.0f;
        return;
    }
    const float t = __int_as_float(h.y);
    const SurfaceHit s = surface_detail(sc, h.x, v3f4(rays[2 * i]), v3f4(rays[2 * i + 1]), t);
    const float v[14] = {static_cast<float>(s.prim), t, s.P.x, s.P.y, s.P.z, s.Ng.x, s.Ng.y, s.Ng.z,
                         s.Ns.x, s.Ns.y, s.Ns.z, s.uvw.x, s.uvw.y, s.uvw.z};
    for (int k = 0; k < 14; ++k) o[k] = v[k];
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_zero_f4(float4* p, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_zero_i32(int32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

}  // namespace mpt
