// Launch-argument PODs shared by the kernels (kernels.hip) and the host driver
// (render.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dscene.hpp"

namespace mpt {

// Work counters and statistics, one int/uint64 slot each.
enum StatSlot {
    ST_CLOSEST_RAYS = 0,  // valid continuation / camera rays traced
    ST_MIS_RAYS,          // valid MIS-BSDF rays traced
    ST_SHADOW_RAYS,       // valid shadow rays traced
    ST_BOXES,             // slab tests of closest-hit queries
    ST_PRIMS,             // primitive tests of closest-hit queries
    ST_BOXES_ANY,         // slab tests of any-hit queries
    ST_PRIMS_ANY,         // primitive tests of any-hit queries
    ST_ERRORS,            // traversal stack overruns (must stay 0)
    ST_MIS_RESOLVED,      // MIS-BSDF queries answered by the picked light's own record (not traced)
    ST_SHADOW_RESOLVED,   // shadow queries answered untraced: their term is exactly zero
    ST_COUNT
};

// The traversal tests queued leaves once LEAF_DEFER are queued in a wave, or
// LEAF_PARK of its lanes have ended traversal and wait on queued leaves, or no
// lane is left traversing (a multiple of 8: the queue's extra entries are
// spread over the wave's 64 threads). 1080p frame, ms per launch: LEAF_DEFER
// 24 / 32 / 40 / 48 / 56 / 64 -> 1.22 / 1.21 / 1.20 / 1.18-1.19 / 1.20 / 1.21;
// LEAF_PARK 4 / 8 / 12 at 48 -> 1.25 / 1.18 / 1.19.
constexpr int LEAF_DEFER = 48;
// Rays a traversal wave reserves with one atomic on the launch's work counter
// (1080p frame, ms per launch: 64 / 128 / 256 / 512 -> 3.19 / 2.40 / 2.14 / 2.07
// when a wave refilled once 56 lanes lacked a ray; 1.93 refilling any lane).
constexpr int TRACE_CHUNK = 512;
// A launch of fewer than TRACE_SPREAD * TRACE_CHUNK rays per wave of its grid
// reserves n / (TRACE_SPREAD * waves) rays at a time, but no fewer than
// TRACE_CHUNK_MIN, so that small launches still spread over many waves
// (sequential pass-by-pass renders of the 64-scene batch, s: a fixed 512 6.03;
// n / (SPREAD * waves) with SPREAD 4 / 2 / 1 and at least 16 rays 5.09 / 4.68 /
// 4.35; SPREAD 1 and at least 4 / 32 / 64 rays 4.61 / 4.27 / 4.17; the 1080p
// frame's launches within 1%).
constexpr int TRACE_SPREAD = 1;
constexpr int TRACE_CHUNK_MIN = 64;
constexpr int LEAF_PARK = 8;

// LDS bytes per traversal thread beyond its stack entries: its ray (two
// float4), its best-hit key and W + LEAF_DEFER / 64 leaf-queue entries (int2).
__host__ __device__ constexpr int coop_lds_bytes(int W) { return 32 + 8 + 8 * W + LEAF_DEFER / 8; }

// A batch of rays for the persistent traversal kernel. Ray i < n0 is
// (org[i * stride], dir[i * stride]) -> hit[i] (closest hit, or any hit in a
// TRACE_ANY launch). A renderer launch (TRACE_MIXED) then traces the *lcount
// rays of its queue's ray list: ray n0 + k is (lorg[k], ldir[k]), answered into
// the queue slot lq[k] names: lq >= 0 a shadow ray, any hit -> hit3[lq]; lq < 0
// a MIS-BSDF ray, closest hit -> hit2[lq & 0x7FFFFFFF]. The list holds only rays
// that are present, so no lane draws an absent one. A ray with dir.w (tmax) < 0
// is absent and reports a miss.
struct TraceArgs {
    DevScene       sc;
    const float4*  org;
    const float4*  dir;
    int32_t        stride;
    const int32_t* count;        // n0 = *count (device) when non-null, else n0 = n_host
    int64_t        n_host;
    int2*          hit;          // closest: (leaf or prim, t bits); any: x = 0/1
    const int32_t* lcount;       // rays in the list (device; null: no list)
    const float4*  lorg;
    const float4*  ldir;
    const int32_t* lq;
    const int32_t* lroot;        // per list ray root node, null = node 0
    int2*          hit2;         // MIS-BSDF rays' closest hits
    int2*          hit3;         // shadow rays' any hits
    int32_t        want_prim;    // closest: write the creation index instead of the leaf position
    int32_t        mis_emitters; // MIS-BSDF rays: report a hit on a primitive without an area light as -2
    const int32_t* root;         // per-ray root node of the first n0, null = node 0
    int32_t*       work;         // atomic work counter, zero on entry
    int32_t*       spill;        // per-lane stack spill, (max_stack - S) entries per lane
    unsigned long long* stats;
    int32_t        stat_slot;    // ST_CLOSEST_RAYS or ST_SHADOW_RAYS
};

// Path state per path (one path per pixel in a sample pass) and the ping-pong
// queues of pending path vertices. The queue's entry count and its ray list's
// length sit side by side in one 64-bit word (qcount[2k], qcount[2k + 1]), so
// that shading reserves both with one atomic.
struct Queue {
    int32_t* path;     // path id
    float4*  org;      // vertex position, tmin
    float4*  dir;      // continuation ray (w = tmax, < 0 absent)
    float4*  mis_dir;  // MIS-BSDF ray (written only when it is traced)
    float4*  lt;       // light-sample term of DirectLightEvaluator x throughput / pick pdf (xyz);
                       // w (int bits): the picked light, -2 - it when the shadow ray is absent, -1 no NEE
    float4*  fac;      // MIS-BSDF factor f |L.N| x throughput / pick pdf (xyz), sample pdf (w; -1: no ray)
    int2*    hit;      // continuation hit
    int2*    mis_hit;
    int2*    occ;      // shadow: x = occluded
    // the ray list: the entries' shadow rays and traced MIS-BSDF rays, compacted
    // (up to two per entry; TraceArgs)
    float4*  lorg;
    float4*  ldir;
    int32_t* lq;
    int32_t* lroot;    // batches: the ray's BVH root
    // streaming renders: the path's state travels with its pending vertex (read
    // and written in queue order; pass by pass it sits in PathRec per pixel)
    float4*  prad;     // radiance so far (xyz)
    float4*  pbeta;    // throughput (xyz); w: emission counts at the next hit
    int4*    pmeta;    // scene, pixel, sample, bounce
};

// Per-vertex factors of the path-replay adjoint, stored beside Queue when a
// vertex is shaded and applied when its NEE rays and continuation resolve at the
// next bounce. "Texel" is the vertex's differentiable parameter: the albedo
// texel of a Lambertian vertex (3 channels) or the roughness texel of a GGX
// vertex (1 channel).
struct AdjQueue {
    int2*   par;   // (texture element index of the texel, channels) or (-1, 0)
    float4* sL;    // light term:   dLterm / dtexel (per channel)
    float4* eL;    // light term:   dLterm / dLe
    float4* sB;    // MIS-BSDF term: dLterm / dtexel = sB * Le * w (w, Le known at resolution)
    float4* eB;    // MIS-BSDF term: dLterm / dLe = eB * w
    float4* sC;    // continuation: d ln f / dtexel (xyz), w = 1 when the continuation ray was traced
    float4* A;     // streaming: the path's adjoint weight (pass by pass: AdjArgs::A per pixel)
};

struct AdjArgs {
    AdjQueue      q[2];
    const float4* samples;  // the forward pass's radiance per sample: [pixel] of this pass (pass by
                            // pass) or [sample - base][pixel] (streaming)
    float4*       A;        // pass by pass: per-pixel adjoint weight dLoss / dL_sample
    const float*  dimg;     // dLoss / dI, npix x 3
    const float4* film;     // the forward film (sum w per pixel)
    float*        gtex;     // dLoss / d(texture element)
    float*        glight;   // dLoss / dLe, nlights x 3
};

// One scene of a batch: where its tables sit in the merged scene, and its
// camera, film and renderer settings (the per-scene fields of DevScene).
struct BatchView {
    int32_t root;                // merged node index of the scene's BVH root
    int32_t light_base, nlights, cluster, cdf_base;
    int32_t env_light;           // merged index of the scene's environment light, -1 none
    int32_t estimator;           // 0 pt, 1 vpt
    int32_t img_w, img_h, tile_x, tile_y;
    int32_t filter;
    float   filter_hw, filter_hh, g_expx, g_expy, g_edge, amplitude, mitchell_b, mitchell_c;
    int32_t sampler, spp, strata, max_depth;   // spp / strata: the sampler's (stratification)
    int32_t sample_base, nsamples;             // samples [sample_base, sample_base + nsamples) are rendered
    uint32_t key_lo, key_hi;     // Philox key (the scene's seed)
    int64_t film_base;           // first pixel of the scene in the merged film
    int64_t work_base;           // first work item of the scene (items: sample-major pixel x sample)
    float   f2c[12], c2w[12], cam_pos[3];
};

// Streaming state of the batched renderer: paths live in the vertex queues;
// new camera paths are drawn from one global work counter.
struct BatchArgs {
    const BatchView* views;
    int32_t  nviews;
    int32_t  capacity;           // paths in flight (= queue entries)
    int64_t  total_work;
    int32_t* root[2];            // per queue entry: BVH root of the entry's scene
    float4*  film;               // merged linear film (sum w*rgb, sum w)
    float4*  keep;               // one scene: each sample's radiance [sample - base][pixel], or null
    int32_t* ctr;                // [1] done flag, [2..3] trace work counters
    unsigned long long* work_next;
    unsigned long long* stats;
};

// Modes of the shading kernel (bit 0: path-replay adjoint, bit 1: streaming
// wavefront of mixed-bounce paths; else one sample pass of one scene).
enum { SHADE_FWD = 0, SHADE_ADJ = 1, SHADE_BATCH = 2, SHADE_BATCH_ADJ = 3 };

// The state of one path of a pass-by-pass render, one 64-byte record per pixel
// so that a shading step reads and writes one line per path and the film
// gather finds each pixel's radiance (streaming renders carry the path state
// in the queues instead).
struct __align__(64) PathRec {
    float4 rad;    // radiance so far (xyz)
    float4 beta;   // throughput (xyz); w: emission counts at the next hit
    int4   meta;   // unused (pads the record to 64 bytes)
    float2 fpos;   // film position of the path's sample
    float2 pad;
};

struct PassArgs {
    DevScene sc;
    uint2    key;          // Philox key (seed)
    int32_t  sample;       // sample index of this pass
    int32_t  max_depth;
    int32_t  npaths;
    PathRec* path;         // pass by pass: per pixel (streaming: null, the state is in the queues)
    Queue    q[2];
    int32_t* qcount;       // [2 (max_depth + 2)]: entries and list rays of each queue (Queue)
    int32_t* work;         // traversal work counters, 3 per bounce
    int32_t  nwork;
    float4*  film;         // sum w*rgb, sum w
    unsigned long long* stats;   // StatSlot counters (shading adds ST_MIS_RESOLVED)
};

}  // namespace mpt
