// Host-driver internals shared by render.hip (one scene: render, trace,
// adjoint, C ABI) and batch.hip (the batched renderer): device buffers,
// traversal launch shapes, the context and scene objects.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/mpt.h"
#include "../host/bvh.hpp"
#include "../host/crsd.hpp"
#include "../host/scene.hpp"
#include "kernels.hpp"

namespace mpt {

template <int W, int S, int MODE>
__global__ void k_trace(TraceArgs a);
template <bool ADJ>
__global__ void k_camera(PassArgs a, AdjArgs g);
template <int MODE>
__global__ void k_shade(PassArgs a, AdjArgs g, BatchArgs bt, int32_t step);
template <bool ADJ>
__global__ void k_bregen(PassArgs a, BatchArgs bt, AdjArgs g, int32_t qc);
__global__ void k_bcommit(PassArgs a, BatchArgs bt, int32_t qc);
__global__ void k_film(PassArgs a);
__global__ void k_srgb8(const float4* film, int64_t npix, uint8_t* out);
__global__ void k_detail(DevScene sc, const float4* rays, const int2* hits, int64_t n, float* out);
__global__ void k_zero_u64(unsigned long long* p, int n);
__global__ void k_zero_f4(float4* p, int64_t n);
__global__ void k_zero_i32(int32_t* p, int n);

// Traversal launch shape: BVH width, per-lane stack entries kept in LDS, threads per block.
struct TraceCfg {
    int width = 4, stack = 24, block = 128, bpc = 6;   // bpc: persistent blocks per CU
};

// mode: TRACE_CLOSEST, TRACE_ANY or TRACE_MIXED
template <int W, int S>
void launch_ws(int mode, int blocks, int block, size_t lds, hipStream_t st, const TraceArgs& a) {
    if (mode == TRACE_ANY) hipLaunchKernelGGL((k_trace<W, S, TRACE_ANY>), dim3(blocks), dim3(block), lds, st, a);
    else if (mode == TRACE_MIXED) hipLaunchKernelGGL((k_trace<W, S, TRACE_MIXED>), dim3(blocks), dim3(block), lds, st, a);
    else hipLaunchKernelGGL((k_trace<W, S, TRACE_CLOSEST>), dim3(blocks), dim3(block), lds, st, a);
}

inline void launch_trace_cfg(const TraceCfg& c, const TraceArgs& a, int mode, int blocks, hipStream_t st) {
    const size_t lds = (sizeof(int32_t) * static_cast<size_t>(c.stack) + static_cast<size_t>(coop_lds_bytes(c.width))) *
                       static_cast<size_t>(c.block);
    const int key = c.width * 100 + c.stack;
    switch (key) {
        case 416: launch_ws<4, 16>(mode, blocks, c.block, lds, st, a); break;
        case 420: launch_ws<4, 20>(mode, blocks, c.block, lds, st, a); break;
        case 820: launch_ws<8, 20>(mode, blocks, c.block, lds, st, a); break;
        case 432: launch_ws<4, 32>(mode, blocks, c.block, lds, st, a); break;
        case 816: launch_ws<8, 16>(mode, blocks, c.block, lds, st, a); break;
        case 824: launch_ws<8, 24>(mode, blocks, c.block, lds, st, a); break;
        case 832: launch_ws<8, 32>(mode, blocks, c.block, lds, st, a); break;
        default: launch_ws<4, 24>(mode, blocks, c.block, lds, st, a); break;
    }
}

inline TraceCfg normalize_cfg(mpt_bvh_opts& o) {
    if (o.width != 8) o.width = 4;
    if (o.leaf_max <= 0) o.leaf_max = 2;
    if (o.leaf_max > 64) o.leaf_max = 64;
    // defaults per width (measured on the 1.31M-primitive 1080p frame): BVH4's
    // traversal fits 117 VGPRs, so 8 blocks of 128 per CU (4 waves per SIMD) with
    // 20 LDS stack entries per lane (the LDS of 8 blocks then fits the CU's 160 KB;
    // 16, 20 or 24 entries and 7 or 8 blocks within 0.2%); BVH8's fits 125 VGPRs
    // but its leaf queue (8 entries per lane) leaves LDS for 7 blocks with 16
    // entries (3.39 ms per launch against 3.49 for 6 blocks with 24)
    if (o.stack_lds != 16 && o.stack_lds != 20 && o.stack_lds != 24 && o.stack_lds != 32)
        o.stack_lds = o.width == 4 ? 20 : 16;
    if (o.block != 256) o.block = 128;
    if (o.blocks_per_cu < 2 || o.blocks_per_cu > 32) o.blocks_per_cu = o.width == 4 ? 8 : 7;
    TraceCfg c;
    c.width = o.width;
    c.stack = o.stack_lds;
    c.block = o.block;
    c.bpc = o.blocks_per_cu;
    return c;
}

// Allocation-failure injection for tests: with MPT_FAIL_ALLOC_AFTER=k in the
// environment, the allocation k + 1 allocations after the variable was set (or
// changed) fails once with hipErrorOutOfMemory. Unset: no effect.
inline bool alloc_fault_injected() {
    static std::string seen;
    static long remaining = -1;
    const char* e = std::getenv("MPT_FAIL_ALLOC_AFTER");
    if (!e) {
        seen.clear();
        remaining = -1;
        return false;
    }
    if (seen != e) {
        seen = e;
        remaining = std::atol(e);
    }
    if (remaining < 0) return false;
    if (remaining-- == 0) return true;
    return false;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Replaces the buffer by one of n bytes. On failure the buffer is empty
    // (p null, bytes 0), so a size check against `bytes` never passes on a
    // buffer that does not exist.
    hipError_t alloc(size_t n) {
        release();
        if (n == 0) n = 16;
        void* q = nullptr;
        if (alloc_fault_injected()) return hipErrorOutOfMemory;
        const hipError_t e = hipMalloc(&q, n);
        if (e != hipSuccess) return e;
        p = q;
        bytes = n;
        return hipSuccess;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

// A queue's arrays, bytes per entry: path, org, dir, mis_dir, lt, fac, hit,
// mis_hit, occ, then its ray list of up to two rays per entry (lorg, ldir, lq,
// lroot), then the streaming path state (prad, pbeta, pmeta).
constexpr int QBUFS = 16;
constexpr size_t kQueueBytes[QBUFS] = {4, 16, 16, 16, 16, 16, 8, 8, 8, 32, 32, 8, 8, 16, 16, 16};

inline Queue queue_view(const DevBuf* b) {
    Queue Q;
    Q.path = b[0].as<int32_t>();
    Q.org = b[1].as<float4>();
    Q.dir = b[2].as<float4>();
    Q.mis_dir = b[3].as<float4>();
    Q.lt = b[4].as<float4>();
    Q.fac = b[5].as<float4>();
    Q.hit = b[6].as<int2>();
    Q.mis_hit = b[7].as<int2>();
    Q.occ = b[8].as<int2>();
    Q.lorg = b[9].as<float4>();
    Q.ldir = b[10].as<float4>();
    Q.lq = b[11].as<int32_t>();
    Q.lroot = b[12].as<int32_t>();
    Q.prad = b[13].as<float4>();
    Q.pbeta = b[14].as<float4>();
    Q.pmeta = b[15].as<int4>();
    return Q;
}

// The rays of a renderer's traversal launch on queue Q: its entries'
// continuation rays, then its ray list (shadow and MIS-BSDF rays); counts at
// qc2[0] (entries) and qc2[1] (list rays). Launched TRACE_MIXED (TRACE_CLOSEST
// only while the list is empty: a list's shadow rays need the any-hit lanes).
inline TraceArgs queue_trace_args(const DevScene& sc, const Queue& Q, const int32_t* qc2, bool roots) {
    TraceArgs t{};
    t.sc = sc;
    t.org = Q.org;
    t.dir = Q.dir;
    t.stride = 1;
    t.count = qc2;
    t.hit = Q.hit;
    t.lcount = qc2 + 1;
    t.lorg = Q.lorg;
    t.ldir = Q.ldir;
    t.lq = Q.lq;
    t.lroot = roots ? Q.lroot : nullptr;
    t.hit2 = Q.mis_hit;
    t.hit3 = Q.occ;
    t.want_prim = 0;
    t.mis_emitters = 1;
    t.stat_slot = ST_CLOSEST_RAYS;
    return t;
}

// Buffers of the streaming wavefront (batch.hip): two vertex queues that carry
// the paths in flight, counters, and the recorded HIP graph.
struct Streamer {
    DevBuf ctr, work_next, qcount, stats, spill, views;
    DevBuf qbuf[2][QBUFS], rootq[2];
    DevBuf adjq[2][7];                // adjoint jobs
    int32_t capacity_alloc = 0, adj_alloc = 0;
    int64_t spill_lanes = 0;
    hipGraphExec_t exec = nullptr;
    std::vector<uint8_t> exec_key;   // what the recorded graph was recorded for
    int32_t* done_host = nullptr;    // pinned
    std::vector<hipEvent_t> events;
    ~Streamer() {
        if (exec) (void)hipGraphExecDestroy(exec);
        for (auto e : events) (void)hipEventDestroy(e);
        if (done_host) (void)hipHostFree(done_host);
    }
};

// One streaming render: scenes (views) over one device scene.
struct StreamJob {
    DevScene sc;
    const BatchView* views = nullptr;   // host array of nviews
    int32_t nviews = 0;
    int64_t total_work = 0;
    float4* film = nullptr;
    int64_t film_px = 0;
    bool    clear_film = true;
    float4* keep = nullptr;
    TraceCfg cfg;
    int32_t capacity_log2 = 20, graph_steps = 16;
    bool    use_graph = false;   // record the steps into a HIP graph (else direct launches)
    bool    timing = false;      // HIP events per kernel kind (direct launches only)
    int32_t max_depth = 0;       // largest over the views
    bool    adjoint = false;     // path-replay adjoint: adj.samples / dimg / film / gtex / glight set,
    AdjArgs adj{};               // film untouched (its weights normalize dLoss/dI)
};

}  // namespace mpt

struct mpt_ctx {
    int device = 0;
    int num_cus = 0;
    void* trace_scratch = nullptr;   // this device's address of the no-allocation trace scratch array
    std::string error;
    std::unique_ptr<mpt::Streamer> streamer;   // path state of single-scene streaming renders
};

struct mpt_scene {
    mpt_ctx* ctx = nullptr;
    mpt::HostScene host;
    mpt::BinaryBvh bin;        // the tree the wide BVH is collapsed from
    mpt::BinaryBvh ref_bin;    // the reference's tree, built on first request
    bool ref_built = false;
    mpt::WideBvh wide;
    mpt::DevScene sc{};
    hipStream_t stream = nullptr;
    // scene arrays
    mpt::DevBuf nodes, tris, prim_info, prim_n, prim_uvw, mats, texs, tex_data, lights, power_cdf, env_cdf;
    // film + pass state
    mpt::DevBuf film, path, qcount, work, stats, spill, srgb;
    mpt::DevBuf qbuf[2][mpt::QBUFS];
    int64_t npaths = 0;
    // parameters that can change after the upload (differentiable rendering)
    std::vector<mpt::DevMaterial> h_mats;
    std::vector<float> h_cdf;
    std::vector<float> h_envcdf;
    std::vector<mpt::DevLight> h_lights;
    std::vector<mpt::DevTexture> h_texs;
    std::vector<mpt::F3> h_rad;
    int64_t tex_elems = 0;
    uint64_t param_version = 0;
    // per-sample radiance kept by a forward render for the adjoint
    mpt::DevBuf samples;
    bool keep_valid = false;
    uint64_t keep_seed = 0, keep_version = 0;
    int32_t keep_begin = 0, keep_end = 0, keep_depth = 0;
    // adjoint buffers
    mpt::DevBuf adjq[2][6], adjA, dimg, gtex, glight;
    int64_t adj_npaths = 0;
    int64_t spill_lanes = 0;
    int32_t max_depth_alloc = -1;
    mpt_scene_stats sstats{};
    mpt_stats rstats{};
    mpt_bvh_opts opts{};
    mpt::TraceCfg cfg;
    std::vector<hipEvent_t> events;
    ~mpt_scene() {
        for (auto e : events) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace mpt {

inline int fail(mpt_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->error = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                      \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(ctx, MPT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// The streaming wavefront (batch.hip): runs job on `stream`, fills out's ray
// counts and times.
int stream_render(mpt_ctx* ctx, hipStream_t stream, Streamer& st, const StreamJob& job, mpt_stats* out);

template <typename T>
inline int upload_vec(mpt_ctx* ctx, DevBuf& b, const std::vector<T>& v) {
    HIP_TRY(ctx, b.alloc(v.size() * sizeof(T)));
    if (!v.empty()) HIP_TRY(ctx, hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return MPT_OK;
}

inline double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace mpt
