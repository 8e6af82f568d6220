// Batched renderer: many scenes in one streaming wavefront on one GPU.
//
// The scenes' device tables are merged into one DevScene (node child indices,
// leaf positions, primitive ids, material / texture / light indices offset),
// so the traversal kernel is the single-scene one with a per-ray root node.
// A path lives in the vertex queues (one pending vertex per path in flight,
// carrying the path's state): each step appends new camera paths up to the
// queue's capacity from one global work counter over (scene, sample, pixel)
// (k_bregen / k_bcommit), traces the queue's continuation rays and its ray
// list (shadow and MIS-BSDF rays), and shades (k_shade<SHADE_BATCH>), where
// finished paths splat into their scene's film and leave the queue, and
// pending ones move to the next queue. Steps are recorded once into a HIP graph
// and the graph is relaunched until the device's done flag is raised, so the
// GPU stays full across scene boundaries and path-length tails instead of
// draining at the end of every sample pass of every scene.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <thread>

#include "driver.hpp"

namespace mpt {

// ---- merging kernels ------------------------------------------------------------------------

__global__ void k_fix_tris(float4* tris, int64_t n, int32_t prim_base, int32_t light_base) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t bits = __float_as_uint(tris[3 * j].w);
    const uint32_t id = (bits & 0x7FFFFFFFu) + static_cast<uint32_t>(prim_base);
    tris[3 * j].w = __uint_as_float(id | (bits & 0x80000000u));
    const int32_t l = __float_as_int(tris[3 * j + 2].w);   // the primitive's area light
    tris[3 * j + 2].w = __int_as_float(l >= 0 ? l + light_base : -1);
}

__global__ void k_fix_info(int4* info, int64_t n, int32_t mat_base, int32_t light_base) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int4 v = info[j];
    v.x += mat_base;
    v.y = v.y >= 0 ? v.y + light_base : -1;
    info[j] = v;
}

__global__ void k_batch_init(int32_t* ctr, int32_t* qcount, unsigned long long* work_next, unsigned long long* stats) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctr[0] = 0;
        ctr[1] = 0;
        ctr[2] = ctr[3] = 0;
        qcount[0] = qcount[1] = qcount[2] = qcount[3] = 0;
        *work_next = 0ull;
    }
    if (i < ST_COUNT) stats[i] = 0ull;
}

}  // namespace mpt

struct mpt_batch {
    mpt_ctx* ctx = nullptr;
    std::vector<mpt_scene*> scenes;
    hipStream_t stream = nullptr;
    mpt_batch_opts opts{};
    // merged scene; nodes per BVH width (index 0: 4, 1: 8), built on demand
    mpt::DevBuf nodes[2];
    int64_t nnodes[2] = {0, 0};
    int32_t max_stack[2] = {0, 0};
    std::vector<int32_t> roots[2];
    mpt::DevBuf tris, prim_info, prim_n, prim_uvw, mats, texs, tex_data, lights, power_cdf, env_cdf, film;
    mpt::DevScene sc{};
    std::vector<mpt::BatchView> hviews;
    int64_t nprims = 0, total_px = 0;
    mpt::Streamer st;
    mpt_stats rstats{};
    ~mpt_batch() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

using namespace mpt;

void normalize_opts(mpt_batch_opts& o) {
    if (o.width != 8) o.width = 4;
    if (o.stack_lds != 16 && o.stack_lds != 20 && o.stack_lds != 24 && o.stack_lds != 32)
        o.stack_lds = o.width == 4 ? 20 : 16;
    if (o.block != 256) o.block = 128;
    if (o.capacity_log2 < 16 || o.capacity_log2 > 24) o.capacity_log2 = 21;
    if (o.graph_steps < 2 || o.graph_steps > 64) o.graph_steps = 16;
    if (o.blocks_per_cu < 2 || o.blocks_per_cu > 32) o.blocks_per_cu = o.width == 4 ? 8 : 7;
    o.graph_steps &= ~1;
}

// Merged wide-BVH nodes for one width: each scene's nodes (collapsed again
// from its binary tree when the scene was built with the other width) with
// interior children offset by the scene's first node and leaves by its first
// primitive record.
int merge_nodes(mpt_batch* B, int wi) {
    if (B->nnodes[wi] > 0) return MPT_OK;
    const int width = wi == 0 ? 4 : 8;
    const size_t nb = static_cast<size_t>(width == 4 ? sizeof(QNode<4>) : sizeof(QNode<8>));
    const size_t child_at = width == 4 ? offsetof(QNode<4>, child) : offsetof(QNode<8>, child);
    const size_t count_at = width == 4 ? offsetof(QNode<4>, count) : offsetof(QNode<8>, count);
    const size_t ns = B->scenes.size();
    std::vector<WideBvh> other(ns);
    std::vector<std::thread> th;
    for (size_t i = 0; i < ns; ++i) {
        mpt_scene* S = B->scenes[i];
        if (S->wide.width != width)
            th.emplace_back([S, &other, i, width]() { collapse_wide(S->host, S->bin, width, S->opts.leaf_max, &other[i]); });
    }
    for (auto& t : th) t.join();
    for (size_t i = 0; i < ns; ++i)
        if (B->scenes[i]->wide.width != width && !other[i].ok)
            return fail(B->ctx, MPT_ERR_UNSUPPORTED, "a BVH node cannot be encoded");
    int64_t total = 0;
    std::vector<int64_t> nbase(ns), tbase(ns);
    int64_t tb = 0;
    int32_t ms = 0;
    for (size_t i = 0; i < ns; ++i) {
        const WideBvh& w = B->scenes[i]->wide.width == width ? B->scenes[i]->wide : other[i];
        nbase[i] = total;
        tbase[i] = tb;
        total += w.nnodes;
        tb += B->scenes[i]->host.nprims();
        ms = std::max(ms, w.max_stack);
    }
    if (total >= (int64_t{1} << 31) || tb >= (int64_t{1} << 31))
        return fail(B->ctx, MPT_ERR_UNSUPPORTED, "batch larger than 2^31 nodes or primitives");
    std::vector<uint8_t> merged(static_cast<size_t>(total) * nb);
    for (size_t i = 0; i < ns; ++i) {
        const WideBvh& w = B->scenes[i]->wide.width == width ? B->scenes[i]->wide : other[i];
        uint8_t* dst = merged.data() + static_cast<size_t>(nbase[i]) * nb;
        std::memcpy(dst, w.nodes.data(), static_cast<size_t>(w.nnodes) * nb);
        for (int64_t k = 0; k < w.nnodes; ++k) {
            int32_t* child = reinterpret_cast<int32_t*>(dst + static_cast<size_t>(k) * nb + child_at);
            const uint16_t* count = reinterpret_cast<const uint16_t*>(dst + static_cast<size_t>(k) * nb + count_at);
            for (int c = 0; c < width; ++c) {
                if (count[c] > 0) child[c] += static_cast<int32_t>(tbase[i]);
                else if (child[c] >= 0) child[c] += static_cast<int32_t>(nbase[i]);
            }
        }
    }
    HIP_TRY(B->ctx, B->nodes[wi].alloc(merged.size()));
    HIP_TRY(B->ctx, hipMemcpy(B->nodes[wi].p, merged.data(), merged.size(), hipMemcpyHostToDevice));
    B->nnodes[wi] = total;
    B->max_stack[wi] = ms;
    B->roots[wi].assign(nbase.begin(), nbase.end());
    return MPT_OK;
}

int merge_tables(mpt_batch* B) {
    mpt_ctx* ctx = B->ctx;
    const size_t ns = B->scenes.size();
    int64_t np = 0, nmat = 0, ntexel = 0, nlight = 0, ncdf = 0, npx = 0;
    bool any_n = false, any_u = false;
    std::vector<DevMaterial> mats;
    std::vector<DevTexture> texs;
    std::vector<DevLight> lights;
    std::vector<float> cdf, envcdf;
    std::vector<int64_t> pbase(ns), tdbase(ns);
    for (size_t i = 0; i < ns; ++i) {
        mpt_scene* S = B->scenes[i];
        any_n |= !S->host.prim_n.empty();
        any_u |= !S->host.prim_uvw.empty();
    }
    B->hviews.assign(ns, BatchView{});
    int64_t work = 0;
    for (size_t i = 0; i < ns; ++i) {
        mpt_scene* S = B->scenes[i];
        const HostScene& h = S->host;
        pbase[i] = np;
        tdbase[i] = ntexel;
        const int32_t tex0 = static_cast<int32_t>(texs.size());
        for (DevMaterial m : S->h_mats) {
            if (m.albedo_tex >= 0) m.albedo_tex += tex0;
            if (m.rough_tex >= 0) m.rough_tex += tex0;
            mats.push_back(m);
        }
        for (DevTexture t : S->h_texs) {
            t.offset += ntexel;
            if (t.odd >= 0) {
                t.odd += tex0;
                t.even += tex0;
            }
            texs.push_back(t);
        }
        const int32_t env0 = static_cast<int32_t>(envcdf.size());
        for (DevLight l : S->h_lights) {
            if (l.type == LIGHT_TYPE_ENV) {
                l.env_tex += tex0;
                l.env_cdf += env0;
            } else {
                l.prim += static_cast<int32_t>(np);
                l.rec += static_cast<int32_t>(np);   // the scene's records start at np in the merged array
            }
            lights.push_back(l);
        }
        envcdf.insert(envcdf.end(), S->h_envcdf.begin(), S->h_envcdf.end());
        BatchView& v = B->hviews[i];
        const DevScene& d = S->sc;
        v.light_base = static_cast<int32_t>(nlight);
        v.env_light = d.env_light >= 0 ? static_cast<int32_t>(nlight) + d.env_light : -1;
        v.estimator = d.estimator;
        v.nlights = static_cast<int32_t>(h.lights.size());
        v.cluster = h.cluster;
        v.cdf_base = static_cast<int32_t>(ncdf);
        cdf.insert(cdf.end(), S->h_cdf.begin(), S->h_cdf.end());
        v.img_w = d.img_w;
        v.img_h = d.img_h;
        v.tile_x = d.tile_x;
        v.tile_y = d.tile_y;
        v.filter = d.filter;
        v.filter_hw = d.filter_hw;
        v.filter_hh = d.filter_hh;
        v.g_expx = d.g_expx;
        v.g_expy = d.g_expy;
        v.g_edge = d.g_edge;
        v.amplitude = d.amplitude;
        v.mitchell_b = d.mitchell_b;
        v.mitchell_c = d.mitchell_c;
        v.sampler = d.sampler;
        v.spp = d.spp;
        v.strata = d.strata;
        v.max_depth = h.max_depth;
        v.film_base = npx;
        v.work_base = work;
        std::memcpy(v.f2c, d.f2c, sizeof(v.f2c));
        std::memcpy(v.c2w, d.c2w, sizeof(v.c2w));
        std::memcpy(v.cam_pos, d.cam_pos, sizeof(v.cam_pos));
        np += h.nprims();
        nmat += static_cast<int64_t>(S->h_mats.size());
        ntexel += S->tex_elems;
        nlight += static_cast<int64_t>(h.lights.size());
        ncdf += static_cast<int64_t>(S->h_cdf.size());
        npx += static_cast<int64_t>(h.width) * h.height;
        work += static_cast<int64_t>(h.width) * h.height * h.spp;
    }
    if (np > MPT_MAX_PRIMITIVES || nmat >= (int64_t{1} << 31) || ntexel >= (int64_t{1} << 30))
        return fail(ctx, MPT_ERR_UNSUPPORTED, "batch tables too large (at most MPT_MAX_PRIMITIVES primitives in all)");
    B->nprims = np;
    B->total_px = npx;
    HIP_TRY(ctx, B->tris.alloc(sizeof(float4) * 3 * static_cast<size_t>(np)));
    HIP_TRY(ctx, B->prim_info.alloc(sizeof(int4) * static_cast<size_t>(np)));
    HIP_TRY(ctx, B->prim_n.alloc(any_n ? sizeof(float) * 9 * static_cast<size_t>(np) : 16));
    HIP_TRY(ctx, B->prim_uvw.alloc(any_u ? sizeof(float) * 9 * static_cast<size_t>(np) : 16));
    HIP_TRY(ctx, B->tex_data.alloc(sizeof(float) * static_cast<size_t>(std::max<int64_t>(1, ntexel))));
    for (size_t i = 0; i < ns; ++i) {
        mpt_scene* S = B->scenes[i];
        const int64_t n = S->host.nprims();
        if (n == 0) continue;
        float4* t = B->tris.as<float4>() + 3 * pbase[i];
        int4* inf = B->prim_info.as<int4>() + pbase[i];
        HIP_TRY(ctx, hipMemcpyAsync(t, S->tris.p, sizeof(float4) * 3 * static_cast<size_t>(n), hipMemcpyDeviceToDevice, B->stream));
        HIP_TRY(ctx, hipMemcpyAsync(inf, S->prim_info.p, sizeof(int4) * static_cast<size_t>(n), hipMemcpyDeviceToDevice, B->stream));
        const int blocks = static_cast<int>((n + 255) / 256);
        int32_t mat0 = 0, light0 = 0;
        for (size_t k = 0; k < i; ++k) {
            mat0 += static_cast<int32_t>(B->scenes[k]->h_mats.size());
            light0 += static_cast<int32_t>(B->scenes[k]->host.lights.size());
        }
        hipLaunchKernelGGL(k_fix_tris, dim3(blocks), dim3(256), 0, B->stream, t, n, static_cast<int32_t>(pbase[i]),
                           light0);
        hipLaunchKernelGGL(k_fix_info, dim3(blocks), dim3(256), 0, B->stream, inf, n, mat0, light0);
        HIP_TRY(ctx, hipGetLastError());
        if (!S->host.prim_n.empty())
            HIP_TRY(ctx, hipMemcpyAsync(B->prim_n.as<float>() + 9 * pbase[i], S->prim_n.p, sizeof(float) * 9 * static_cast<size_t>(n),
                                        hipMemcpyDeviceToDevice, B->stream));
        if (!S->host.prim_uvw.empty())
            HIP_TRY(ctx, hipMemcpyAsync(B->prim_uvw.as<float>() + 9 * pbase[i], S->prim_uvw.p,
                                        sizeof(float) * 9 * static_cast<size_t>(n), hipMemcpyDeviceToDevice, B->stream));
        if (S->tex_elems > 0)
            HIP_TRY(ctx, hipMemcpyAsync(B->tex_data.as<float>() + tdbase[i], S->tex_data.p,
                                        sizeof(float) * static_cast<size_t>(S->tex_elems), hipMemcpyDeviceToDevice, B->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(B->stream));
    int rc;
    if ((rc = upload_vec(ctx, B->mats, mats)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, B->texs, texs)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, B->lights, lights)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, B->power_cdf, cdf)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, B->env_cdf, envcdf)) != MPT_OK) return rc;
    HIP_TRY(ctx, B->film.alloc(sizeof(float4) * static_cast<size_t>(std::max<int64_t>(1, npx))));
    DevScene& d = B->sc;
    d = DevScene{};
    d.nprims = np;
    d.tris = B->tris.as<float4>();
    d.prim_info = B->prim_info.as<int4>();
    d.prim_n = B->prim_n.as<float>();
    d.prim_uvw = B->prim_uvw.as<float>();
    d.mats = B->mats.as<DevMaterial>();
    d.texs = B->texs.as<DevTexture>();
    d.tex_data = B->tex_data.as<float>();
    d.lights = B->lights.as<DevLight>();
    d.power_cdf = B->power_cdf.as<float>();
    d.env_cdf = B->env_cdf.as<float>();
    d.env_light = -1;   // per scene: BatchView::env_light
    d.nlights = static_cast<int32_t>(nlight);
    return MPT_OK;
}

// Streaming buffers for `cap` paths in flight.
int ensure_stream_buffers(mpt_ctx* ctx, Streamer& S, int32_t cap) {
    if (!S.done_host) {
        HIP_TRY(ctx, hipHostMalloc(reinterpret_cast<void**>(&S.done_host), sizeof(int32_t), hipHostMallocDefault));
        *S.done_host = 0;
    }
    if (cap <= S.capacity_alloc) return MPT_OK;   // larger buffers serve smaller jobs
    if (S.exec) {   // the recorded graph holds the old buffers
        (void)hipGraphExecDestroy(S.exec);
        S.exec = nullptr;
    }
    const size_t n = static_cast<size_t>(cap);
    S.capacity_alloc = 0;   // until every buffer below has its new size
    for (int q = 0; q < 2; ++q) {
        for (int k = 0; k < QBUFS; ++k) HIP_TRY(ctx, S.qbuf[q][k].alloc(kQueueBytes[k] * n));
        HIP_TRY(ctx, S.rootq[q].alloc(sizeof(int32_t) * n));
    }
    if (!S.ctr.p) {
        HIP_TRY(ctx, S.ctr.alloc(sizeof(int32_t) * 4));
        HIP_TRY(ctx, S.work_next.alloc(sizeof(unsigned long long)));
        HIP_TRY(ctx, S.qcount.alloc(sizeof(int32_t) * 4));
        HIP_TRY(ctx, S.stats.alloc(sizeof(unsigned long long) * ST_COUNT));
    }
    S.capacity_alloc = cap;
    return MPT_OK;
}

enum { SK_REGEN = 0, SK_CLOSEST, SK_ANY, SK_SHADE, SK_KINDS };

struct StepShape {
    TraceCfg cfg;
    int tblocks, sblocks, rblocks;
};

// One wavefront step on queue qc. With events (2 per kernel kind), each kind's
// launches are bracketed for timing.
int enqueue_step(mpt_ctx* ctx, hipStream_t stream, Streamer& S, const StepShape& sh, PassArgs& a, BatchArgs& bt,
                 const AdjArgs* g, int qc, hipEvent_t* ev) {
    auto rec = [&](int k, int e) -> hipError_t { return ev ? hipEventRecord(ev[2 * k + e], stream) : hipSuccess; };
    HIP_TRY(ctx, rec(SK_REGEN, 0));
    if (g) hipLaunchKernelGGL(k_bregen<true>, dim3(sh.rblocks), dim3(256), 0, stream, a, bt, *g, qc);
    else hipLaunchKernelGGL(k_bregen<false>, dim3(sh.rblocks), dim3(256), 0, stream, a, bt, AdjArgs{}, qc);
    hipLaunchKernelGGL(k_bcommit, dim3(1), dim3(64), 0, stream, a, bt, qc);
    HIP_TRY(ctx, rec(SK_REGEN, 1));
    // continuation rays, then the list's shadow rays (any-hit lanes beside
    // closest-hit lanes) and traced MIS-BSDF rays, in one launch
    TraceArgs t = queue_trace_args(a.sc, a.q[qc], a.qcount + 2 * qc, true);
    t.root = S.rootq[qc].as<int32_t>();
    t.work = bt.ctr + 2;
    t.spill = a.sc.max_stack > sh.cfg.stack ? S.spill.as<int32_t>() : nullptr;
    t.stats = bt.stats;
    HIP_TRY(ctx, rec(SK_CLOSEST, 0));
    launch_trace_cfg(sh.cfg, t, TRACE_MIXED, sh.tblocks, stream);
    HIP_TRY(ctx, rec(SK_CLOSEST, 1));
    HIP_TRY(ctx, rec(SK_ANY, 0));   // no launch of its own: an empty interval
    HIP_TRY(ctx, rec(SK_ANY, 1));
    HIP_TRY(ctx, rec(SK_SHADE, 0));
    if (g) hipLaunchKernelGGL(k_shade<SHADE_BATCH_ADJ>, dim3(sh.sblocks), dim3(256), 0, stream, a, *g, bt, qc);
    else hipLaunchKernelGGL(k_shade<SHADE_BATCH>, dim3(sh.sblocks), dim3(256), 0, stream, a, AdjArgs{}, bt, qc);
    HIP_TRY(ctx, rec(SK_SHADE, 1));
    HIP_TRY(ctx, hipGetLastError());
    return MPT_OK;
}

}  // namespace

namespace mpt {

int stream_render(mpt_ctx* ctx, hipStream_t stream, Streamer& S, const StreamJob& job, mpt_stats* out) {
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    const int32_t cap = 1 << job.capacity_log2;
    if ((rc = ensure_stream_buffers(ctx, S, cap)) != MPT_OK) return rc;
    // the views buffer keeps its address while its size does (a recorded graph holds it)
    const size_t vbytes = sizeof(BatchView) * static_cast<size_t>(job.nviews);
    if (S.views.bytes != vbytes) {
        HIP_TRY(ctx, S.views.alloc(vbytes));
        if (S.exec) {
            (void)hipGraphExecDestroy(S.exec);
            S.exec = nullptr;
        }
    }
    HIP_TRY(ctx, hipMemcpyAsync(S.views.p, job.views, vbytes, hipMemcpyHostToDevice, stream));
    StepShape sh;
    sh.cfg = job.cfg;
    {
        const int64_t want = (3 * static_cast<int64_t>(cap) + sh.cfg.block - 1) / sh.cfg.block;
        sh.tblocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(want, static_cast<int64_t>(ctx->num_cus) * sh.cfg.bpc)));
        sh.sblocks = static_cast<int>(std::min<int64_t>((cap + 255) / 256, static_cast<int64_t>(ctx->num_cus) * 8));
        sh.rblocks = (cap + 255) / 256;
        const int64_t extra = job.sc.max_stack - sh.cfg.stack;
        const int64_t lanes = static_cast<int64_t>(sh.tblocks) * sh.cfg.block;
        if (extra > 0 && S.spill.bytes < sizeof(int32_t) * static_cast<size_t>(lanes * extra)) {
            HIP_TRY(ctx, S.spill.alloc(sizeof(int32_t) * static_cast<size_t>(lanes * extra)));
            S.spill_lanes = lanes;
            if (S.exec) {
                (void)hipGraphExecDestroy(S.exec);
                S.exec = nullptr;
            }
        }
    }
    PassArgs a{};
    a.sc = job.sc;
    a.path = nullptr;   // the path state travels in the queues
    for (int q = 0; q < 2; ++q) a.q[q] = queue_view(S.qbuf[q]);
    a.qcount = S.qcount.as<int32_t>();
    a.npaths = cap;
    BatchArgs bt{};
    bt.views = S.views.as<BatchView>();
    bt.nviews = job.nviews;
    bt.capacity = cap;
    bt.total_work = job.total_work;
    bt.root[0] = S.rootq[0].as<int32_t>();
    bt.root[1] = S.rootq[1].as<int32_t>();
    bt.film = job.film;
    bt.keep = job.keep;
    bt.ctr = S.ctr.as<int32_t>();
    bt.work_next = S.work_next.as<unsigned long long>();
    bt.stats = S.stats.as<unsigned long long>();
    a.stats = bt.stats;
    const int steps = job.graph_steps;
    AdjArgs g = job.adj;
    const AdjArgs* gp = nullptr;
    if (job.adjoint) {
        if (job.use_graph) return fail(ctx, MPT_ERR_INVALID, "the adjoint runs without a recorded graph");
        if (cap > S.adj_alloc) {
            S.adj_alloc = 0;   // until every buffer below has its new size
            const size_t sizes[7] = {8, 16, 16, 16, 16, 16, 16};
            for (int q = 0; q < 2; ++q)
                for (int k = 0; k < 7; ++k) HIP_TRY(ctx, S.adjq[q][k].alloc(sizes[k] * static_cast<size_t>(cap)));
            S.adj_alloc = cap;
        }
        for (int q = 0; q < 2; ++q) {
            g.q[q].par = S.adjq[q][0].as<int2>();
            g.q[q].sL = S.adjq[q][1].as<float4>();
            g.q[q].eL = S.adjq[q][2].as<float4>();
            g.q[q].sB = S.adjq[q][3].as<float4>();
            g.q[q].eB = S.adjq[q][4].as<float4>();
            g.q[q].sC = S.adjq[q][5].as<float4>();
            g.q[q].A = S.adjq[q][6].as<float4>();
        }
        g.A = nullptr;
        bt.film = nullptr;
        bt.keep = nullptr;
        gp = &g;
    }

    if (job.use_graph) {
        // re-record when anything the graph holds by value changed
        std::vector<uint8_t> key(sizeof(PassArgs) + sizeof(BatchArgs) + sizeof(StepShape) + sizeof(int));
        std::memcpy(key.data(), &a, sizeof(PassArgs));
        std::memcpy(key.data() + sizeof(PassArgs), &bt, sizeof(BatchArgs));
        std::memcpy(key.data() + sizeof(PassArgs) + sizeof(BatchArgs), &sh, sizeof(StepShape));
        std::memcpy(key.data() + sizeof(PassArgs) + sizeof(BatchArgs) + sizeof(StepShape), &steps, sizeof(int));
        if (!S.exec || key != S.exec_key) {
            if (S.exec) (void)hipGraphExecDestroy(S.exec);
            S.exec = nullptr;
            hipGraph_t graph = nullptr;
            HIP_TRY(ctx, hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            int erc = MPT_OK;
            for (int k = 0; k < steps && erc == MPT_OK; ++k)
                erc = enqueue_step(ctx, stream, S, sh, a, bt, nullptr, k & 1, nullptr);
            const hipError_t ec = hipStreamEndCapture(stream, &graph);
            if (erc != MPT_OK) return erc;
            if (ec != hipSuccess) return fail(ctx, MPT_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
            const hipError_t ei = hipGraphInstantiate(&S.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) return fail(ctx, MPT_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ei));
            S.exec_key = key;
        }
    }
    const size_t nev = static_cast<size_t>(2 + (job.timing ? 2 * SK_KINDS * steps : 0));
    while (S.events.size() < nev) {
        hipEvent_t e;
        HIP_TRY(ctx, hipEventCreate(&e));
        S.events.push_back(e);
    }
    // reset: counters, film
    hipLaunchKernelGGL(k_batch_init, dim3(1), dim3(64), 0, stream, bt.ctr, a.qcount, bt.work_next, bt.stats);
    if (job.clear_film && !job.adjoint)
        hipLaunchKernelGGL(k_zero_f4, dim3((job.film_px + 255) / 256), dim3(256), 0, stream, job.film, job.film_px);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipEventRecord(S.events[0], stream));
    // every path in flight finishes within max-depth + 2 steps, after which its
    // slot is refilled: ceil(work / cap) + 1 such windows bound the step count
    const int64_t bound = (job.total_work / cap + 2) * static_cast<int64_t>(job.max_depth + 2) + 2;
    const int64_t max_chunks = 2 * (bound / steps + 2);
    int64_t chunks = 0;
    double ms[SK_KINDS] = {0, 0, 0, 0};
    while (true) {
        if (job.use_graph) {
            HIP_TRY(ctx, hipGraphLaunch(S.exec, stream));
        } else {
            for (int k = 0; k < steps; ++k) {
                hipEvent_t* ev = job.timing ? &S.events[2 + 2 * SK_KINDS * k] : nullptr;
                if ((rc = enqueue_step(ctx, stream, S, sh, a, bt, gp, k & 1, ev)) != MPT_OK) return rc;
            }
        }
        HIP_TRY(ctx, hipMemcpyAsync(S.done_host, bt.ctr + 1, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIP_TRY(ctx, hipStreamSynchronize(stream));
        ++chunks;
        if (job.timing && !job.use_graph) {
            for (int k = 0; k < steps; ++k)
                for (int kind = 0; kind < SK_KINDS; ++kind) {
                    float t = 0.f;
                    const size_t e = 2 + 2 * SK_KINDS * static_cast<size_t>(k) + 2 * kind;
                    HIP_TRY(ctx, hipEventElapsedTime(&t, S.events[e], S.events[e + 1]));
                    ms[kind] += t;
                }
        }
        if (*S.done_host) break;
        if (chunks > max_chunks) return fail(ctx, MPT_ERR_DEVICE, "streaming renderer did not drain");
    }
    HIP_TRY(ctx, hipEventRecord(S.events[1], stream));
    HIP_TRY(ctx, hipEventSynchronize(S.events[1]));
    float gms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&gms, S.events[0], S.events[1]));
    unsigned long long st[ST_COUNT];
    HIP_TRY(ctx, hipMemcpy(st, S.stats.p, sizeof(st), hipMemcpyDeviceToHost));
    mpt_stats r{};
    r.samples = job.total_work;
    for (int32_t i = 0; i < job.nviews; ++i)
        if (job.views[i].max_depth > 0)
            r.camera_rays += static_cast<int64_t>(job.views[i].img_w) * job.views[i].img_h * job.views[i].nsamples;
    r.continuation_rays = static_cast<int64_t>(st[ST_CLOSEST_RAYS]) - r.camera_rays;
    r.mis_rays = static_cast<int64_t>(st[ST_MIS_RAYS]);
    r.mis_resolved = static_cast<int64_t>(st[ST_MIS_RESOLVED]);
    r.shadow_resolved = static_cast<int64_t>(st[ST_SHADOW_RESOLVED]);
    r.shadow_rays = static_cast<int64_t>(st[ST_SHADOW_RAYS]);
    r.box_tests = static_cast<int64_t>(st[ST_BOXES]);
    r.tri_tests = static_cast<int64_t>(st[ST_PRIMS]);
    r.box_tests_any = static_cast<int64_t>(st[ST_BOXES_ANY]);
    r.tri_tests_any = static_cast<int64_t>(st[ST_PRIMS_ANY]);
    r.closest_launches = chunks * steps;   // mixed launches: their shadow rays need no launch of their own
    r.anyhit_launches = 0;
    r.ms_camera = ms[SK_REGEN];
    r.ms_closest = ms[SK_CLOSEST];
    r.ms_anyhit = ms[SK_ANY];
    r.ms_shade = ms[SK_SHADE];
    r.ms_gpu = gms;
    r.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    r.path_slots = cap;
    if (out) *out = r;
    if (st[ST_ERRORS]) return fail(ctx, MPT_ERR_DEVICE, "traversal stack overrun");
    return MPT_OK;
}

}  // namespace mpt

namespace {

int do_batch_render(mpt_batch* B, const uint64_t* seeds, int32_t spp_override, double* gpu_ms) {
    mpt_ctx* ctx = B->ctx;
    mpt_batch_opts& o = B->opts;
    normalize_opts(o);
    const int wi = o.width == 8 ? 1 : 0;
    int rc;
    if ((rc = merge_nodes(B, wi)) != MPT_OK) return rc;
    int64_t work = 0;
    int32_t maxd = 0;
    for (size_t i = 0; i < B->scenes.size(); ++i) {
        BatchView& v = B->hviews[i];
        const HostScene& h = B->scenes[i]->host;
        v.root = B->roots[wi][i];
        const uint64_t sd = seeds ? seeds[i] : 0;
        v.key_lo = static_cast<uint32_t>(sd);
        v.key_hi = static_cast<uint32_t>(sd >> 32);
        v.spp = spp_override > 0 ? spp_override : h.spp;
        v.strata = strata_base(v.spp);
        v.sample_base = 0;
        v.nsamples = v.spp;
        v.work_base = work;
        work += static_cast<int64_t>(h.width) * h.height * v.nsamples;
        maxd = std::max(maxd, h.max_depth);
    }
    StreamJob job;
    job.sc = B->sc;
    job.sc.nodes = B->nodes[wi].p;
    job.sc.width = o.width;
    job.sc.nnodes = B->nnodes[wi];
    job.sc.max_stack = B->max_stack[wi];
    job.views = B->hviews.data();
    job.nviews = static_cast<int32_t>(B->hviews.size());
    job.total_work = work;
    job.film = B->film.as<float4>();
    job.film_px = B->total_px;
    job.clear_film = true;
    job.cfg.width = o.width;
    job.cfg.stack = o.stack_lds;
    job.cfg.block = o.block;
    job.cfg.bpc = o.blocks_per_cu;
    job.capacity_log2 = o.capacity_log2;
    job.graph_steps = o.graph_steps;
    job.use_graph = true;
    job.max_depth = maxd;
    if ((rc = stream_render(ctx, B->stream, B->st, job, &B->rstats)) != MPT_OK) return rc;
    if (gpu_ms) *gpu_ms = B->rstats.ms_gpu;
    return MPT_OK;
}

}  // namespace

extern "C" {

int mpt_batch_create(mpt_ctx* ctx, mpt_scene* const* scenes, int32_t n, const mpt_batch_opts* opts, mpt_batch** out) {
    if (!ctx || !scenes || n <= 0 || !out) return fail(ctx, MPT_ERR_INVALID, "null argument");
    if (ctx->device < 0) return fail(ctx, MPT_ERR_DEVICE, "host-only context has no device");
    std::unique_ptr<mpt_batch> B(new mpt_batch());
    B->ctx = ctx;
    for (int32_t i = 0; i < n; ++i) {
        if (!scenes[i] || scenes[i]->ctx != ctx) return fail(ctx, MPT_ERR_INVALID, "scene of another context");
        B->scenes.push_back(scenes[i]);
    }
    if (opts) B->opts = *opts;
    normalize_opts(B->opts);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    int rc;
    if ((rc = merge_tables(B.get())) != MPT_OK) return rc;
    if ((rc = merge_nodes(B.get(), B->opts.width == 8 ? 1 : 0)) != MPT_OK) return rc;
    *out = B.release();
    return MPT_OK;
}

void mpt_batch_destroy(mpt_batch* b) { delete b; }

int mpt_batch_render(mpt_batch* b, const uint64_t* seeds, int32_t spp_override, mpt_stats* stats) {
    if (!b) return MPT_ERR_INVALID;
    HIP_TRY(b->ctx, hipSetDevice(b->ctx->device));
    const int rc = do_batch_render(b, seeds, spp_override, nullptr);
    if (rc == MPT_OK && stats) *stats = b->rstats;
    return rc;
}

int mpt_batch_film(mpt_batch* b, int32_t i, float* film_host) {
    if (!b || !film_host || i < 0 || i >= static_cast<int32_t>(b->scenes.size())) return MPT_ERR_INVALID;
    HIP_TRY(b->ctx, hipSetDevice(b->ctx->device));
    const BatchView& v = b->hviews[i];
    HIP_TRY(b->ctx, hipMemcpy(film_host, b->film.as<float4>() + v.film_base,
                              sizeof(float4) * static_cast<size_t>(v.img_w) * v.img_h, hipMemcpyDeviceToHost));
    return MPT_OK;
}

int mpt_batch_get_opts(const mpt_batch* b, mpt_batch_opts* out) {
    if (!b || !out) return MPT_ERR_INVALID;
    *out = b->opts;
    return MPT_OK;
}

int mpt_batch_set_opts(mpt_batch* b, const mpt_batch_opts* o) {
    if (!b || !o) return MPT_ERR_INVALID;
    b->opts = *o;
    normalize_opts(b->opts);
    return MPT_OK;
}

int mpt_batch_autotune(mpt_batch* b, const uint64_t* seeds, int32_t trial_spp, mpt_batch_opts* best,
                       mpt_batch_trial* trials, int32_t max_trials, int32_t* ntrials) {
    if (!b) return MPT_ERR_INVALID;
    HIP_TRY(b->ctx, hipSetDevice(b->ctx->device));
    if (trial_spp <= 0) trial_spp = 1;
    int32_t nt = 0;
    mpt_batch_opts cur = b->opts;
    normalize_opts(cur);
    double best_ms = 1e30;
    auto trial = [&](const mpt_batch_opts& o, double* ms) -> int {
        b->opts = o;
        double t = 0.0;
        // the first run of a shape records its graph; time the second
        int rc = do_batch_render(b, seeds, trial_spp, &t);
        if (rc == MPT_OK) rc = do_batch_render(b, seeds, trial_spp, &t);
        if (rc != MPT_OK) return rc;
        *ms = t;
        if (trials && nt < max_trials) trials[nt] = mpt_batch_trial{b->opts, t};
        ++nt;
        return MPT_OK;
    };
    int rc;
    if ((rc = trial(cur, &best_ms)) != MPT_OK) return rc;
    // coordinate descent: paths in flight, BVH width, LDS stack entries, block size
    for (int pass = 0; pass < 2; ++pass) {
        bool improved = false;
        for (int dim = 0; dim < 5; ++dim) {
            std::vector<mpt_batch_opts> cands;
            if (dim == 0)
                for (int c : {19, 20, 21, 22, 23, 24}) { mpt_batch_opts o = cur; o.capacity_log2 = c; cands.push_back(o); }
            if (dim == 1)
                for (int w : {4, 8}) { mpt_batch_opts o = cur; o.width = w; cands.push_back(o); }
            if (dim == 2)
                for (int s : {16, 20, 24, 32}) { mpt_batch_opts o = cur; o.stack_lds = s; cands.push_back(o); }
            if (dim == 3)
                for (int k : {128, 256}) { mpt_batch_opts o = cur; o.block = k; cands.push_back(o); }
            if (dim == 4)
                for (int k : {4, 6, 7, 8, 12}) { mpt_batch_opts o = cur; o.blocks_per_cu = k; cands.push_back(o); }
            for (const mpt_batch_opts& o : cands) {
                if (std::memcmp(&o, &cur, sizeof(o)) == 0) continue;
                double ms = 0.0;
                if ((rc = trial(o, &ms)) != MPT_OK) return rc;
                if (ms < best_ms * 0.98) {
                    best_ms = ms;
                    cur = b->opts;
                    improved = true;
                }
            }
        }
        if (!improved) break;
    }
    b->opts = cur;
    if (best) *best = cur;
    if (ntrials) *ntrials = nt;
    return MPT_OK;
}

}  // extern "C"
