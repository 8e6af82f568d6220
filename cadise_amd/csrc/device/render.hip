// Host driver of the device path and the C ABI of include/mpt.h.
//
// DeviceScene owns every HBM allocation of one scene: the wide BVH, the
// leaf-ordered primitive records, shading tables, the film, and the per-pass
// path state / ping-pong vertex queues. render() runs, per sample pass,
//   k_camera -> [k_trace closest (continuation + MIS-BSDF rays),
//                k_trace any (shadow rays), k_shade] x (max_depth + 1) -> k_film
// on one stream; queue sizes stay on the device, so the host never waits
// inside a pass.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "driver.hpp"
#include "../host/picture.hpp"

namespace {

using namespace mpt;


// Build the BVHs and upload every scene table.
int build_and_upload(mpt_scene* S) {
    mpt_ctx* ctx = S->ctx;
    const HostScene& h = S->host;
    if (h.nprims() > MPT_MAX_PRIMITIVES)
        return fail(ctx, MPT_ERR_UNSUPPORTED, "more than MPT_MAX_PRIMITIVES primitives");
    const bool host_only = ctx->device < 0;
    if (!host_only) {
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        if (!S->stream) HIP_TRY(ctx, hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
    }
    mpt_bvh_opts& o = S->opts;
    S->cfg = normalize_cfg(o);

    auto t0 = std::chrono::steady_clock::now();
    build_binary(h, false, &S->bin);
    S->sstats.build_ms = ms_since(t0);
    t0 = std::chrono::steady_clock::now();
    collapse_wide(h, S->bin, o.width, o.leaf_max, &S->wide);
    S->sstats.collapse_ms = ms_since(t0);
    if (!S->wide.ok) return fail(ctx, MPT_ERR_UNSUPPORTED, "a BVH node cannot be encoded (a leaf of more than 65535 "
                                                          "coincident primitives, or coordinates beyond binary32)");

    const int64_t n = h.nprims();
    S->sstats.nprims = n;
    S->sstats.nlights = static_cast<int64_t>(h.lights.size());
    S->sstats.binary_nodes = static_cast<int64_t>(S->bin.nodes.size());
    S->sstats.wide_nodes = S->wide.nnodes;
    S->sstats.binary_depth = S->bin.depth;
    S->sstats.wide_depth = S->wide.depth;
    S->sstats.width = S->wide.width;
    S->sstats.leaf_max = o.leaf_max;
    S->sstats.stack_lds = S->cfg.stack;
    S->sstats.trace_block = S->cfg.block;
    S->sstats.blocks_per_cu = S->cfg.bpc;
    S->sstats.image_width = h.width;
    S->sstats.image_height = h.height;
    S->sstats.spp = h.spp;
    S->sstats.max_depth = h.max_depth;
    S->sstats.ntextures = static_cast<int64_t>(h.textures.size());
    S->sstats.tex_elements = 0;
    for (const Texture& t : h.textures)
        S->sstats.tex_elements += (t.data.empty() ? 1 : static_cast<int64_t>(t.width) * t.height) * t.channels;
    if (host_only) return MPT_OK;
    // leaf-ordered primitive records
    std::vector<float4> tris(static_cast<size_t>(3 * n));
    for (int64_t j = 0; j < n; ++j) {
        const int32_t p = S->bin.ordered[j];
        const float* v = &h.prim_v[9 * p];
        if (h.prim_type[p] == PRIM_SPHERE) {
            const uint32_t bits = static_cast<uint32_t>(p) | 0x80000000u;
            float w;
            std::memcpy(&w, &bits, 4);
            tris[3 * j + 0] = make_float4(v[0], v[1], v[2], w);
            tris[3 * j + 1] = make_float4(v[3], 0.f, 0.f, 1.f);
            tris[3 * j + 2] = make_float4(0.f, 0.f, 0.f, 0.f);
            continue;
        }
        float A[3], e1[3], e2[3];
        const bool rect = h.prim_type[p] == PRIM_RECTANGLE;
        for (int k = 0; k < 3; ++k) {
            if (rect) {
                A[k] = v[3 + k];
                e1[k] = v[k] - v[3 + k];
                e2[k] = v[6 + k] - v[3 + k];
            } else {
                A[k] = v[k];
                e1[k] = v[3 + k] - v[k];
                e2[k] = v[6 + k] - v[k];
            }
        }
        uint32_t bits = static_cast<uint32_t>(p) | (rect ? 0x80000000u : 0u);
        float w;
        std::memcpy(&w, &bits, 4);
        tris[3 * j + 0] = make_float4(A[0], A[1], A[2], w);
        tris[3 * j + 1] = make_float4(e1[0], e1[1], e1[2], 0.f);
        tris[3 * j + 2] = make_float4(e2[0], e2[1], e2[2], 0.f);
    }
    // per-primitive shading info; a primitive's area light is the last one bound to it
    std::vector<int4> info(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        const bool has_n = !h.prim_n.empty() && h.prim_type[i] == PRIM_TRIANGLE;
        const bool has_u = !h.prim_uvw.empty() && h.prim_type[i] == PRIM_TRIANGLE && !std::isnan(h.prim_uvw[9 * i]);
        info[i] = make_int4(h.prim_mat[i], -1, has_n ? 1 : 0, has_u ? 1 : 0);
    }
    for (size_t l = 0; l < h.lights.size(); ++l)
        if (h.lights[l].type == LIGHT_AREA) info[h.lights[l].prim].y = static_cast<int32_t>(l);
    // the record's last word: the primitive's area light (-1 none), so that a hit's
    // light is read from the record the hit already names
    for (int64_t j = 0; j < n; ++j) {
        uint32_t bits;
        std::memcpy(&bits, &tris[3 * j].w, 4);
        const int32_t l = info[bits & 0x7FFFFFFFu].y;
        std::memcpy(&tris[3 * j + 2].w, &l, 4);
    }

    std::vector<DevMaterial> mats(h.materials.size());
    for (size_t i = 0; i < mats.size(); ++i) {
        const Material& m = h.materials[i];
        DevMaterial d{};
        d.type = m.type;
        d.albedo_tex = m.albedo_tex;
        d.rough_tex = m.rough_tex;
        d.albedo[0] = m.albedo.x; d.albedo[1] = m.albedo.y; d.albedo[2] = m.albedo.z;
        d.f0[0] = m.f0.x; d.f0[1] = m.f0.y; d.f0[2] = m.f0.z;
        d.roughness = m.roughness;
        d.fresnel = m.fresnel;
        d.exponent = m.exponent;
        d.ratio = m.ratio;
        d.ior_outer = m.ior_outer;
        d.ior_inner = m.ior_inner;
        mats[i] = d;
    }
    std::vector<DevTexture> texs(h.textures.size());
    std::vector<float> texdata;
    for (size_t i = 0; i < texs.size(); ++i) {
        const Texture& t = h.textures[i];
        DevTexture d{};
        d.offset = static_cast<int64_t>(texdata.size());
        d.width = t.width;
        d.height = t.height;
        d.channels = t.channels;
        d.wrap_clamp = t.wrap_clamp;
        d.odd = t.odd;
        d.even = t.even;
        d.odd_size = t.odd_size;
        d.even_size = t.even_size;
        if (t.data.empty()) {
            // a constant texture as a 1x1 image: u * 1 and v * 1 floor to 0
            d.width = d.height = 1;
            texdata.push_back(t.value.x);
            if (t.channels == 3) { texdata.push_back(t.value.y); texdata.push_back(t.value.z); }
        } else {
            texdata.insert(texdata.end(), t.data.begin(), t.data.end());
        }
        texs[i] = d;
    }
    const std::vector<F3> rad = light_radiance(h);
    // leaf position of each primitive's record (the records follow the binary BVH's order)
    std::vector<int32_t> leaf_pos(S->bin.ordered.size());
    for (size_t j = 0; j < S->bin.ordered.size(); ++j) leaf_pos[static_cast<size_t>(S->bin.ordered[j])] = static_cast<int32_t>(j);
    std::vector<DevLight> lights(h.lights.size());
    std::vector<float> envcdf;
    int32_t env_light = -1;
    for (size_t l = 0; l < lights.size(); ++l) {
        DevLight d{};
        if (h.lights[l].type == LIGHT_ENVIRONMENT) {
            const Texture& t = h.textures[h.lights[l].env_tex];
            d.prim = -1;
            d.rec = -1;
            d.type = LIGHT_TYPE_ENV;
            d.env_tex = h.lights[l].env_tex;
            d.env_w = t.width;
            d.env_h = t.height;
            d.env_cdf = static_cast<int32_t>(envcdf.size());
            const EnvTables e = env_tables(t);
            envcdf.insert(envcdf.end(), e.cdf.begin(), e.cdf.end());
            env_light = static_cast<int32_t>(l);
            lights[l] = d;
            continue;
        }
        const int32_t p = h.lights[l].prim;
        const float* v = &h.prim_v[9 * p];
        d.prim = p;
        d.rec = leaf_pos[static_cast<size_t>(p)];
        d.type = h.prim_type[p];
        d.back_face = h.lights[l].back_face;
        d.radiance[0] = rad[l].x; d.radiance[1] = rad[l].y; d.radiance[2] = rad[l].z;
        d.pdfA = 1.0f / prim_area(h, p);
        for (int k = 0; k < 3; ++k) {
            if (d.type == PRIM_RECTANGLE) {
                d.a[k] = v[3 + k];
                d.e1[k] = v[k] - v[3 + k];
                d.e2[k] = v[6 + k] - v[3 + k];
            } else {
                d.a[k] = v[k];
                d.e1[k] = v[3 + k] - v[k];
                d.e2[k] = v[6 + k] - v[k];
            }
        }
        d.has_normals = 0;
        if (d.type == PRIM_TRIANGLE && !h.prim_n.empty()) {
            // face normal unless a non-zero vertex normal was given (Triangle::setNa..c)
            const float cx = d.e1[1] * d.e2[2] - d.e1[2] * d.e2[1];
            const float cy = d.e1[2] * d.e2[0] - d.e1[0] * d.e2[2];
            const float cz = d.e1[0] * d.e2[1] - d.e1[1] * d.e2[0];
            float N[3] = {0.f, 1.f, 0.f};
            if (!(cx == 0.f && cy == 0.f && cz == 0.f)) {
                float l2 = 0.f;
                l2 += cx * cx;
                l2 += cy * cy;
                l2 += cz * cz;
                const float inv = 1.f / std::sqrt(l2);
                N[0] = cx * inv; N[1] = cy * inv; N[2] = cz * inv;
            }
            for (int q = 0; q < 3; ++q) {
                const float* vn = &h.prim_n[9 * p + 3 * q];
                const bool zero = vn[0] == 0.f && vn[1] == 0.f && vn[2] == 0.f;
                for (int k = 0; k < 3; ++k) d.n[3 * q + k] = zero ? N[k] : vn[k];
            }
            d.has_normals = 1;
        }
        lights[l] = d;
    }
    const std::vector<float> cdf = power_cdf(h, rad);

    int rc;
    if ((rc = upload_vec(ctx, S->nodes, S->wide.nodes)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, S->tris, tris)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, S->prim_info, info)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, S->prim_n, h.prim_n)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, S->prim_uvw, h.prim_uvw)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, S->mats, mats)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, S->texs, texs)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, S->tex_data, texdata)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, S->lights, lights)) != MPT_OK) return rc;
    if ((rc = upload_vec(ctx, S->env_cdf, envcdf)) != MPT_OK) return rc;
    S->h_envcdf = envcdf;
    S->h_lights = lights;
    S->h_mats = mats;
    S->h_cdf = cdf;
    S->h_texs = texs;
    S->h_rad = rad;
    S->tex_elems = static_cast<int64_t>(texdata.size());
    if ((rc = upload_vec(ctx, S->power_cdf, cdf)) != MPT_OK) return rc;
    HIP_TRY(ctx, S->stats.alloc(sizeof(unsigned long long) * ST_COUNT));

    DevScene& sc = S->sc;
    sc = DevScene{};
    sc.nodes = S->nodes.p;
    sc.width = S->wide.width;
    sc.max_stack = S->wide.max_stack;
    sc.nnodes = S->wide.nnodes;
    sc.nprims = n;
    sc.tris = S->tris.as<float4>();
    sc.prim_info = S->prim_info.as<int4>();
    sc.prim_n = S->prim_n.as<float>();
    sc.prim_uvw = S->prim_uvw.as<float>();
    sc.mats = S->mats.as<DevMaterial>();
    sc.texs = S->texs.as<DevTexture>();
    sc.tex_data = S->tex_data.as<float>();
    sc.lights = S->lights.as<DevLight>();
    sc.power_cdf = S->power_cdf.as<float>();
    sc.nlights = static_cast<int32_t>(h.lights.size());
    sc.env_light = env_light;
    sc.estimator = h.estimator;
    sc.env_cdf = S->env_cdf.as<float>();
    sc.cluster = h.cluster;
    float f2c[16], c2w[16];
    camera_matrices(h, f2c, c2w);
    std::memcpy(sc.f2c, f2c, sizeof(sc.f2c));
    std::memcpy(sc.c2w, c2w, sizeof(sc.c2w));
    sc.cam_pos[0] = h.cam_pos.x; sc.cam_pos[1] = h.cam_pos.y; sc.cam_pos[2] = h.cam_pos.z;
    sc.img_w = h.width;
    sc.img_h = h.height;
    sc.tile_x = h.tile_x;
    sc.tile_y = h.tile_y;
    sc.filter = h.filter;
    sc.filter_hw = h.filter_w * 0.5f;   // Filter::filterHalfSize: size / 2
    sc.filter_hh = h.filter_h * 0.5f;
    sc.g_expx = -1.0f / (2.0f * h.sigma_x * h.sigma_x);
    sc.g_expy = -1.0f / (2.0f * h.sigma_y * h.sigma_y);
    sc.amplitude = h.amplitude;
    sc.g_edge = h.amplitude * std::exp(sc.g_expx * sc.filter_hw * sc.filter_hw + sc.g_expy * sc.filter_hh * sc.filter_hh);
    sc.mitchell_b = h.mitchell_b;
    sc.mitchell_c = h.mitchell_c;
    sc.sampler = h.sampler;
    sc.spp = h.spp;
    sc.strata = strata_base(h.spp);

    const int64_t npix = static_cast<int64_t>(h.width) * h.height;
    HIP_TRY(ctx, S->film.alloc(sizeof(float4) * static_cast<size_t>(npix)));
    HIP_TRY(ctx, hipMemset(S->film.p, 0, S->film.bytes));

    int64_t bytes = 0;
    for (DevBuf* b : {&S->nodes, &S->tris, &S->prim_info, &S->prim_n, &S->prim_uvw, &S->mats, &S->texs, &S->tex_data,
                      &S->lights, &S->power_cdf, &S->film})
        bytes += static_cast<int64_t>(b->bytes);
    S->sstats.device_bytes = bytes;
    return MPT_OK;
}

int trace_blocks(mpt_scene* S, int64_t nrays) {
    const int64_t want = (nrays + S->cfg.block - 1) / S->cfg.block;
    const int64_t cap = static_cast<int64_t>(S->ctx->num_cus) * S->cfg.bpc;
    return static_cast<int>(std::max<int64_t>(1, std::min(want, cap)));
}

// Spill space for traversal stacks deeper than the LDS part.
int ensure_spill(mpt_scene* S, int blocks) {
    const int64_t extra = S->sc.max_stack - S->cfg.stack;
    if (extra <= 0) return MPT_OK;
    const int64_t lanes = static_cast<int64_t>(blocks) * S->cfg.block;
    if (lanes <= S->spill_lanes) return MPT_OK;
    S->spill_lanes = 0;
    HIP_TRY(S->ctx, S->spill.alloc(sizeof(int32_t) * static_cast<size_t>(lanes * extra)));
    S->spill_lanes = lanes;
    return MPT_OK;
}

int launch_trace(mpt_scene* S, TraceArgs a, int mode, int blocks) {
    a.spill = S->sc.max_stack > S->cfg.stack ? S->spill.as<int32_t>() : nullptr;
    launch_trace_cfg(S->cfg, a, mode, blocks, S->stream);
    HIP_TRY(S->ctx, hipGetLastError());
    return MPT_OK;
}

int read_stats(mpt_scene* S, unsigned long long* out) {
    HIP_TRY(S->ctx, hipMemcpyAsync(out, S->stats.p, sizeof(unsigned long long) * ST_COUNT, hipMemcpyDeviceToHost, S->stream));
    HIP_TRY(S->ctx, hipStreamSynchronize(S->stream));
    return MPT_OK;
}

int zero_stats(mpt_scene* S) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, S->stream, S->stats.as<unsigned long long>(), ST_COUNT);
    HIP_TRY(S->ctx, hipGetLastError());
    return MPT_OK;
}

// want_prim: hits carry the primitive's creation id (else its leaf position).
int do_trace(mpt_scene* S, const mpt_ray* rays_dev, size_t n, int any_hit, mpt_hit* hits_dev, hipStream_t stream,
             int want_prim = 1) {
    if (n == 0) return MPT_OK;
    hipStream_t keep = S->stream;
    if (stream) S->stream = stream;
    int rc = MPT_OK;
    const int blocks = trace_blocks(S, static_cast<int64_t>(n));
    DevBuf work;
    do {
        if ((rc = ensure_spill(S, blocks)) != MPT_OK) break;
        if (work.alloc(sizeof(int32_t)) != hipSuccess) { rc = fail(S->ctx, MPT_ERR_DEVICE, "hipMalloc work counter"); break; }
        if (hipMemsetAsync(work.p, 0, sizeof(int32_t), S->stream) != hipSuccess) { rc = fail(S->ctx, MPT_ERR_DEVICE, "memset"); break; }
        if ((rc = zero_stats(S)) != MPT_OK) break;
        TraceArgs a{};
        a.sc = S->sc;
        a.org = reinterpret_cast<const float4*>(rays_dev);
        a.dir = reinterpret_cast<const float4*>(rays_dev) + 1;
        a.stride = 2;
        a.count = nullptr;
        a.n_host = static_cast<int64_t>(n);
        a.hit = reinterpret_cast<int2*>(hits_dev);
        a.want_prim = want_prim;
        a.work = work.as<int32_t>();
        a.stats = S->stats.as<unsigned long long>();
        a.stat_slot = any_hit ? ST_SHADOW_RAYS : ST_CLOSEST_RAYS;
        if ((rc = launch_trace(S, a, any_hit != 0 ? TRACE_ANY : TRACE_CLOSEST, blocks)) != MPT_OK) break;
        if (hipStreamSynchronize(S->stream) != hipSuccess) { rc = fail(S->ctx, MPT_ERR_DEVICE, "trace failed"); break; }
        unsigned long long st[ST_COUNT];
        if ((rc = read_stats(S, st)) != MPT_OK) break;
        S->rstats = mpt_stats{};
        S->rstats.box_tests = static_cast<int64_t>(st[any_hit ? ST_BOXES_ANY : ST_BOXES]);
        S->rstats.tri_tests = static_cast<int64_t>(st[any_hit ? ST_PRIMS_ANY : ST_PRIMS]);
        S->rstats.box_tests_any = static_cast<int64_t>(st[ST_BOXES_ANY]);
        S->rstats.tri_tests_any = static_cast<int64_t>(st[ST_PRIMS_ANY]);
        if (any_hit) S->rstats.shadow_rays = static_cast<int64_t>(st[ST_SHADOW_RAYS]);
        else S->rstats.camera_rays = static_cast<int64_t>(st[ST_CLOSEST_RAYS]);
        if (st[ST_ERRORS]) rc = fail(S->ctx, MPT_ERR_DEVICE, "traversal stack overrun");
    } while (false);
    S->stream = keep;
    return rc;
}

// Per-pass buffers for npix paths.
int ensure_pass_buffers(mpt_scene* S, int64_t npix, int32_t max_depth) {
    mpt_ctx* ctx = S->ctx;
    if (npix != S->npaths) {
        S->npaths = 0;   // until every buffer below has its new size
        HIP_TRY(ctx, S->path.alloc(sizeof(PathRec) * npix));
        for (int q = 0; q < 2; ++q)
            for (int k = 0; k < QBUFS; ++k) HIP_TRY(ctx, S->qbuf[q][k].alloc(kQueueBytes[k] * static_cast<size_t>(npix)));
        S->npaths = npix;
    }
    if (max_depth != S->max_depth_alloc) {
        S->max_depth_alloc = -1;
        HIP_TRY(ctx, S->qcount.alloc(sizeof(int32_t) * 2 * (max_depth + 2)));
        HIP_TRY(ctx, S->work.alloc(sizeof(int32_t) * 3 * (max_depth + 2)));
        S->max_depth_alloc = max_depth;
    }
    return MPT_OK;
}

PassArgs pass_args(mpt_scene* S, uint64_t seed, int32_t max_depth) {
    PassArgs a{};
    a.sc = S->sc;
    a.key = make_uint2(static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    a.max_depth = max_depth;
    a.npaths = static_cast<int32_t>(S->npaths);
    a.path = S->path.as<PathRec>();
    for (int q = 0; q < 2; ++q) a.q[q] = queue_view(S->qbuf[q]);
    a.qcount = S->qcount.as<int32_t>();
    a.work = S->work.as<int32_t>();
    a.nwork = 3 * (max_depth + 2);
    a.film = S->film.as<float4>();
    a.stats = S->stats.as<unsigned long long>();
    return a;
}

enum EvKind { EV_CLOSEST = 0, EV_ANY, EV_SHADE, EV_CAMERA, EV_FILM, EV_KINDS };

int do_render_passes(mpt_scene* S, const mpt_render_params* p) {
    mpt_ctx* ctx = S->ctx;
    const auto t0 = std::chrono::steady_clock::now();
    const HostScene& h = S->host;
    const int32_t max_depth = (p && p->max_depth > 0) ? p->max_depth : h.max_depth;
    const int32_t spp_begin = p ? std::max(0, p->spp_begin) : 0;
    const int32_t spp_end = (p && p->spp_end > 0) ? p->spp_end : h.spp;
    const uint64_t seed = p ? p->seed : 0;
    const int64_t npix = static_cast<int64_t>(h.width) * h.height;
    int rc;
    if ((rc = ensure_pass_buffers(S, npix, max_depth)) != MPT_OK) return rc;
    if (!p || !p->accumulate) {
        hipLaunchKernelGGL(k_zero_f4, dim3((npix + 255) / 256), dim3(256), 0, S->stream, S->film.as<float4>(), npix);
        HIP_TRY(ctx, hipGetLastError());
    }
    if ((rc = zero_stats(S)) != MPT_OK) return rc;
    const bool keep = p && (p->flags & MPT_RENDER_KEEP_SAMPLES);
    S->keep_valid = false;
    if (keep) {
        const size_t want = sizeof(float4) * static_cast<size_t>(npix) * static_cast<size_t>(std::max(1, spp_end - spp_begin));
        if (S->samples.bytes < want) HIP_TRY(ctx, S->samples.alloc(want));
    }
    const int tblocks = trace_blocks(S, 2 * npix);
    if ((rc = ensure_spill(S, tblocks)) != MPT_OK) return rc;
    const int sblocks = static_cast<int>(std::min<int64_t>((npix + 255) / 256, static_cast<int64_t>(ctx->num_cus) * 8));
    const int ntiles = ((h.width + h.tile_x - 1) / h.tile_x) * ((h.height + h.tile_y - 1) / h.tile_y);
    const int cam_threads = static_cast<int>(std::max<int64_t>(npix, 3 * (max_depth + 2) + 1));

    const int passes = std::max(0, spp_end - spp_begin);
    // events: (start, stop) per launch, grouped by kind, created as needed
    std::vector<std::pair<int, size_t>> marks;   // (kind, index of start event)
    size_t ev = 0;
    auto record = [&]() -> hipError_t {
        if (ev == S->events.size()) {
            hipEvent_t e;
            const hipError_t err = hipEventCreate(&e);
            if (err != hipSuccess) return err;
            S->events.push_back(e);
        }
        return hipEventRecord(S->events[ev++], S->stream);
    };
    auto mark = [&](int kind) -> hipError_t {
        marks.push_back({kind, ev});
        return record();
    };
    auto stop = [&]() -> hipError_t { return record(); };

    for (int32_t k = spp_begin; k < spp_end; ++k) {
        PassArgs a = pass_args(S, seed, max_depth);
        a.sample = k;
        HIP_TRY(ctx, mark(EV_CAMERA));
        hipLaunchKernelGGL(k_camera<false>, dim3((cam_threads + 255) / 256), dim3(256), 0, S->stream, a, AdjArgs{});
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, stop());
        for (int32_t b = 0; b <= max_depth; ++b) {
            // continuation rays, then (from bounce 1) the previous vertices' shadow
            // rays and traced MIS-BSDF rays from the queue's list
            TraceArgs t = queue_trace_args(S->sc, a.q[b & 1], a.qcount + 2 * b, false);
            t.work = a.work + 3 * b;
            t.stats = S->stats.as<unsigned long long>();
            HIP_TRY(ctx, mark(EV_CLOSEST));
            // bounce 0's list is empty: camera rays only, closest hit
            if ((rc = launch_trace(S, t, b > 0 ? TRACE_MIXED : TRACE_CLOSEST, tblocks)) != MPT_OK) return rc;
            HIP_TRY(ctx, stop());
            HIP_TRY(ctx, mark(EV_SHADE));
            hipLaunchKernelGGL(k_shade<SHADE_FWD>, dim3(sblocks), dim3(256), 0, S->stream, a, AdjArgs{}, BatchArgs{}, b);
            HIP_TRY(ctx, hipGetLastError());
            HIP_TRY(ctx, stop());
            // deep limits (the reference's default max-depth is 10000): every 16
            // bounces, stop once no path is left for the next one
            if ((b & 15) == 15 && b < max_depth) {
                int32_t left = 0;
                HIP_TRY(ctx, hipMemcpyAsync(&left, a.qcount + 2 * (b + 1), sizeof(int32_t), hipMemcpyDeviceToHost, S->stream));
                HIP_TRY(ctx, hipStreamSynchronize(S->stream));
                if (left == 0) break;
            }
        }
        HIP_TRY(ctx, mark(EV_FILM));
        hipLaunchKernelGGL(k_film, dim3(ntiles), dim3(256), 0, S->stream, a);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, stop());
        if (keep)
            HIP_TRY(ctx, hipMemcpy2DAsync(S->samples.as<float4>() + static_cast<int64_t>(k - spp_begin) * npix,
                                          sizeof(float4), S->path.p, sizeof(PathRec), sizeof(float4),
                                          static_cast<size_t>(npix), hipMemcpyDeviceToDevice, S->stream));
    }
    HIP_TRY(ctx, hipStreamSynchronize(S->stream));
    unsigned long long st[ST_COUNT];
    if ((rc = read_stats(S, st)) != MPT_OK) return rc;
    double ms[EV_KINDS] = {0, 0, 0, 0, 0};
    for (auto& m : marks) {
        float t = 0.f;
        HIP_TRY(ctx, hipEventElapsedTime(&t, S->events[m.second], S->events[m.second + 1]));
        ms[m.first] += t;
    }
    mpt_stats& r = S->rstats;
    r = mpt_stats{};
    r.samples = npix * passes;
    r.camera_rays = npix * (max_depth > 0 ? passes : 0);
    r.continuation_rays = static_cast<int64_t>(st[ST_CLOSEST_RAYS]) - r.camera_rays;
    r.mis_rays = static_cast<int64_t>(st[ST_MIS_RAYS]);
    r.mis_resolved = static_cast<int64_t>(st[ST_MIS_RESOLVED]);
    r.shadow_resolved = static_cast<int64_t>(st[ST_SHADOW_RESOLVED]);
    r.shadow_rays = static_cast<int64_t>(st[ST_SHADOW_RAYS]);
    r.box_tests = static_cast<int64_t>(st[ST_BOXES]);
    r.tri_tests = static_cast<int64_t>(st[ST_PRIMS]);
    r.box_tests_any = static_cast<int64_t>(st[ST_BOXES_ANY]);
    r.tri_tests_any = static_cast<int64_t>(st[ST_PRIMS_ANY]);
    r.closest_launches = static_cast<int64_t>(passes) * (max_depth + 1);
    r.anyhit_launches = 0;   // shadow rays ride in the closest-hit launches
    r.ms_closest = ms[EV_CLOSEST];
    r.ms_anyhit = ms[EV_ANY];
    r.ms_shade = ms[EV_SHADE];
    r.ms_camera = ms[EV_CAMERA];
    r.ms_film = ms[EV_FILM];
    r.ms_gpu = ms[EV_CLOSEST] + ms[EV_ANY] + ms[EV_SHADE] + ms[EV_CAMERA] + ms[EV_FILM];
    r.ms_total = ms_since(t0);
    if (st[ST_ERRORS]) return fail(ctx, MPT_ERR_DEVICE, "traversal stack overrun");
    if (keep && !(p->accumulate)) {
        S->keep_valid = true;
        S->keep_seed = seed;
        S->keep_begin = spp_begin;
        S->keep_end = spp_end;
        S->keep_depth = max_depth;
        S->keep_version = S->param_version;
    }
    return MPT_OK;
}

// The scene as the one view of a streaming render: samples
// [spp_begin, spp_begin + nsamples) with the given seed and depth.
BatchView scene_view(const mpt_scene* S, uint64_t seed, int32_t spp_begin, int32_t nsamples, int32_t max_depth) {
    BatchView v{};
    const DevScene& d = S->sc;
    v.root = 0;
    v.light_base = 0;
    v.nlights = d.nlights;
    v.cluster = d.cluster;
    v.cdf_base = 0;
    v.env_light = d.env_light;
    v.estimator = d.estimator;
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
    v.max_depth = max_depth;
    v.sample_base = spp_begin;
    v.nsamples = nsamples;
    v.key_lo = static_cast<uint32_t>(seed);
    v.key_hi = static_cast<uint32_t>(seed >> 32);
    v.film_base = 0;
    v.work_base = 0;
    std::memcpy(v.f2c, d.f2c, sizeof(v.f2c));
    std::memcpy(v.c2w, d.c2w, sizeof(v.c2w));
    std::memcpy(v.cam_pos, d.cam_pos, sizeof(v.cam_pos));
    return v;
}

// The default render: the scene's sample passes [begin, end) as one streaming
// wavefront (batch.hip) over the context's streaming buffers.
int do_render(mpt_scene* S, const mpt_render_params* p) {
    if (p && (p->flags & MPT_RENDER_PASSWISE)) return do_render_passes(S, p);
    mpt_ctx* ctx = S->ctx;
    const HostScene& h = S->host;
    const int32_t max_depth = (p && p->max_depth > 0) ? p->max_depth : h.max_depth;
    const int32_t spp_begin = p ? std::max(0, p->spp_begin) : 0;
    const int32_t spp_end = (p && p->spp_end > 0) ? p->spp_end : h.spp;
    const uint64_t seed = p ? p->seed : 0;
    const int64_t npix = static_cast<int64_t>(h.width) * h.height;
    const bool keep = p && (p->flags & MPT_RENDER_KEEP_SAMPLES);
    const int32_t nsamples = std::max(0, spp_end - spp_begin);
    S->keep_valid = false;
    if (keep) {
        const size_t want = sizeof(float4) * static_cast<size_t>(npix) * static_cast<size_t>(std::max(1, nsamples));
        if (S->samples.bytes < want) HIP_TRY(ctx, S->samples.alloc(want));
    }
    if (!ctx->streamer) ctx->streamer.reset(new Streamer());
    const BatchView v = scene_view(S, seed, spp_begin, nsamples, max_depth);
    StreamJob job;
    job.sc = S->sc;
    job.views = &v;
    job.nviews = 1;
    job.total_work = npix * nsamples;
    job.film = S->film.as<float4>();
    job.film_px = npix;
    job.clear_film = !(p && p->accumulate);
    job.keep = keep ? S->samples.as<float4>() : nullptr;
    job.cfg = S->cfg;
    int lg = 16;
    // up to 16M paths in flight (~8 GB of queues): larger traversal launches
    // amortize each launch's ramp and straggler tail (1080p, 32 passes: 99.8 /
    // 94.6 / 91.3 ms of traversal with 4M / 8M / 16M paths)
    while (lg < 24 && (int64_t{1} << lg) < job.total_work) ++lg;
    job.capacity_log2 = lg;
    job.graph_steps = 16;
    job.use_graph = false;
    job.timing = true;
    job.max_depth = max_depth;
    const int rc = stream_render(ctx, S->stream, *ctx->streamer, job, &S->rstats);
    if (rc != MPT_OK) return rc;
    if (keep && !(p->accumulate)) {
        S->keep_valid = true;
        S->keep_seed = seed;
        S->keep_begin = spp_begin;
        S->keep_end = spp_end;
        S->keep_depth = max_depth;
        S->keep_version = S->param_version;
    }
    return MPT_OK;
}

// Path-replay adjoint: the forward pass's kernels in their ADJ form over the
// same passes, reading the kept per-sample radiance and the film's weights.
int do_adjoint_passes(mpt_scene* S, const mpt_render_params* p, const float* d_image, float* d_tex, float* d_light) {
    mpt_ctx* ctx = S->ctx;
    const auto t0 = std::chrono::steady_clock::now();
    const HostScene& h = S->host;
    const int32_t max_depth = (p && p->max_depth > 0) ? p->max_depth : h.max_depth;
    const int32_t spp_begin = p ? std::max(0, p->spp_begin) : 0;
    const int32_t spp_end = (p && p->spp_end > 0) ? p->spp_end : h.spp;
    const uint64_t seed = p ? p->seed : 0;
    const int64_t npix = static_cast<int64_t>(h.width) * h.height;
    if (S->tex_elems >= (int64_t{1} << 30)) return fail(ctx, MPT_ERR_UNSUPPORTED, "textures above 2^30 elements");
    int rc;
    if (!(S->keep_valid && S->keep_seed == seed && S->keep_begin == spp_begin && S->keep_end == spp_end &&
          S->keep_depth == max_depth && S->keep_version == S->param_version)) {
        mpt_render_params fp{seed, spp_begin, spp_end, max_depth, 0, MPT_RENDER_KEEP_SAMPLES};
        if ((rc = do_render(S, &fp)) != MPT_OK) return rc;
    }
    const mpt_stats fwd = S->rstats;
    if ((rc = ensure_pass_buffers(S, npix, max_depth)) != MPT_OK) return rc;
    if (S->adj_npaths != npix) {
        S->adj_npaths = 0;   // until every buffer below has its new size
        const size_t sizes[6] = {8, 16, 16, 16, 16, 16};
        for (int q = 0; q < 2; ++q)
            for (int k = 0; k < 6; ++k) HIP_TRY(ctx, S->adjq[q][k].alloc(sizes[k] * static_cast<size_t>(npix)));
        HIP_TRY(ctx, S->adjA.alloc(sizeof(float4) * static_cast<size_t>(npix)));
        HIP_TRY(ctx, S->dimg.alloc(sizeof(float) * 3 * static_cast<size_t>(npix)));
        S->adj_npaths = npix;
    }
    HIP_TRY(ctx, S->gtex.alloc(sizeof(float) * static_cast<size_t>(std::max<int64_t>(1, S->tex_elems))));
    HIP_TRY(ctx, S->glight.alloc(sizeof(float) * 3 * static_cast<size_t>(std::max<size_t>(1, h.lights.size()))));
    HIP_TRY(ctx, hipMemcpyAsync(S->dimg.p, d_image, sizeof(float) * 3 * static_cast<size_t>(npix), hipMemcpyHostToDevice,
                                S->stream));
    {
        const int64_t nt = std::max<int64_t>(1, S->tex_elems);
        hipLaunchKernelGGL(k_zero_i32, dim3((nt + 255) / 256), dim3(256), 0, S->stream, S->gtex.as<int32_t>(),
                           static_cast<int>(nt));
        const int nl = static_cast<int>(3 * std::max<size_t>(1, h.lights.size()));
        hipLaunchKernelGGL(k_zero_i32, dim3((nl + 255) / 256), dim3(256), 0, S->stream, S->glight.as<int32_t>(), nl);
        HIP_TRY(ctx, hipGetLastError());
    }
    if ((rc = zero_stats(S)) != MPT_OK) return rc;
    const int tblocks = trace_blocks(S, 2 * npix);
    if ((rc = ensure_spill(S, tblocks)) != MPT_OK) return rc;
    const int sblocks = static_cast<int>(std::min<int64_t>((npix + 255) / 256, static_cast<int64_t>(ctx->num_cus) * 8));
    const int cam_threads = static_cast<int>(std::max<int64_t>(npix, 3 * (max_depth + 2) + 1));
    while (S->events.size() < 2) {
        hipEvent_t e;
        HIP_TRY(ctx, hipEventCreate(&e));
        S->events.push_back(e);
    }
    HIP_TRY(ctx, hipEventRecord(S->events[0], S->stream));
    for (int32_t k = spp_begin; k < spp_end; ++k) {
        PassArgs a = pass_args(S, seed, max_depth);
        a.sample = k;
        AdjArgs g{};
        for (int q = 0; q < 2; ++q) {
            g.q[q].par = S->adjq[q][0].as<int2>();
            g.q[q].sL = S->adjq[q][1].as<float4>();
            g.q[q].eL = S->adjq[q][2].as<float4>();
            g.q[q].sB = S->adjq[q][3].as<float4>();
            g.q[q].eB = S->adjq[q][4].as<float4>();
            g.q[q].sC = S->adjq[q][5].as<float4>();
        }
        g.samples = S->samples.as<float4>() + static_cast<int64_t>(k - spp_begin) * npix;
        g.A = S->adjA.as<float4>();
        g.dimg = S->dimg.as<float>();
        g.film = S->film.as<float4>();
        g.gtex = S->gtex.as<float>();
        g.glight = S->glight.as<float>();
        hipLaunchKernelGGL(k_camera<true>, dim3((cam_threads + 255) / 256), dim3(256), 0, S->stream, a, g);
        HIP_TRY(ctx, hipGetLastError());
        for (int32_t b = 0; b <= max_depth; ++b) {
            TraceArgs t = queue_trace_args(S->sc, a.q[b & 1], a.qcount + 2 * b, false);
            t.work = a.work + 3 * b;
            t.stats = S->stats.as<unsigned long long>();
            if ((rc = launch_trace(S, t, b > 0 ? TRACE_MIXED : TRACE_CLOSEST, tblocks)) != MPT_OK) return rc;
            hipLaunchKernelGGL(k_shade<SHADE_ADJ>, dim3(sblocks), dim3(256), 0, S->stream, a, g, BatchArgs{}, b);
            HIP_TRY(ctx, hipGetLastError());
        }
    }
    HIP_TRY(ctx, hipEventRecord(S->events[1], S->stream));
    if (d_tex && S->tex_elems > 0)
        HIP_TRY(ctx, hipMemcpyAsync(d_tex, S->gtex.p, sizeof(float) * static_cast<size_t>(S->tex_elems),
                                    hipMemcpyDeviceToHost, S->stream));
    if (d_light && !h.lights.empty())
        HIP_TRY(ctx, hipMemcpyAsync(d_light, S->glight.p, sizeof(float) * 3 * h.lights.size(), hipMemcpyDeviceToHost,
                                    S->stream));
    HIP_TRY(ctx, hipStreamSynchronize(S->stream));
    unsigned long long st[ST_COUNT];
    if ((rc = read_stats(S, st)) != MPT_OK) return rc;
    float adj_ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&adj_ms, S->events[0], S->events[1]));
    mpt_stats& r = S->rstats;
    r = fwd;
    r.ms_total = ms_since(t0);
    r.ms_adjoint = adj_ms;
    if (st[ST_ERRORS]) return fail(ctx, MPT_ERR_DEVICE, "traversal stack overrun");
    return MPT_OK;
}

// Path-replay adjoint on the streaming wavefront: the forward render's paths
// (same seed, passes, depth) regenerated into the queues, each carrying its
// adjoint weight, shaded by k_shade<SHADE_BATCH_ADJ>.
int do_adjoint(mpt_scene* S, const mpt_render_params* p, const float* d_image, float* d_tex, float* d_light) {
    if (p && (p->flags & MPT_RENDER_PASSWISE)) return do_adjoint_passes(S, p, d_image, d_tex, d_light);
    mpt_ctx* ctx = S->ctx;
    const auto t0 = std::chrono::steady_clock::now();
    const HostScene& h = S->host;
    const int32_t max_depth = (p && p->max_depth > 0) ? p->max_depth : h.max_depth;
    const int32_t spp_begin = p ? std::max(0, p->spp_begin) : 0;
    const int32_t spp_end = (p && p->spp_end > 0) ? p->spp_end : h.spp;
    const uint64_t seed = p ? p->seed : 0;
    const int64_t npix = static_cast<int64_t>(h.width) * h.height;
    const int32_t nsamples = std::max(0, spp_end - spp_begin);
    if (S->tex_elems >= (int64_t{1} << 30)) return fail(ctx, MPT_ERR_UNSUPPORTED, "textures above 2^30 elements");
    int rc;
    if (!(S->keep_valid && S->keep_seed == seed && S->keep_begin == spp_begin && S->keep_end == spp_end &&
          S->keep_depth == max_depth && S->keep_version == S->param_version)) {
        mpt_render_params fp{seed, spp_begin, spp_end, max_depth, 0, MPT_RENDER_KEEP_SAMPLES};
        if ((rc = do_render(S, &fp)) != MPT_OK) return rc;
    }
    const mpt_stats fwd = S->rstats;
    if (S->dimg.bytes < sizeof(float) * 3 * static_cast<size_t>(npix))
        HIP_TRY(ctx, S->dimg.alloc(sizeof(float) * 3 * static_cast<size_t>(npix)));
    HIP_TRY(ctx, S->gtex.alloc(sizeof(float) * static_cast<size_t>(std::max<int64_t>(1, S->tex_elems))));
    HIP_TRY(ctx, S->glight.alloc(sizeof(float) * 3 * static_cast<size_t>(std::max<size_t>(1, h.lights.size()))));
    HIP_TRY(ctx, hipMemcpyAsync(S->dimg.p, d_image, sizeof(float) * 3 * static_cast<size_t>(npix), hipMemcpyHostToDevice,
                                S->stream));
    {
        const int64_t nt = std::max<int64_t>(1, S->tex_elems);
        hipLaunchKernelGGL(k_zero_i32, dim3((nt + 255) / 256), dim3(256), 0, S->stream, S->gtex.as<int32_t>(),
                           static_cast<int>(nt));
        const int nl = static_cast<int>(3 * std::max<size_t>(1, h.lights.size()));
        hipLaunchKernelGGL(k_zero_i32, dim3((nl + 255) / 256), dim3(256), 0, S->stream, S->glight.as<int32_t>(), nl);
        HIP_TRY(ctx, hipGetLastError());
    }
    if (!ctx->streamer) ctx->streamer.reset(new Streamer());
    const BatchView v = scene_view(S, seed, spp_begin, nsamples, max_depth);
    StreamJob job;
    job.sc = S->sc;
    job.views = &v;
    job.nviews = 1;
    job.total_work = npix * nsamples;
    job.film = S->film.as<float4>();
    job.film_px = npix;
    job.clear_film = false;
    job.cfg = S->cfg;
    int lg = 16;
    while (lg < 22 && (int64_t{1} << lg) < job.total_work) ++lg;
    job.capacity_log2 = lg;
    job.graph_steps = 16;
    job.use_graph = false;
    job.timing = true;
    job.max_depth = max_depth;
    job.adjoint = true;
    job.adj.samples = S->samples.as<float4>();
    job.adj.dimg = S->dimg.as<float>();
    job.adj.film = S->film.as<float4>();
    job.adj.gtex = S->gtex.as<float>();
    job.adj.glight = S->glight.as<float>();
    mpt_stats ast{};
    if ((rc = stream_render(ctx, S->stream, *ctx->streamer, job, &ast)) != MPT_OK) return rc;
    if (d_tex && S->tex_elems > 0)
        HIP_TRY(ctx, hipMemcpyAsync(d_tex, S->gtex.p, sizeof(float) * static_cast<size_t>(S->tex_elems),
                                    hipMemcpyDeviceToHost, S->stream));
    if (d_light && !h.lights.empty())
        HIP_TRY(ctx, hipMemcpyAsync(d_light, S->glight.p, sizeof(float) * 3 * h.lights.size(), hipMemcpyDeviceToHost,
                                    S->stream));
    HIP_TRY(ctx, hipStreamSynchronize(S->stream));
    mpt_stats& r = S->rstats;
    r = fwd;
    r.ms_total = ms_since(t0);
    r.ms_adjoint = ast.ms_gpu;
    return MPT_OK;
}

int finish_scene(mpt_ctx* ctx, std::unique_ptr<mpt_scene>& S, const mpt_bvh_opts* opts, mpt_scene** out) {
    if (opts) S->opts = *opts;
    S->ctx = ctx;
    const int rc = build_and_upload(S.get());
    if (rc != MPT_OK) return rc;
    *out = S.release();
    return MPT_OK;
}

// ---- fixed scratch of mpt_trace_device_noalloc ------------------------------------
//
// One array per device, held until the process ends: the traversal's work
// counter (word 0), its statistics (bytes 256..), and from byte 4096 the
// per-lane spill slices of traversal stacks deeper than the LDS part. Host code
// uses it only through the device address hipGetSymbolAddress returns (kept
// per device in kScratchPtr and in each context); the array's name is not a
// device address on the host.
constexpr size_t kScratchBytes = MPT_TRACE_SCRATCH_BYTES;
constexpr size_t kScratchSpillOffset = 4096;
static_assert(256 + sizeof(unsigned long long) * ST_COUNT <= kScratchSpillOffset, "scratch header");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
void* g_scratch_ptr[kMaxDevices] = {};

// With `device` current: this device's address of the scratch array (looked up once).
int scratch_address(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return MPT_ERR_DEVICE;
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MPT_ERR_DEVICE;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return MPT_OK;
}

}  // namespace

// ---- C ABI --------------------------------------------------------------------------

extern "C" {

const char* mpt_version(void) { return "mpt 0.1 gfx950"; }

int mpt_create(int device, mpt_ctx** out) {
    if (!out) return MPT_ERR_INVALID;
    *out = nullptr;
    if (device == -1) {
        auto* ctx = new mpt_ctx();
        ctx->device = -1;
        ctx->num_cus = 0;
        *out = ctx;
        return MPT_OK;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MPT_ERR_DEVICE;
    if (device < 0 || device >= n) return MPT_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return MPT_ERR_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MPT_ERR_DEVICE;
    auto* ctx = new mpt_ctx();
    ctx->device = device;
    ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    // the no-allocation trace's scratch: only that entry point fails without it
    if (scratch_address(device, &ctx->trace_scratch) != MPT_OK) {
        (void)hipGetLastError();
        ctx->trace_scratch = nullptr;
    }
    *out = ctx;
    return MPT_OK;
}

void mpt_destroy(mpt_ctx* ctx) { delete ctx; }

const char* mpt_last_error(const mpt_ctx* ctx) { return ctx ? ctx->error.c_str() : "no context"; }

int mpt_scene_parse_crsd(mpt_ctx* ctx, const char* text, size_t len, const mpt_bvh_opts* opts, mpt_scene** out) {
    if (!ctx || !text || !out) return fail(ctx, MPT_ERR_INVALID, "null argument");
    *out = nullptr;
    auto S = std::make_unique<mpt_scene>();
    const CrsdResult r = parse_crsd(std::string(text, len), "", &S->host);
    if (r.code != MPT_OK) return fail(ctx, r.code, r.message);
    return finish_scene(ctx, S, opts, out);
}

int mpt_scene_load_crsd(mpt_ctx* ctx, const char* path, const mpt_bvh_opts* opts, mpt_scene** out) {
    if (!ctx || !path || !out) return fail(ctx, MPT_ERR_INVALID, "null argument");
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(ctx, MPT_ERR_IO, std::string("cannot open ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string base(path);
    const size_t slash = base.find_last_of('/');
    base = slash == std::string::npos ? std::string() : base.substr(0, slash);
    auto S = std::make_unique<mpt_scene>();
    const CrsdResult r = parse_crsd(ss.str(), base, &S->host);
    if (r.code != MPT_OK) return fail(ctx, r.code, r.message);
    return finish_scene(ctx, S, opts, out);
}

int mpt_scene_create(mpt_ctx* ctx, const mpt_scene_desc* d, const mpt_bvh_opts* opts, mpt_scene** out) {
    if (!ctx || !d || !out) return fail(ctx, MPT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (d->nprims < 0 || (d->nprims > 0 && (!d->prim_type || !d->prim_v || !d->prim_mat)))
        return fail(ctx, MPT_ERR_INVALID, "primitive arrays missing");
    if (d->nmats <= 0 || !d->mats) return fail(ctx, MPT_ERR_INVALID, "no materials");
    if (d->nlights <= 0 || !d->lights) return fail(ctx, MPT_ERR_INVALID, "no lights");
    if (d->width <= 0 || d->height <= 0 || d->tile_x <= 0 || d->tile_y <= 0 || d->spp <= 0 || d->max_depth < 0)
        return fail(ctx, MPT_ERR_INVALID, "bad film / renderer settings");
    if (d->nprims > MPT_MAX_PRIMITIVES)
        return fail(ctx, MPT_ERR_UNSUPPORTED, "more than MPT_MAX_PRIMITIVES primitives");
    auto S = std::make_unique<mpt_scene>();
    HostScene& h = S->host;
    const int64_t n = d->nprims;
    h.prim_type.assign(d->prim_type, d->prim_type + n);
    h.prim_v.assign(d->prim_v, d->prim_v + 9 * n);
    h.prim_mat.assign(d->prim_mat, d->prim_mat + n);
    if (d->prim_n) h.prim_n.assign(d->prim_n, d->prim_n + 9 * n);
    if (d->prim_uvw) h.prim_uvw.assign(d->prim_uvw, d->prim_uvw + 9 * n);
    for (int64_t i = 0; i < n; ++i) {
        if (h.prim_type[i] != PRIM_TRIANGLE && h.prim_type[i] != PRIM_RECTANGLE && h.prim_type[i] != PRIM_SPHERE)
            return fail(ctx, MPT_ERR_INVALID, "bad primitive type");
        if (h.prim_mat[i] < 0 || h.prim_mat[i] >= d->nmats) return fail(ctx, MPT_ERR_INVALID, "bad material index");
    }
    std::vector<std::pair<const float*, int32_t>> seen;   // the same array in several materials is one texture
    std::function<int32_t(const mpt_texture_desc&, int, int)> add_tex = [&](const mpt_texture_desc& t, int channels,
                                                                           int depth) -> int32_t {
        if (t.odd && t.even) {
            if (depth >= kMaxTextureNesting) return -2;
            const int32_t odd = add_tex(*t.odd, channels, depth + 1);
            const int32_t even = add_tex(*t.even, channels, depth + 1);
            if (odd == -2 || even == -2) return -2;
            Texture x;
            x.channels = channels;
            x.odd = odd;
            x.even = even;
            x.odd_size = 1.0f / t.odd_number;
            x.even_size = 1.0f / t.even_number;
            h.textures.push_back(std::move(x));
            return static_cast<int32_t>(h.textures.size() - 1);
        }
        if (!t.data) {
            if (depth == 0) return -1;   // a material's constant lives in the material
            Texture x;
            x.channels = channels;
            x.value = {t.value[0], channels == 3 ? t.value[1] : t.value[0], channels == 3 ? t.value[2] : t.value[0]};
            h.textures.push_back(std::move(x));
            return static_cast<int32_t>(h.textures.size() - 1);
        }
        for (const auto& sn : seen)
            if (sn.first == t.data && h.textures[sn.second].channels == channels &&
                h.textures[sn.second].width == t.width && h.textures[sn.second].height == t.height)
                return sn.second;
        seen.push_back({t.data, static_cast<int32_t>(h.textures.size())});
        Texture x;
        x.width = t.width;
        x.height = t.height;
        x.channels = channels;
        x.wrap_clamp = t.wrap_clamp;
        const size_t cnt = static_cast<size_t>(t.width) * t.height * t.channels;
        if (t.channels == channels) {
            x.data.assign(t.data, t.data + cnt);
        } else {
            x.data.resize(static_cast<size_t>(t.width) * t.height * channels);
            for (size_t px = 0; px < static_cast<size_t>(t.width) * t.height; ++px)
                for (int c = 0; c < channels; ++c) x.data[px * channels + c] = t.data[px * t.channels + std::min(c, t.channels - 1)];
        }
        h.textures.push_back(std::move(x));
        return static_cast<int32_t>(h.textures.size() - 1);
    };
    for (int32_t i = 0; i < d->nmats; ++i) {
        const mpt_material_desc& m = d->mats[i];
        Material x;
        x.type = m.type;
        x.albedo = {m.albedo.value[0], m.albedo.value[1], m.albedo.value[2]};
        x.albedo_tex = add_tex(m.albedo, 3, 0);
        x.f0 = {m.f0[0], m.f0[1], m.f0[2]};
        x.roughness = m.roughness.value[0];
        x.rough_tex = add_tex(m.roughness, 1, 0);
        if (x.albedo_tex == -2 || x.rough_tex == -2)
            return fail(ctx, MPT_ERR_UNSUPPORTED, "checkerboards nested deeper than 8");
        x.exponent = m.exponent;
        x.ratio = m.ratio;
        x.fresnel = m.fresnel;
        x.ior_outer = m.ior_outer;
        x.ior_inner = m.ior_inner;
        if (x.type < MAT_LAMBERT || x.type > MAT_ABSORB) return fail(ctx, MPT_ERR_INVALID, "material type");
        h.materials.push_back(x);
    }
    for (int32_t i = 0; i < d->nlights; ++i) {
        const mpt_light_desc& l = d->lights[i];
        if (l.type == MPT_LIGHT_ENVIRONMENT) {
            for (const Light& o : h.lights)
                if (o.type == LIGHT_ENVIRONMENT) return fail(ctx, MPT_ERR_UNSUPPORTED, "more than one environment light");
            if (!l.env.data || l.env.width <= 0 || l.env.height <= 0 || l.env.channels != 3)
                return fail(ctx, MPT_ERR_INVALID, "environment light needs an H x W x 3 image");
            Texture t;
            t.width = l.env.width;
            t.height = l.env.height;
            t.channels = 3;
            t.data.assign(l.env.data, l.env.data + static_cast<size_t>(t.width) * t.height * 3);
            h.textures.push_back(std::move(t));
            Light x;
            x.type = LIGHT_ENVIRONMENT;
            x.env_tex = static_cast<int32_t>(h.textures.size() - 1);
            h.lights.push_back(x);
            continue;
        }
        if (l.type != MPT_LIGHT_AREA) return fail(ctx, MPT_ERR_INVALID, "light type");
        if (l.prim < 0 || l.prim >= n) return fail(ctx, MPT_ERR_INVALID, "light primitive out of range");
        if (h.prim_type[l.prim] == PRIM_SPHERE)
            return fail(ctx, MPT_ERR_UNSUPPORTED, "area light on a sphere (the reference cannot sample it)");
        Light x;
        x.prim = l.prim;
        x.color = {l.color[0], l.color[1], l.color[2]};
        x.watt = l.watt;
        x.back_face = l.back_face;
        h.lights.push_back(x);
        if (!(prim_area(h, l.prim) > 0.f)) return fail(ctx, MPT_ERR_INVALID, "zero-area emitter");
    }
    h.cluster = d->cluster;
    if (d->estimator != MPT_ESTIMATOR_PT && d->estimator != MPT_ESTIMATOR_VPT)
        return fail(ctx, MPT_ERR_INVALID, "estimator");
    h.estimator = d->estimator;
    h.cam_pos = {d->cam_pos[0], d->cam_pos[1], d->cam_pos[2]};
    h.cam_dir = {d->cam_dir[0], d->cam_dir[1], d->cam_dir[2]};
    h.cam_up = {d->cam_up[0], d->cam_up[1], d->cam_up[2]};
    h.fov = d->fov_deg;
    h.sensor_width_mm = d->sensor_width_mm > 0.f ? d->sensor_width_mm : 36.f;
    h.width = d->width;
    h.height = d->height;
    h.tile_x = d->tile_x;
    h.tile_y = d->tile_y;
    h.filter = d->filter;
    h.filter_w = d->filter_w;
    h.filter_h = d->filter_h;
    h.sigma_x = d->sigma_x;
    h.sigma_y = d->sigma_y;
    h.amplitude = d->amplitude;
    h.mitchell_b = d->mitchell_b;
    h.mitchell_c = d->mitchell_c;
    h.sampler = d->sampler;
    h.spp = d->spp;
    h.max_depth = d->max_depth;
    h.split_sah = d->split_sah;
    if (d->output_filename) h.output = d->output_filename;
    return finish_scene(ctx, S, opts, out);
}

void mpt_scene_destroy(mpt_scene* s) {
    if (!s) return;
    if (s->ctx && s->ctx->device >= 0) (void)hipSetDevice(s->ctx->device);
    delete s;
}

int mpt_scene_camera(const mpt_scene* s, float* f2c, float* c2w) {
    if (!s || !f2c || !c2w) return MPT_ERR_INVALID;
    camera_matrices(s->host, f2c, c2w);
    return MPT_OK;
}

int mpt_scene_light_radiance(const mpt_scene* s, float* out) {
    if (!s || !out) return MPT_ERR_INVALID;
    const std::vector<F3> r = light_radiance(s->host);
    for (size_t i = 0; i < r.size(); ++i) {
        out[3 * i] = r[i].x;
        out[3 * i + 1] = r[i].y;
        out[3 * i + 2] = r[i].z;
    }
    return MPT_OK;
}

int mpt_scene_light_pick_pdf(const mpt_scene* s, float* out) {
    if (!s || !out) return MPT_ERR_INVALID;
    const size_t n = s->host.lights.size();
    if (s->host.cluster == 1) {
        // the radiance set by mpt_scene_set_light_radiance, else the parsed one
        const std::vector<float> cdf = power_cdf(s->host, s->h_rad.empty() ? light_radiance(s->host) : s->h_rad);
        for (size_t i = 0; i < n; ++i) out[i] = cdf[i + 1] - cdf[i];
    } else {
        for (size_t i = 0; i < n; ++i) out[i] = 1.0f / static_cast<float>(n);
    }
    return MPT_OK;
}

int mpt_scene_stats_get(const mpt_scene* s, mpt_scene_stats* out) {
    if (!s || !out) return MPT_ERR_INVALID;
    *out = s->sstats;
    return MPT_OK;
}

int mpt_scene_binary_bvh(mpt_scene* s, int64_t* count, float* bounds, int32_t* is_leaf, int64_t* a, int64_t* b,
                         int32_t* ordered) {
    if (!s || !count) return MPT_ERR_INVALID;
    if (!s->ref_built) {
        build_binary(s->host, true, &s->ref_bin);
        s->ref_built = true;
    }
    const BinaryBvh& bin = s->ref_bin;
    *count = static_cast<int64_t>(bin.nodes.size());
    if (!bounds) return MPT_OK;
    for (size_t i = 0; i < bin.nodes.size(); ++i) {
        const BinaryNode& n = bin.nodes[i];
        std::memcpy(bounds + 6 * i, n.bmin, 12);
        std::memcpy(bounds + 6 * i + 3, n.bmax, 12);
        if (is_leaf) is_leaf[i] = n.leaf;
        if (a) a[i] = n.a;
        if (b) b[i] = n.b;
    }
    if (ordered) std::memcpy(ordered, bin.ordered.data(), bin.ordered.size() * sizeof(int32_t));
    return MPT_OK;
}

#define REQUIRE_DEVICE(s)                                                                          \
    do {                                                                                           \
        if ((s)->ctx->device < 0) return fail((s)->ctx, MPT_ERR_DEVICE, "host-only context has no device"); \
    } while (0)

int mpt_trace_device(mpt_scene* s, const mpt_ray* rays_dev, size_t n, int any_hit, mpt_hit* hits_dev, void* stream) {
    if (!s) return MPT_ERR_INVALID;
    if (n > MPT_TRACE_MAX_RAYS) return fail(s->ctx, MPT_ERR_INVALID, "at most MPT_TRACE_MAX_RAYS rays per call");
    if (n > 0 && (!rays_dev || !hits_dev)) return fail(s->ctx, MPT_ERR_INVALID, "null ray / hit buffer");
    REQUIRE_DEVICE(s);
    HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    return do_trace(s, rays_dev, n, any_hit, hits_dev, static_cast<hipStream_t>(stream));
}

int mpt_trace(mpt_scene* s, const mpt_ray* rays, size_t n, int any_hit, mpt_hit* hits) {
    if (!s) return MPT_ERR_INVALID;
    if (n > MPT_TRACE_MAX_RAYS) return fail(s->ctx, MPT_ERR_INVALID, "at most MPT_TRACE_MAX_RAYS rays per call");
    if (n > 0 && (!rays || !hits)) return fail(s->ctx, MPT_ERR_INVALID, "null ray / hit buffer");
    REQUIRE_DEVICE(s);
    if (n == 0) return MPT_OK;
    HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    DevBuf r, h;
    HIP_TRY(s->ctx, r.alloc(n * sizeof(mpt_ray)));
    HIP_TRY(s->ctx, h.alloc(n * sizeof(mpt_hit)));
    HIP_TRY(s->ctx, hipMemcpy(r.p, rays, n * sizeof(mpt_ray), hipMemcpyHostToDevice));
    const int rc = do_trace(s, r.as<mpt_ray>(), n, any_hit, h.as<mpt_hit>(), nullptr);
    if (rc != MPT_OK) return rc;
    HIP_TRY(s->ctx, hipMemcpy(hits, h.p, n * sizeof(mpt_hit), hipMemcpyDeviceToHost));
    return MPT_OK;
}

int mpt_intersect_detail(mpt_scene* s, const mpt_ray* rays, size_t n, float* out) {
    if (!s) return MPT_ERR_INVALID;
    if (n > MPT_TRACE_MAX_RAYS) return fail(s->ctx, MPT_ERR_INVALID, "at most MPT_TRACE_MAX_RAYS rays per call");
    if (n > 0 && (!rays || !out)) return fail(s->ctx, MPT_ERR_INVALID, "null ray / output buffer");
    REQUIRE_DEVICE(s);
    if (n == 0) return MPT_OK;
    HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    DevBuf r, h, o;
    HIP_TRY(s->ctx, r.alloc(n * sizeof(mpt_ray)));
    HIP_TRY(s->ctx, h.alloc(n * sizeof(mpt_hit)));
    HIP_TRY(s->ctx, o.alloc(n * 14 * sizeof(float)));
    HIP_TRY(s->ctx, hipMemcpy(r.p, rays, n * sizeof(mpt_ray), hipMemcpyHostToDevice));
    // closest hits by leaf position, then the detail of each
    const int rc = do_trace(s, r.as<mpt_ray>(), n, 0, h.as<mpt_hit>(), nullptr, 0);
    if (rc != MPT_OK) return rc;
    const int64_t nn = static_cast<int64_t>(n);
    hipLaunchKernelGGL(k_detail, dim3(static_cast<unsigned>((nn + 255) / 256)), dim3(256), 0, s->stream, s->sc,
                       r.as<float4>(), h.as<int2>(), nn, o.as<float>());
    HIP_TRY(s->ctx, hipGetLastError());
    HIP_TRY(s->ctx, hipStreamSynchronize(s->stream));
    HIP_TRY(s->ctx, hipMemcpy(out, o.p, n * 14 * sizeof(float), hipMemcpyDeviceToHost));
    return MPT_OK;
}

int mpt_trace_device_noalloc(mpt_scene* s, const mpt_ray* rays_dev, size_t n, int any_hit, mpt_hit* hits_dev,
                             void* stream) {
    if (!s) return MPT_ERR_INVALID;
    if (n > MPT_TRACE_MAX_RAYS) return fail(s->ctx, MPT_ERR_INVALID, "at most MPT_TRACE_MAX_RAYS rays per call");
    if (n > 0 && (!rays_dev || !hits_dev)) return fail(s->ctx, MPT_ERR_INVALID, "null ray / hit buffer");
    REQUIRE_DEVICE(s);
    if (n == 0) return MPT_OK;
    mpt_ctx* ctx = s->ctx;
    if (!ctx->trace_scratch) return fail(ctx, MPT_ERR_DEVICE, "no scratch array on this device");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    unsigned char* base = static_cast<unsigned char*>(ctx->trace_scratch);
    int blocks = trace_blocks(s, static_cast<int64_t>(n));
    const int64_t extra = s->sc.max_stack - s->cfg.stack;
    if (extra > 0) {
        const int64_t entries = static_cast<int64_t>((kScratchBytes - kScratchSpillOffset) / sizeof(int32_t));
        const int64_t fit = entries / (extra * s->cfg.block);   // whole blocks whose lanes' spill slices fit
        if (fit < 1)
            return fail(ctx, MPT_ERR_INVALID, "the tree's traversal stack (" + std::to_string(s->sc.max_stack) +
                                                  " entries) needs more spill than MPT_TRACE_SCRATCH_BYTES holds");
        blocks = static_cast<int>(std::min<int64_t>(blocks, fit));
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    // counter and statistics to zero with a fill kernel (a caller may capture this)
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned long long*>(base),
                       static_cast<int>(kScratchSpillOffset / sizeof(unsigned long long)));
    HIP_TRY(ctx, hipGetLastError());
    TraceArgs a{};
    a.sc = s->sc;
    a.org = reinterpret_cast<const float4*>(rays_dev);
    a.dir = reinterpret_cast<const float4*>(rays_dev) + 1;
    a.stride = 2;
    a.n_host = static_cast<int64_t>(n);
    a.hit = reinterpret_cast<int2*>(hits_dev);
    a.want_prim = 1;
    a.work = reinterpret_cast<int32_t*>(base);
    a.stats = reinterpret_cast<unsigned long long*>(base + 256);
    a.stat_slot = any_hit ? ST_SHADOW_RAYS : ST_CLOSEST_RAYS;
    a.spill = extra > 0 ? reinterpret_cast<int32_t*>(base + kScratchSpillOffset) : nullptr;
    launch_trace_cfg(s->cfg, a, any_hit != 0 ? TRACE_ANY : TRACE_CLOSEST, blocks, st);
    HIP_TRY(ctx, hipGetLastError());
    return MPT_OK;
}

int mpt_render(mpt_scene* s, const mpt_render_params* p, float* film_host) {
    if (!s) return MPT_ERR_INVALID;
    REQUIRE_DEVICE(s);
    HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    const int rc = do_render(s, p);
    if (rc != MPT_OK) return rc;
    if (film_host) {
        HIP_TRY(s->ctx, hipMemcpy(film_host, s->film.p, s->film.bytes, hipMemcpyDeviceToHost));
    }
    return MPT_OK;
}

int mpt_film_device(mpt_scene* s, void** film_dev) {
    if (!s || !film_dev) return MPT_ERR_INVALID;
    *film_dev = s->film.p;
    return MPT_OK;
}

int mpt_film_srgb8(mpt_scene* s, uint8_t* out) {
    if (!s || !out) return MPT_ERR_INVALID;
    REQUIRE_DEVICE(s);
    HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    const int64_t npix = static_cast<int64_t>(s->host.width) * s->host.height;
    HIP_TRY(s->ctx, s->srgb.alloc(static_cast<size_t>(3 * npix)));
    hipLaunchKernelGGL(k_srgb8, dim3((npix + 255) / 256), dim3(256), 0, s->stream, s->film.as<float4>(), npix,
                       s->srgb.as<uint8_t>());
    HIP_TRY(s->ctx, hipGetLastError());
    HIP_TRY(s->ctx, hipStreamSynchronize(s->stream));
    HIP_TRY(s->ctx, hipMemcpy(out, s->srgb.p, static_cast<size_t>(3 * npix), hipMemcpyDeviceToHost));
    return MPT_OK;
}

int mpt_film_srgb8_device(mpt_ctx* ctx, const float* film_dev, int64_t npix, uint8_t* out_dev, void* stream) {
    if (!ctx) return MPT_ERR_INVALID;
    if (ctx->device < 0) return fail(ctx, MPT_ERR_DEVICE, "host-only context has no device");
    if (npix < 0) return fail(ctx, MPT_ERR_INVALID, "negative pixel count");
    if (npix == 0) return MPT_OK;
    if (!film_dev || !out_dev) return fail(ctx, MPT_ERR_INVALID, "null film / byte buffer");
    if (npix > (static_cast<int64_t>(1) << 40)) return fail(ctx, MPT_ERR_INVALID, "at most 2^40 pixels per call");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int64_t blocks = (npix + 255) / 256;
    const unsigned grid = static_cast<unsigned>(std::min<int64_t>(blocks, 1 << 20));
    hipLaunchKernelGGL(k_srgb8, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream),
                       reinterpret_cast<const float4*>(film_dev), npix, out_dev);
    HIP_TRY(ctx, hipGetLastError());
    return MPT_OK;
}

int mpt_film_save(mpt_scene* s, const char* path) {
    if (!s) return MPT_ERR_INVALID;
    REQUIRE_DEVICE(s);
    // NULL: the film block's output-filename (filmInstantiator.cpp:24), as Film::save
    const std::string p = path ? std::string(path) : s->host.output;
    const int w = s->host.width, h = s->host.height;
    const int kind = picture_kind(p);
    if (kind == PICTURE_NONE)
        return fail(s->ctx, MPT_ERR_UNSUPPORTED, "output format of " + p + " (png, jpg, ppm; pfm for linear films)");
    HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    std::string err;
    if (kind == PICTURE_PFM) {
        std::vector<float> film(static_cast<size_t>(4) * w * h);
        HIP_TRY(s->ctx, hipMemcpy(film.data(), s->film.p, film.size() * sizeof(float), hipMemcpyDeviceToHost));
        return write_pfm(p, film.data(), w, h, &err) ? MPT_OK : fail(s->ctx, MPT_ERR_IO, err);
    }
    std::vector<uint8_t> rgb(static_cast<size_t>(3) * w * h);
    const int rc = mpt_film_srgb8(s, rgb.data());
    if (rc != MPT_OK) return rc;
    return write_picture8(p, rgb.data(), w, h, &err) ? MPT_OK : fail(s->ctx, MPT_ERR_IO, err);
}

int mpt_scene_output_filename(const mpt_scene* s, char* buf, size_t n, size_t* needed) {
    if (!s) return MPT_ERR_INVALID;
    const std::string& o = s->host.output;
    if (needed) *needed = o.size() + 1;
    if (!buf) return MPT_OK;
    if (n < o.size() + 1) return fail(s->ctx, MPT_ERR_INVALID, "buffer shorter than the output file name");
    std::memcpy(buf, o.c_str(), o.size() + 1);
    return MPT_OK;
}

int mpt_scene_texture_info(const mpt_scene* s, int32_t index, int64_t* width, int64_t* height, int64_t* channels,
                           int64_t* offset) {
    if (!s) return MPT_ERR_INVALID;
    const auto& T = s->host.textures;
    if (index < 0 || index >= static_cast<int32_t>(T.size())) return fail(s->ctx, MPT_ERR_INVALID, "texture index");
    int64_t off = 0;
    for (int32_t i = 0; i <= index; ++i) {
        const int64_t w = T[i].data.empty() ? 1 : T[i].width, hh = T[i].data.empty() ? 1 : T[i].height;
        if (i == index) {
            if (width) *width = w;
            if (height) *height = hh;
            if (channels) *channels = T[i].channels;
            if (offset) *offset = off;
        }
        off += w * hh * T[i].channels;
    }
    return MPT_OK;
}

int mpt_scene_get_texture(const mpt_scene* s, int32_t index, float* out) {
    if (!s || !out) return MPT_ERR_INVALID;
    int64_t w = 0, hh = 0, c = 0, off = 0;
    const int rc = mpt_scene_texture_info(s, index, &w, &hh, &c, &off);
    if (rc != MPT_OK) return rc;
    const Texture& t = s->host.textures[index];
    if (t.data.empty()) {
        out[0] = t.value.x;
        if (c == 3) {
            out[1] = t.value.y;
            out[2] = t.value.z;
        }
    } else {
        std::copy(t.data.begin(), t.data.end(), out);
    }
    return MPT_OK;
}

int mpt_scene_set_texture(mpt_scene* s, int32_t index, const float* data) {
    if (!s || !data) return MPT_ERR_INVALID;
    int64_t w = 0, hh = 0, c = 0, off = 0;
    const int rc = mpt_scene_texture_info(s, index, &w, &hh, &c, &off);
    if (rc != MPT_OK) return rc;
    const size_t n = static_cast<size_t>(w * hh * c);
    for (const Light& l : s->host.lights)
        if (l.type == LIGHT_ENVIRONMENT && l.env_tex == index)
            return fail(s->ctx, MPT_ERR_UNSUPPORTED, "an environment light's texture is not a parameter");
    Texture& t = s->host.textures[index];
    if (t.checker()) return fail(s->ctx, MPT_ERR_UNSUPPORTED, "a checkerboard has no texels of its own");
    if (t.data.empty()) {
        t.value = {data[0], c == 3 ? data[1] : data[0], c == 3 ? data[2] : data[0]};
    } else {
        t.data.assign(data, data + n);
    }
    ++s->param_version;
    if (s->ctx->device < 0) return MPT_OK;
    HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    HIP_TRY(s->ctx, hipMemcpyAsync(s->tex_data.as<float>() + off, data, n * sizeof(float), hipMemcpyHostToDevice,
                                   s->stream));
    HIP_TRY(s->ctx, hipStreamSynchronize(s->stream));
    return MPT_OK;
}

int mpt_scene_set_light_radiance(mpt_scene* s, int32_t light, const float* rgb) {
    if (!s || !rgb) return MPT_ERR_INVALID;
    if (light < 0 || light >= static_cast<int32_t>(s->host.lights.size())) return fail(s->ctx, MPT_ERR_INVALID, "light index");
    if (s->host.lights[light].type == LIGHT_ENVIRONMENT)
        return fail(s->ctx, MPT_ERR_UNSUPPORTED, "an environment light's radiance is its texture");
    ++s->param_version;
    if (s->ctx->device < 0) return MPT_OK;
    s->h_rad[light] = F3{rgb[0], rgb[1], rgb[2]};
    DevLight& d = s->h_lights[light];
    d.radiance[0] = rgb[0]; d.radiance[1] = rgb[1]; d.radiance[2] = rgb[2];
    HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    HIP_TRY(s->ctx, hipMemcpy(s->lights.p, s->h_lights.data(), sizeof(DevLight) * s->h_lights.size(),
                              hipMemcpyHostToDevice));
    // the power cluster's CDF follows the lights' flux
    const std::vector<float> cdf = power_cdf(s->host, s->h_rad);
    s->h_cdf = cdf;
    HIP_TRY(s->ctx, hipMemcpy(s->power_cdf.p, cdf.data(), sizeof(float) * cdf.size(), hipMemcpyHostToDevice));
    return MPT_OK;
}

int mpt_render_adjoint(mpt_scene* s, const mpt_render_params* p, const float* d_image, float* d_tex, float* d_light) {
    if (!s || !d_image) return MPT_ERR_INVALID;
    REQUIRE_DEVICE(s);
    HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    // the environment's radiance texture is not a differentiated parameter; the
    // replayed derivatives cover Lambertian and GGX vertices of the pt estimator
    if (s->sc.env_light >= 0) return fail(s->ctx, MPT_ERR_UNSUPPORTED, "adjoint of a scene with an environment light");
    if (s->host.estimator != EST_PT) return fail(s->ctx, MPT_ERR_UNSUPPORTED, "adjoint of the vpt estimator");
    for (const Material& m : s->host.materials)
        if (m.type != MAT_LAMBERT && m.type != MAT_GGX)
            return fail(s->ctx, MPT_ERR_UNSUPPORTED, "adjoint of materials other than matte-lambertian and GGX");
    return do_adjoint(s, p, d_image, d_tex, d_light);
}

int mpt_release_stream_buffers(mpt_ctx* ctx) {
    if (!ctx) return MPT_ERR_INVALID;
    if (ctx->device >= 0) HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (ctx->streamer) HIP_TRY(ctx, hipDeviceSynchronize());
    ctx->streamer.reset();
    return MPT_OK;
}

int mpt_stats_get(const mpt_scene* s, mpt_stats* out) {
    if (!s || !out) return MPT_ERR_INVALID;
    *out = s->rstats;
    return MPT_OK;
}

}  // extern "C"
