// Test infrastructure: a C-callable harness around the compiled reference
// (xh5a5n6k6/cadise, built from its own sources by oracle/Makefile into
// oracle/_ref/). It is the checker's checker: golden vectors under
// tests/golden/ come from here, and bench.py times it as the CPU baseline.
// Nothing in the product loads it.
//
// Scene assembly follows RenderDatabase::setUpData / prepareRender
// (source/core/renderDatabase.cpp:39-131) with the reference's own
// Tokenizer and instantiator::make* factories. Rendering follows
// SamplingRenderer::render + EstimatorTileWork::work
// (samplingRenderer.cpp:33-68, estimatorTileWork.cpp:32-61) but re-keys the
// random stream per (pixel, sample) so that results are reproducible.
//
// Harness-only CRSD extension (used for generated multi-million-triangle
// scenes, whose vector3r-array text would take minutes to tokenize):
//   primitive { [type string "triangle-mesh"] [positions-file string "x.f32"] ... }
// loads raw little-endian float32 triples (9 per triangle) and builds the
// triangles exactly as createTriangleMesh does (primitiveInstantiator.cpp:64-90).
#include "core/camera/camera.h"
#include "core/estimator/energyEstimator.h"
#include "core/film/film.h"
#include "core/film/filmTile.h"
#include "core/film/filter/filter.h"
#include "core/instantiator/instantiator.h"
#include "core/integral-tool/sample/bsdfSample.h"
#include "core/integral-tool/sample/directLightSample.h"
#include "core/intersector/accelerator/accelerator.h"
#include "core/intersector/accelerator/bvh/bvhBinaryNode.h"
#include "core/intersector/accelerator/bvh/bvhBuilder.h"
#include "core/intersector/accelerator/bvh/bvhLinearNode.h"
#include "core/intersector/primitive/primitive.h"
#include "core/intersector/primitive/triangleMesh.h"
#include "core/intersector/primitiveInfo.h"
#include "core/light/light.h"
#include "core/light/category/areaLight.h"
#include "math/random.h"
#include "core/ray.h"
#include "core/sampler/sampleRecord2D.h"
#include "core/sampler/sampler.h"
#include "core/scene.h"
#include "core/spectrum/spectrum.h"
#include "core/surface/bsdf/bsdf.h"
#include "core/surface/transportInfo.h"
#include "core/surfaceDetail.h"
#include "core/surfaceIntersection.h"
#include "core/imaging/tImage.h"
#include "file-io/path.h"
#include "file-io/pictureLoader.h"
#include "file-io/scene-description/sdData.h"
#include "file-io/scene-description/tokenizer.h"
#include "math/tAabb.h"
#include "math/type/imageType.h"
#include "math/type/mapType.h"

#include "ref_random.h"

extern "C" {
#include "../seqrng.h"
}

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace cadise;

namespace {

// Per-thread ray-query counters: [0] closest-hit, [1] any-hit.
thread_local int64_t t_rayCount[2] = {0, 0};

// Forwards to the reference BVH and counts queries; used so that the CPU
// baseline reports Mrays/s the same way the GPU tracer does.
class CountingAccelerator : public Accelerator {
public:
    explicit CountingAccelerator(std::shared_ptr<Accelerator> inner) : _inner(std::move(inner)) {}
    void evaluateBound(AABB3R* const out_bound) const override { _inner->evaluateBound(out_bound); }
    bool isIntersecting(Ray& ray, PrimitiveInfo& info) const override {
        t_rayCount[0] += 1;
        return _inner->isIntersecting(ray, info);
    }
    bool isOccluded(const Ray& ray) const override {
        t_rayCount[1] += 1;
        return _inner->isOccluded(ray);
    }
    std::shared_ptr<Accelerator> _inner;
};

struct RefScene {
    std::vector<std::shared_ptr<Intersector>> intersectors;
    std::vector<std::shared_ptr<Light>>       lights;
    StringKeyMap<Bsdf>                        bsdfs;
    StringKeyMap<Primitive>                   primitives;
    StringKeyMap<TTexture<real>>              realTextures;
    StringKeyMap<TTexture<Spectrum>>          spectrumTextures;
    std::shared_ptr<SdData> filmData, cameraData, rendererData, acceleratorData, clusterData;
    // SdData hands each value out once (TSdDataUnit::value() moves it,
    // tSdDataUnit.ipp), so the harness keeps the block text and tokenizes a
    // fresh copy for the keys it reads itself.
    std::string filmText, rendererText;

    std::shared_ptr<Film>            film;
    std::shared_ptr<Filter>          filter;
    std::shared_ptr<Camera>          camera;
    std::shared_ptr<Accelerator>     accelerator;
    std::shared_ptr<CountingAccelerator> counting;
    std::shared_ptr<LightCluster>    cluster;
    std::shared_ptr<Scene>           scene;
    std::shared_ptr<EnergyEstimator> estimator;
    std::shared_ptr<Sampler>         sampler;

    // the environment light's background sphere and the index RenderDatabase
    // records for it (renderDatabase.cpp:181-189)
    std::shared_ptr<Primitive> background;
    size_t                     envIndex = static_cast<size_t>(-1);

    std::unordered_map<const Primitive*, int32_t> primIndex;
    std::unordered_map<const Light*, int32_t>     lightIndex;
    int32_t width = 0, height = 0, tileX = 16, tileY = 16, spp = 0, maxDepth = 0;
    double  buildSeconds = 0.0;
    std::vector<std::string> errors;
};

std::string strip_comments(const char* text, size_t len) {
    // SdParser::_parseSdToString (sdParser.cpp:37-61): '#' to end of line is dropped
    std::string out;
    out.reserve(len);
    for (size_t i = 0; i < len; ++i) {
        if (text[i] == '#') {
            while (i < len && text[i] != '\n') ++i;
            continue;
        }
        out.push_back(text[i]);
    }
    return out;
}

bool load_f32(const std::string& path, std::vector<float>* out) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    const std::streamsize bytes = f.tellg();
    f.seekg(0);
    out->resize(static_cast<size_t>(bytes) / sizeof(float));
    f.read(reinterpret_cast<char*>(out->data()), bytes);
    return static_cast<bool>(f);
}

void make_mesh_from_files(RefScene* s, const std::shared_ptr<SdData>& data) {
    std::vector<float> pos, nrm, uvw;
    const std::string pfile(data->findString("positions-file"));
    if (!load_f32(pfile, &pos)) {
        s->errors.push_back("cannot read " + pfile);
        return;
    }
    const std::string nfile(data->findString("normals-file"));
    const std::string ufile(data->findString("uvws-file"));
    if (!nfile.empty()) load_f32(nfile, &nrm);
    if (!ufile.empty()) load_f32(ufile, &uvw);
    auto to_vec = [](const std::vector<float>& a) {
        std::vector<Vector3R> v(a.size() / 3);
        for (size_t i = 0; i < v.size(); ++i) v[i] = Vector3R(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
        return v;
    };
    const auto bsdfName = data->findString("bsdf");
    const auto bsdf     = s->bsdfs.find(bsdfName);
    if (bsdf == s->bsdfs.end()) {
        s->errors.push_back("unknown bsdf in mesh");
        return;
    }
    std::vector<std::shared_ptr<Primitive>> tris;
    TriangleMesh mesh(bsdf->second, to_vec(pos), to_vec(nrm), to_vec(uvw));
    mesh.transformToTriangles(&tris);
    const std::string name(data->findString("name"));
    for (auto& t : tris) {
        s->intersectors.push_back(t);
        s->primitives.insert(std::pair<std::string, std::shared_ptr<Primitive>>(name, t));
    }
}

void set_up(RefScene* s, const std::shared_ptr<SdData>& data, std::string_view text) {
    switch (data->classType()) {
        case ESdClassType::FILM: s->filmData = data; s->filmText = std::string(text); break;
        case ESdClassType::CAMERA: s->cameraData = data; break;
        case ESdClassType::RENDERER: s->rendererData = data; s->rendererText = std::string(text); break;
        case ESdClassType::ACCELERATOR: s->acceleratorData = data; break;
        case ESdClassType::LIGHT_CLUSTER: s->clusterData = data; break;
        case ESdClassType::TEXTURE_REAL: {
            auto tex = instantiator::makeRealTexture(data, s->realTextures, s->spectrumTextures);
            s->realTextures.insert(std::pair<std::string, std::shared_ptr<TTexture<real>>>(data->findString("name"), tex));
            break;
        }
        case ESdClassType::TEXTURE_SPECTRUM: {
            auto tex = instantiator::makeSpectrumTexture(data, s->realTextures, s->spectrumTextures);
            s->spectrumTextures.insert(
                std::pair<std::string, std::shared_ptr<TTexture<Spectrum>>>(data->findString("name"), tex));
            break;
        }
        case ESdClassType::MATERIAL: {
            auto bsdf = instantiator::makeBsdf(data, s->realTextures, s->spectrumTextures, s->bsdfs);
            s->bsdfs.insert(std::pair<std::string, std::shared_ptr<Bsdf>>(data->findString("name"), bsdf));
            break;
        }
        case ESdClassType::LIGHT: {
            // RenderDatabase::_setUpLight (renderDatabase.cpp:181-189)
            auto light = instantiator::makeLight(data, s->primitives, s->background);
            if (s->background) s->envIndex = s->lights.size();
            s->lights.push_back(light);
            break;
        }
        case ESdClassType::PRIMITIVE:
            if (std::string_view(text).find("positions-file") != std::string_view::npos) {
                make_mesh_from_files(s, data);
            } else {
                instantiator::makePrimitive(data, s->bsdfs, &s->intersectors, &s->primitives);
            }
            break;
        default: break;
    }
}

Vector3R v3(const float* p) { return Vector3R(p[0], p[1], p[2]); }
void put3(float* p, const Vector3R& v) { p[0] = v.x(); p[1] = v.y(); p[2] = v.z(); }

} // namespace

extern "C" {

void* ref_scene_create(const char* text, size_t len) {
    auto* s = new RefScene();
    const std::string body = strip_comments(text, len);
    const std::string_view view(body);
    Tokenizer tokenizer;
    size_t start = 0, end = 0;
    while ((end = view.find_first_of('}', start)) != std::string_view::npos) {
        const std::string_view block = view.substr(start, end - start + 1);
        set_up(s, tokenizer.tokenize(block), block);
        start = end + 1;
    }
    if (!s->filmData || !s->cameraData || !s->rendererData || !s->acceleratorData || !s->clusterData) {
        s->errors.push_back("missing film/camera/renderer/accelerator/light-cluster block");
        return s;
    }
    for (size_t i = 0; i < s->intersectors.size(); ++i) {
        s->primIndex[dynamic_cast<const Primitive*>(s->intersectors[i].get())] = static_cast<int32_t>(i);
    }
    for (size_t i = 0; i < s->lights.size(); ++i) {
        s->lightIndex[s->lights[i].get()] = static_cast<int32_t>(i);
    }
    // RenderDatabase::prepareRender (renderDatabase.cpp:78-131)
    s->film   = instantiator::makeFilm(s->filmData);
    const auto filmCopy     = tokenizer.tokenize(s->filmText);
    const auto rendererCopy = tokenizer.tokenize(s->rendererText);
    s->filter = instantiator::makeFilter(filmCopy);
    s->camera = instantiator::makeCamera(s->cameraData);
    const auto t0 = std::chrono::steady_clock::now();
    s->accelerator = instantiator::makeAccelerator(s->acceleratorData, s->intersectors);
    s->buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    s->counting  = std::make_shared<CountingAccelerator>(s->accelerator);
    // environment light: scene-bound radius, background sphere (renderDatabase.cpp:100-120)
    if (s->envIndex != static_cast<size_t>(-1)) {
        AABB3R sceneBound;
        s->accelerator->evaluateBound(&sceneBound);
        if (!sceneBound.isEmpty()) s->lights[s->envIndex]->setSceneBoundRadius(sceneBound.extent().length() * 0.5f);
    }
    s->cluster   = instantiator::makeLightCluster(s->clusterData, s->lights);
    s->scene     = std::make_shared<Scene>(s->counting, s->cluster);
    if (s->background) s->scene->setBackgroundSphere(s->background.get());
    s->estimator = instantiator::makeEstimator(s->rendererData);
    s->sampler   = instantiator::makeSampler(s->rendererData);
    s->width     = filmCopy->findInt32("image-width", 1024);
    s->height    = filmCopy->findInt32("image-height", 768);
    s->tileX     = filmCopy->findInt32("tile-size-x", 16);
    s->tileY     = filmCopy->findInt32("tile-size-y", 16);
    s->spp       = rendererCopy->findInt32("sample-number", 4);
    s->maxDepth  = rendererCopy->findInt32("max-depth", 10000);
    s->camera->setResolution(Vector2S(static_cast<std::size_t>(s->width), static_cast<std::size_t>(s->height)));
    s->camera->updateTransform();
    return s;
}

void ref_scene_destroy(void* h) { delete static_cast<RefScene*>(h); }

const char* ref_scene_error(void* h) {
    auto* s = static_cast<RefScene*>(h);
    return s->errors.empty() ? "" : s->errors.front().c_str();
}

void ref_scene_info(void* h, int64_t* out /* nprims, nlights, width, height, spp, maxDepth, tileX, tileY */,
                    double* build_seconds) {
    auto* s = static_cast<RefScene*>(h);
    out[0] = static_cast<int64_t>(s->intersectors.size());
    out[1] = static_cast<int64_t>(s->lights.size());
    out[2] = s->width;
    out[3] = s->height;
    out[4] = s->spp;
    out[5] = s->maxDepth;
    out[6] = s->tileX;
    out[7] = s->tileY;
    if (build_seconds) *build_seconds = s->buildSeconds;
}

// BVH topology as BvhAccelerator's constructor builds it (bvhAccelerator.cpp:10-32).
// Returns the node count; fills the arrays when they are non-null.
int64_t ref_bvh_export(void* h, int split_sah, float* bounds, int32_t* is_leaf, int64_t* a, int64_t* b,
                       int32_t* ordered) {
    auto* s = static_cast<RefScene*>(h);
    const BvhBuilder builder(split_sah ? EBvhSplitMode::SAH : EBvhSplitMode::EQUAL);
    std::vector<std::shared_ptr<Intersector>> orderedIntersectors;
    std::size_t total = 0;
    auto root = builder.buildBinaryNodes(s->intersectors, &orderedIntersectors, &total);
    std::vector<BvhLinearNode> nodes;
    builder.buildLinearNodes(std::move(root), total, &nodes);
    if (bounds) {
        for (size_t i = 0; i < nodes.size(); ++i) {
            const AABB3R& bb = nodes[i].bound();
            put3(bounds + 6 * i, bb.minVertex());
            put3(bounds + 6 * i + 3, bb.maxVertex());
            is_leaf[i] = nodes[i].isLeaf() ? 1 : 0;
            a[i] = static_cast<int64_t>(nodes[i].isLeaf() ? nodes[i].intersectorIndex() : nodes[i].secondChildIndex());
            b[i] = static_cast<int64_t>(nodes[i].isLeaf() ? nodes[i].intersectorCounts() : nodes[i].splitAxis());
        }
        for (size_t i = 0; i < orderedIntersectors.size(); ++i) {
            ordered[i] = s->primIndex.at(dynamic_cast<const Primitive*>(orderedIntersectors[i].get()));
        }
    }
    return static_cast<int64_t>(nodes.size());
}

// Accelerator::isIntersecting / isOccluded on a batch of rays.
// rays: n x 8 floats (ox, oy, oz, minT, dx, dy, dz, maxT). The Ray constructor
// normalizes the direction (ray.cpp:19-27); the direction actually used is
// written to dir_used (n x 3) when non-null.
void ref_trace(void* h, const float* rays, int64_t n, int any_hit, int32_t* out_prim, float* out_t,
               float* dir_used) {
    auto* s = static_cast<RefScene*>(h);
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray(v3(r), v3(r + 4), r[3], r[7]);
        if (dir_used) put3(dir_used + 3 * i, ray.direction());
        if (any_hit) {
            out_prim[i] = s->accelerator->isOccluded(ray) ? 1 : 0;
            out_t[i]    = 0.0f;
        } else {
            PrimitiveInfo info;
            if (s->accelerator->isIntersecting(ray, info)) {
                out_prim[i] = s->primIndex.at(info.primitive());
                out_t[i]    = ray.maxT();
            } else {
                out_prim[i] = -1;
                out_t[i]    = ray.maxT();
            }
        }
    }
}

// Full surface interaction of a closest hit (Scene::isIntersecting, scene.cpp:30-65):
// out per ray: prim, t, P(3), Ng(3), Ns(3), uvw(3) -> 14 floats (prim as float;
// -2 for the background sphere of an environment light, -1 for no hit).
void ref_intersect_detail(void* h, const float* rays, int64_t n, float* out) {
    auto* s = static_cast<RefScene*>(h);
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray(v3(r), v3(r + 4), r[3], r[7]);
        SurfaceIntersection si;
        float* o = out + 14 * i;
        if (s->scene->isIntersecting(ray, si)) {
            // the environment's background sphere is reported as primitive -2
            const auto it = s->primIndex.find(si.primitiveInfo().primitive());
            o[0] = it == s->primIndex.end() ? -2.0f : static_cast<float>(it->second);
            o[1] = ray.maxT();
            put3(o + 2, si.surfaceDetail().position());
            put3(o + 5, si.surfaceDetail().geometryNormal());
            put3(o + 8, si.surfaceDetail().shadingNormal());
            put3(o + 11, si.surfaceDetail().uvw());
        } else {
            for (int k = 0; k < 14; ++k) o[k] = 0.0f;
            o[0] = -1.0f;
        }
    }
}

// PerspectivePinholeCamera::spawnPrimaryRay for film positions (n x 2).
void ref_camera_rays(void* h, const float* film_xy, int64_t n, float* out /* n x 6: o, d */) {
    auto* s = static_cast<RefScene*>(h);
    for (int64_t i = 0; i < n; ++i) {
        Ray ray;
        const Vector2R fp(film_xy[2 * i], film_xy[2 * i + 1]);
        s->camera->spawnPrimaryRay(fp.asType<float64>(), &ray);
        put3(out + 6 * i, ray.origin());
        put3(out + 6 * i + 3, ray.direction());
    }
}

// PtEstimator::estimate per sample, the primary ray spawned at film_xy, with the
// random stream keyed by keys[i]. Writes rgb and the number of uniforms drawn.
void ref_estimate(void* h, const float* film_xy, const uint64_t* keys, int64_t n, float* out_rgb,
                  int64_t* out_draws) {
    auto* s = static_cast<RefScene*>(h);
    for (int64_t i = 0; i < n; ++i) {
        Ray ray;
        const Vector2R fp(film_xy[2 * i], film_xy[2 * i + 1]);
        s->camera->spawnPrimaryRay(fp.asType<float64>(), &ray);
        refharness::rng_set_stream(keys[i]);
        Spectrum radiance;
        s->estimator->estimate(*s->scene, ray, &radiance);
        Vector3R rgb;
        radiance.transformToRgb(&rgb);
        put3(out_rgb + 3 * i, rgb);
        if (out_draws) out_draws[i] = static_cast<int64_t>(refharness::g_rng.draws);
    }
}

// One film tile of the sampling renderer (estimatorTileWork.cpp:32-61): per pixel the
// sampler's jitter comes from stream (seed, pixel, 2^40) and sample k's path from
// stream (seed, pixel, k). Accumulates the tile's sensors into film (w*h*4).
static void render_tile(RefScene* s, int tileIndex, uint64_t seed, float* film) {
    const int ntx = (s->width + s->tileX - 1) / s->tileX;
    const int x0 = (tileIndex % ntx) * s->tileX, y0 = (tileIndex / ntx) * s->tileY;
    const int x1 = std::min(x0 + s->tileX, s->width), y1 = std::min(y0 + s->tileY, s->height);
    FilmTile tile(AABB2I(Vector2I(x0, y0), Vector2I(x1, y1)), s->filter.get());
    for (int iy = y0; iy < y1; ++iy) {
        for (int ix = x0; ix < x1; ++ix) {
            const uint64_t pixel = static_cast<uint64_t>(iy) * s->width + ix;
            refharness::rng_set_stream(seqrng_key(seed, pixel, 1ull << 40));
            auto sampler  = s->sampler->clone(s->sampler->sampleNumber());
            auto sample2D = sampler->requestSample2D();
            for (std::size_t k = 0; k < sampler->sampleNumber(); ++k) {
                const Vector2R jitter = Vector2I(ix, iy).asType<real>() + sample2D->nextSample();
                Ray primary;
                s->camera->spawnPrimaryRay(jitter.asType<float64>(), &primary);
                refharness::rng_set_stream(seqrng_key(seed, pixel, k));
                Spectrum radiance;
                s->estimator->estimate(*s->scene, primary, &radiance);
                tile.addSample(jitter, radiance);
            }
        }
    }
    for (int iy = y0; iy < y1; ++iy) {
        for (int ix = x0; ix < x1; ++ix) {
            const RgbRadianceSensor& sensor = tile.getSensor(ix - x0, iy - y0);
            float* p = film + 4 * (static_cast<size_t>(iy) * s->width + ix);
            p[0] += sensor.r();
            p[1] += sensor.g();
            p[2] += sensor.b();
            p[3] += sensor.weight();
        }
    }
}

// Linear film (sum of w*rgb, sum of w per pixel) of the whole image. Tiles touch
// disjoint pixels (the filter footprint is clamped to the tile), so the result
// does not depend on the thread count.
void ref_render_linear(void* h, uint64_t seed, int nthreads, float* film) {
    auto* s = static_cast<RefScene*>(h);
    const int ntiles = ((s->width + s->tileX - 1) / s->tileX) * ((s->height + s->tileY - 1) / s->tileY);
    for (size_t i = 0; i < static_cast<size_t>(s->width) * s->height * 4; ++i) film[i] = 0.0f;
    std::atomic<int> next(0);
    auto worker = [&]() {
        int t;
        while ((t = next.fetch_add(1)) < ntiles) render_tile(s, t, seed, film);
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < std::max(1, nthreads); ++i) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
}

// CPU baseline timing: pixels [begin, end) in row-major order, spp samples each
// (uniform jitter), work split over nthreads like Parallel::execute's static
// contiguous chunks (parallel.cpp:11-31). Returns wall seconds; ray counts in
// out_rays[0] (closest-hit) and out_rays[1] (any-hit).
double ref_time_render(void* h, int64_t begin, int64_t end, int spp, int nthreads, uint64_t seed,
                       int64_t* out_rays, int64_t* out_samples) {
    auto* s = static_cast<RefScene*>(h);
    nthreads = std::max(1, nthreads);
    std::vector<int64_t> rays(2 * nthreads, 0), samples(nthreads, 0);
    const int64_t total = end - begin;
    const int64_t chunk = (total + nthreads - 1) / nthreads;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int w = 0; w < nthreads; ++w) {
        pool.emplace_back([&, w]() {
            t_rayCount[0] = t_rayCount[1] = 0;
            const int64_t b = begin + w * chunk, e = std::min(end, b + chunk);
            for (int64_t p = b; p < e; ++p) {
                const int ix = static_cast<int>(p % s->width), iy = static_cast<int>(p / s->width);
                for (int k = 0; k < spp; ++k) {
                    refharness::rng_set_stream(seqrng_key(seed, static_cast<uint64_t>(p), static_cast<uint64_t>(k)));
                    const Vector2R fp(ix + cadise::Random::nextReal(), iy + cadise::Random::nextReal());
                    Ray primary;
                    s->camera->spawnPrimaryRay(fp.asType<float64>(), &primary);
                    Spectrum radiance;
                    s->estimator->estimate(*s->scene, primary, &radiance);
                    samples[w] += 1;
                }
            }
            rays[2 * w]     = t_rayCount[0];
            rays[2 * w + 1] = t_rayCount[1];
        });
    }
    for (auto& th : pool) th.join();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    out_rays[0] = out_rays[1] = 0;
    *out_samples = 0;
    for (int w = 0; w < nthreads; ++w) {
        out_rays[0] += rays[2 * w];
        out_rays[1] += rays[2 * w + 1];
        *out_samples += samples[w];
    }
    return secs;
}

// Bsdf::evaluate / evaluatePdfW / evaluateSample of the named material at one
// shading point. in: n x (12 + nu) (Ns, wi, wo, uvw, then the uniforms the sample
// draws, in draw order); out: n x 12 (f(3), pdf, sample_f(3), sample_L(3),
// sample_pdf, valid).
int ref_bsdf_query_n(void* h, const char* material, const float* in, int64_t n, int nu, float* out) {
    auto* s = static_cast<RefScene*>(h);
    const auto it = s->bsdfs.find(std::string(material));
    if (it == s->bsdfs.end()) return -1;
    const Bsdf* bsdf = it->second.get();
    for (int64_t i = 0; i < n; ++i) {
        const float* q = in + (12 + nu) * i;
        float*       o = out + 12 * i;
        SurfaceDetail detail;
        detail.setShadingNormal(v3(q));
        detail.setGeometryNormal(v3(q));
        detail.setUvw(v3(q + 9));
        SurfaceIntersection si;
        si.setSurfaceDetail(detail);
        si.setWi(v3(q + 3));
        si.setWo(v3(q + 6));
        Vector3R f;
        bsdf->evaluate(TransportInfo(), si).transformToRgb(&f);
        put3(o, f);
        o[3] = bsdf->evaluatePdfW(TransportInfo(), si);
        refharness::rng_set_queue(q + 12, static_cast<size_t>(nu));
        BsdfSample sample;
        bsdf->evaluateSample(TransportInfo(ETransportMode::RADIANCE), si, &sample);
        Vector3R sf;
        sample.scatterValue().transformToRgb(&sf);
        put3(o + 4, sf);
        put3(o + 7, sample.scatterDirection());
        o[10] = sample.pdfW();
        o[11] = sample.isValid() ? 1.0f : 0.0f;
    }
    refharness::rng_set_stream(0);
    return 0;
}

int ref_bsdf_query(void* h, const char* material, const float* in, int64_t n, float* out) {
    return ref_bsdf_query_n(h, material, in, n, 2, out);
}

// Light::evaluateDirectSample for light index li towards target P, with the given
// uniforms (triangles take 2; rectangles may take more, rejection loop).
// in: n x (3 + nu); out: n x 11 (emitPos(3), emitN(3), radiance(3), pdfW, valid).
int ref_light_sample(void* h, int li, const float* in, int64_t n, int nu, float* out) {
    auto* s = static_cast<RefScene*>(h);
    if (li < 0 || li >= static_cast<int>(s->lights.size())) return -1;
    const Light* light = s->lights[li].get();
    for (int64_t i = 0; i < n; ++i) {
        const float* q = in + (3 + nu) * i;
        float*       o = out + 11 * i;
        refharness::rng_set_queue(q + 3, static_cast<size_t>(nu));
        DirectLightSample ds;
        ds.setTargetPosition(v3(q));
        light->evaluateDirectSample(&ds);
        put3(o, ds.emitPosition());
        put3(o + 3, ds.emitNormal());
        Vector3R rad;
        ds.radiance().transformToRgb(&rad);
        put3(o + 6, rad);
        o[9]  = ds.pdfW();
        o[10] = ds.isValid() ? 1.0f : 0.0f;
    }
    refharness::rng_set_stream(0);
    return 0;
}

// Emittance and evaluateDirectPdfW of the light bound to the primitive hit by each
// ray, seen from the ray origin (directLightEvaluator.cpp:88-103).
// out: n x 5 (light index or -1, Le(3), pdfW).
void ref_light_hit(void* h, const float* rays, int64_t n, float* out) {
    auto* s = static_cast<RefScene*>(h);
    for (int64_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray(v3(r), v3(r + 4), r[3], r[7]);
        SurfaceIntersection si;
        float* o = out + 5 * i;
        o[0] = -1.0f; o[1] = o[2] = o[3] = o[4] = 0.0f;
        if (!s->scene->isIntersecting(ray, si)) continue;
        const AreaLight* al = si.primitiveInfo().primitive()->areaLight();
        if (!al) continue;
        const Light* l = al;
        o[0] = static_cast<float>(s->lightIndex.at(l));
        Vector3R le;
        al->emittance(si).transformToRgb(&le);
        put3(o + 1, le);
        o[4] = al->evaluateDirectPdfW(si, v3(r));
    }
}

// PictureLoader::loadRgbImage: the texels an image texture of that file holds
// (width x height x 3 floats, as the HdrImage stores them). out may be NULL
// to query the size.
void ref_load_rgb_image(const char* path, int32_t* width, int32_t* height, float* out) {
    const HdrImage img = PictureLoader::loadRgbImage(Path(std::string(path)));
    *width = img.width();
    *height = img.height();
    if (!out) return;
    const real* d = img.rawData();
    for (size_t i = 0; i < img.dataSize(); ++i) out[i] = d[i];
}

// Film::save (film.cpp:159-196) of a given linear film (w*h*4: sum of w*rgb, sum
// of w per pixel) to path, through the scene's own film settings and filter: the
// sensors are loaded into one film tile spanning the image and merged into a
// fresh Film (mergeWithFilmTile adds them to zero sensors), which then writes
// the picture by its extension (PictureSaver::save, pictureSaver.cpp:12-82).
void ref_film_save(void* h, const float* film, int64_t spp, const char* path) {
    auto* s = static_cast<RefScene*>(h);
    Film out(Vector2I(s->width, s->height), Vector2I(s->tileX, s->tileY), Path(std::string(path)), s->filter);
    auto tile = std::make_unique<FilmTile>(AABB2I(Vector2I(0, 0), Vector2I(s->width, s->height)), s->filter.get());
    for (int iy = 0; iy < s->height; ++iy) {
        for (int ix = 0; ix < s->width; ++ix) {
            const float* p = film + 4 * (static_cast<size_t>(iy) * s->width + ix);
            auto& sensor = const_cast<RgbRadianceSensor&>(tile->getSensor(ix, iy));
            sensor.addValue(p[0], p[1], p[2]);
            sensor.addWeight(p[3]);
        }
    }
    out.mergeWithFilmTile(std::move(tile));
    out.save(static_cast<std::size_t>(spp));
}

// Light::approximateFlux and LightCluster::evaluatePickLightPdf of every light.
// out: nlights x 2 (flux, pick pdf).
void ref_light_flux(void* h, float* out) {
    auto* s = static_cast<RefScene*>(h);
    for (size_t i = 0; i < s->lights.size(); ++i) {
        out[2 * i]     = s->lights[i]->approximateFlux();
        out[2 * i + 1] = s->scene->evaluatePickLightPdf(s->lights[i].get());
    }
}

} // extern "C"
