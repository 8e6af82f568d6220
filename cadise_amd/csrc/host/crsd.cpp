// CRSD reader (see crsd.hpp). Numbers go through std::stold then a cast to
// float, as the reference tokenizer does (tokenizer.cpp:134-246), so a scene
// parses to the same binary32 values.
#include "crsd.hpp"
#include "hdr.hpp"
#include "image.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <map>
#include <string_view>
#include <vector>

#include "../../../include/mpt.h"

namespace mpt {

namespace {

struct Unit {
    std::string name, type, value;
};

struct Block {
    std::string       cls;
    std::vector<Unit> units;

    const Unit* find(std::string_view name, std::initializer_list<std::string_view> types) const {
        // SdData::_findData: the first unit of that name in the typed list wins
        for (const Unit& u : units) {
            if (u.name != name) continue;
            for (auto t : types)
                if (u.type == t) return &u;
        }
        return nullptr;
    }
    std::string str(std::string_view name, std::string_view def = "") const {
        const Unit* u = find(name, {"string", "material", "texture", "primitive"});
        return u ? u->value : std::string(def);
    }
    float real(std::string_view name, float def = 0.f) const {
        const Unit* u = find(name, {"real"});
        if (!u) return def;
        try { return static_cast<float>(std::stold(u->value)); } catch (...) { return def; }
    }
    int32_t i32(std::string_view name, int32_t def = 0) const {
        const Unit* u = find(name, {"int32"});
        if (!u) return def;
        try { return static_cast<int32_t>(std::stoll(u->value)); } catch (...) { return def; }
    }
    bool boolean(std::string_view name, bool def = false) const {
        const Unit* u = find(name, {"bool"});
        return u ? (u->value == "true") : def;
    }
    bool has_vec(std::string_view name) const { return find(name, {"rgb", "vector3r"}) != nullptr; }
    F3 vec(std::string_view name, F3 def = F3{}) const;
    std::vector<float> vec_array(std::string_view name) const;
};

// Tokenizer::_parseVector3r (tokenizer.cpp:140-172)
F3 parse_vec3(const std::string& v) {
    size_t s = v.find_first_not_of(' ');
    size_t e = v.find_first_of(' ', s);
    const std::string x = v.substr(s, e - s);
    s = v.find_first_not_of(' ', e + 1);
    e = v.find_first_of(' ', s);
    const std::string y = v.substr(s, e - s);
    s = v.find_first_not_of(' ', e + 1);
    e = v.find_last_not_of(' ');
    const std::string z = v.substr(s, e - s + 1);
    return {static_cast<float>(std::stold(x)), static_cast<float>(std::stold(y)), static_cast<float>(std::stold(z))};
}

F3 Block::vec(std::string_view name, F3 def) const {
    const Unit* u = find(name, {"rgb", "vector3r"});
    if (!u) return def;
    try { return parse_vec3(u->value); } catch (...) { return def; }
}

// Tokenizer::_parseVector3rArray (tokenizer.cpp:204-246): "x y z, x y z, ..."
std::vector<float> Block::vec_array(std::string_view name) const {
    std::vector<float> out;
    const Unit* u = find(name, {"vector3r-array"});
    if (!u) return out;
    const std::string& v = u->value;
    size_t s = 0;
    while ((s = v.find_first_not_of(' ', s)) != std::string::npos) {
        size_t e = v.find_first_of(',', s);
        if (e == std::string::npos) e = v.find_last_not_of(' ') + 1;
        size_t t = v.find_first_of(' ', s);
        const std::string x = v.substr(s, t - s);
        s = v.find_first_not_of(' ', t + 1);
        t = v.find_first_of(' ', s);
        const std::string y = v.substr(s, t - s);
        s = v.find_first_not_of(' ', t + 1);
        const std::string z = v.substr(s, e - s);
        out.push_back(static_cast<float>(std::stold(x)));
        out.push_back(static_cast<float>(std::stold(y)));
        out.push_back(static_cast<float>(std::stold(z)));
        s = e + 1;
    }
    return out;
}

// Tokenizer::tokenize (tokenizer.cpp:12-43) + _parseDataUnit (tokenizer.cpp:45-114)
Block tokenize(std::string_view d) {
    Block b;
    size_t s = 0;
    while (s < d.size() && !std::isalpha(static_cast<unsigned char>(d[s]))) ++s;
    size_t e = d.find_first_of('{', s);
    if (e == std::string_view::npos) return b;
    while (e > s && !std::isalpha(static_cast<unsigned char>(d[e]))) --e;
    b.cls = std::string(d.substr(s, e - s + 1));
    while ((s = d.find_first_of('[', e)) != std::string_view::npos) {
        e = d.find_first_of(']', s);
        if (e == std::string_view::npos) break;
        std::string_view unit = d.substr(s + 1, e - s - 1);
        Unit u;
        size_t a = unit.find_first_not_of(' ');
        size_t z = unit.find_first_of(' ', a);
        u.name = std::string(unit.substr(a, z - a));
        a = unit.find_first_not_of(' ', z + 1);
        z = unit.find_first_of(' ', a);
        u.type = std::string(unit.substr(a, z - a));
        a = unit.find_first_of('"', z);
        z = unit.find_last_of('"');
        if (a != std::string_view::npos && z != std::string_view::npos && z > a)
            u.value = std::string(unit.substr(a + 1, z - a - 1));
        b.units.push_back(std::move(u));
        s = e + 1;
    }
    return b;
}

bool read_f32(const std::string& path, std::vector<float>* out) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return false;
    const std::streamsize bytes = f.tellg();
    f.seekg(0);
    out->resize(static_cast<size_t>(bytes) / sizeof(float));
    f.read(reinterpret_cast<char*>(out->data()), static_cast<std::streamsize>(out->size() * sizeof(float)));
    return static_cast<bool>(f);
}

std::string resolve(const std::string& base, const std::string& p) {
    if (p.empty() || p[0] == '/' || base.empty()) return p;
    return base + "/" + p;
}

struct Db {
    HostScene& s;
    std::string base;
    std::map<std::string, int32_t, std::less<>> mats, real_tex, spec_tex, prims;
    bool have_film = false, have_camera = false, have_renderer = false, have_accel = false, have_cluster = false;
    Block film, camera, renderer, accel, cluster;
    CrsdResult err;

    bool fail(int code, std::string msg) {
        err.code = code;
        err.message = std::move(msg);
        return false;
    }

    // A checkerboard's odd / even texture: SdData::getSpectrumTexture /
    // getRealTexture (sdData.cpp:98-140) give the named texture of the same
    // kind, else a constant of the field's value (default 0).
    int32_t sub_texture(const Block& b, std::string_view key, int channels) {
        const std::string name = b.str(key);
        if (!name.empty()) {
            const auto& tab = channels == 3 ? spec_tex : real_tex;
            auto it = tab.find(name);
            if (it == tab.end()) {
                fail(MPT_ERR_PARSE, "unknown texture <" + name + ">");
                return -1;
            }
            return it->second;
        }
        Texture t;
        t.channels = channels;
        if (channels == 3) t.value = b.vec(key, F3{0.f, 0.f, 0.f});
        else { const float v = b.real(key, 0.0f); t.value = {v, v, v}; }
        s.textures.push_back(std::move(t));
        return static_cast<int32_t>(s.textures.size() - 1);
    }

    int nesting(int32_t id) const {
        const Texture& t = s.textures[id];
        return t.checker() ? 1 + std::max(nesting(t.odd), nesting(t.even)) : 0;
    }

    // instantiator::makeRealTexture / makeSpectrumTexture (textureInstantiator.cpp:118-176)
    int32_t add_texture(const Block& b, int channels) {
        const std::string type = b.str("type");
        Texture t;
        t.channels = channels;
        if (type == "constant") {
            if (channels == 3) t.value = b.vec("value");
            else { const float v = b.real("value"); t.value = {v, v, v}; }
        } else if (type == "image" && channels == 3) {
            // createSpectrumImage (textureInstantiator.cpp:70-103): nearest sampling
            // whatever the sample-mode, repeat wrap unless wrap-mode is "clamp"
            std::string e;
            const std::string fn = resolve(base, b.str("filename"));
            if (!rgb_image_supported(fn)) {
                fail(MPT_ERR_UNSUPPORTED, "image <" + fn + ">: only .png, .tga and .hdr pictures are read");
                return -1;
            }
            if (!load_rgb_image(fn, &t.width, &t.height, &t.data, &e)) {
                fail(MPT_ERR_IO, e);
                return -1;
            }
            t.wrap_clamp = b.str("wrap-mode") == "clamp" ? 1 : 0;
        } else if (type == "checkerboard") {
            // createRealCheckerboard / createSpectrumCheckerboard (textureInstantiator.cpp:31-68)
            const int32_t odd = sub_texture(b, "odd-texture", channels);
            if (odd < 0) return -1;
            const int32_t even = sub_texture(b, "even-texture", channels);
            if (even < 0) return -1;
            if (std::max(nesting(odd), nesting(even)) + 1 > kMaxTextureNesting) {
                fail(MPT_ERR_UNSUPPORTED, "checkerboards nested deeper than 8");
                return -1;
            }
            t.odd = odd;
            t.even = even;
            t.odd_size = 1.0f / b.real("odd-number");
            t.even_size = 1.0f / b.real("even-number");
        } else if (type == "image-raw") {
            t.width = b.i32("width");
            t.height = b.i32("height");
            t.wrap_clamp = b.str("wrap-mode") == "clamp" ? 1 : 0;
            if (!read_f32(resolve(base, b.str("filename")), &t.data) ||
                t.data.size() != static_cast<size_t>(t.width) * t.height * channels) {
                fail(MPT_ERR_IO, "cannot read image-raw texture " + b.str("filename"));
                return -1;
            }
        } else if (type == "image") {
            fail(MPT_ERR_UNSUPPORTED, "real image textures: the reference leaves them unimplemented "
                                      "(textureInstantiator.cpp:133-136)");
            return -1;
        } else {
            fail(MPT_ERR_UNSUPPORTED, "texture type <" + type + "> is not on the path-tracing path");
            return -1;
        }
        s.textures.push_back(std::move(t));
        return static_cast<int32_t>(s.textures.size() - 1);
    }

    // SdData::getSpectrumTexture / getRealTexture (sdData.cpp:98-140): a named
    // texture, else a constant (default 0)
    bool spectrum_param(const Block& b, std::string_view key, int32_t* tex, F3* value) {
        const std::string name = b.str(key);
        if (!name.empty()) {
            auto it = spec_tex.find(name);
            if (it == spec_tex.end()) return fail(MPT_ERR_PARSE, "unknown texture <" + name + ">");
            *tex = it->second;
        } else {
            *value = b.vec(key, F3{0.f, 0.f, 0.f});
        }
        return true;
    }

    // instantiator::makeBsdf (bsdfInstantiator.cpp:28-221)
    bool material(const Block& b) {
        const std::string type = b.str("type");
        Material m;
        if (type == "matte-lambertian") {
            m.type = MAT_LAMBERT;
            if (!spectrum_param(b, "albedo", &m.albedo_tex, &m.albedo)) return false;
        } else if (type == "conductor-microfacet") {
            m.type = MAT_GGX;
            const std::string mf = b.str("microfacet-type", "ggx");
            if (mf == "beckmann" || mf == "blinn-phong")
                return fail(MPT_ERR_UNSUPPORTED, "microfacet-type <" + mf + "> is not on the path-tracing path");
            m.f0 = b.vec("f0", F3{1.f, 1.f, 1.f});
            const std::string tex = b.str("roughness");
            if (!tex.empty()) {
                auto it = real_tex.find(tex);
                if (it == real_tex.end()) return fail(MPT_ERR_PARSE, "unknown texture <" + tex + ">");
                m.rough_tex = it->second;
            } else {
                m.roughness = b.real("roughness", 0.f);
            }
        } else if (type == "phong") {
            m.type = MAT_PHONG;
            m.exponent = b.real("exponent", 32.f);
        } else if (type == "plastic") {
            m.type = MAT_PLASTIC;
            if (!spectrum_param(b, "diffuse-albedo", &m.albedo_tex, &m.albedo)) return false;
            m.exponent = b.real("specular-exponent", 32.f);
            m.ratio = b.real("diffuse-ratio", 0.7f);
        } else if (type == "mixed") {
            // MixedBsdf over two named materials: on this path, Lambertian A and Phong B
            // (the plastic combination)
            auto ia = mats.find(b.str("bsdf-a")), ib = mats.find(b.str("bsdf-b"));
            if (ia == mats.end() || ib == mats.end()) return fail(MPT_ERR_PARSE, "mixed material names an unknown bsdf");
            const Material& A = s.materials[ia->second];
            const Material& B = s.materials[ib->second];
            if (A.type != MAT_LAMBERT || B.type != MAT_PHONG)
                return fail(MPT_ERR_UNSUPPORTED, "mixed materials other than (matte-lambertian, phong)");
            m = A;
            m.type = MAT_PLASTIC;
            m.exponent = B.exponent;
            m.ratio = b.real("ratio", 0.5f);
        } else if (type == "mirror") {
            m.type = MAT_MIRROR;
            if (!spectrum_param(b, "albedo", &m.albedo_tex, &m.albedo)) return false;
            const std::string fr = b.str("fresnel-type", "conductor");
            if (fr == "conductor") {
                m.fresnel = 0;
                m.f0 = b.vec("f0", F3{1.f, 1.f, 1.f});
            } else if (fr == "dielectric") {
                m.fresnel = 1;
                m.ior_outer = b.real("ior-outer", 1.f);
                m.ior_inner = b.real("ior-inner", 1.5f);
            } else {
                return fail(MPT_ERR_PARSE, "unknown fresnel-type <" + fr + "> in mirror");
            }
        } else if (type == "glass" || type == "transmission") {
            m.type = type == "glass" ? MAT_GLASS : MAT_TRANSMISSION;
            if (!spectrum_param(b, "albedo", &m.albedo_tex, &m.albedo)) return false;
            m.fresnel = 1;
            m.ior_outer = b.real("ior-outer", 1.f);
            m.ior_inner = b.real("ior-inner", type == "glass" ? 1.5f : 1.f);
        } else if (type == "absorb") {
            m.type = MAT_ABSORB;
        } else {
            return fail(MPT_ERR_UNSUPPORTED, "material type <" + type + "> is not on the path-tracing path");
        }
        s.materials.push_back(m);
        mats.emplace(b.str("name"), static_cast<int32_t>(s.materials.size() - 1));  // map::insert: first wins
        return true;
    }

    int32_t default_lambert() {
        // LambertianDiffuse() default: constant albedo 0.5 (lambertianDiffuse.cpp:14-17)
        Material m;
        m.type = MAT_LAMBERT;
        m.albedo = {0.5f, 0.5f, 0.5f};
        s.materials.push_back(m);
        return static_cast<int32_t>(s.materials.size() - 1);
    }

    void push_prim(int32_t type, const float* v, int32_t mat, const float* n, const float* uvw) {
        s.prim_type.push_back(type);
        s.prim_v.insert(s.prim_v.end(), v, v + 9);
        s.prim_mat.push_back(mat);
        const size_t k = s.prim_type.size() - 1;
        if (n) {
            if (s.prim_n.empty()) s.prim_n.assign(9 * k, 0.f);
            s.prim_n.insert(s.prim_n.end(), n, n + 9);
        } else if (!s.prim_n.empty()) {
            s.prim_n.insert(s.prim_n.end(), 9, 0.f);
        }
        if (uvw) {
            if (s.prim_uvw.empty()) {
                s.prim_uvw.assign(9 * k, 0.f);
                for (size_t i = 0; i < k; ++i) s.prim_uvw[9 * i] = std::nanf("");
            }
            s.prim_uvw.insert(s.prim_uvw.end(), uvw, uvw + 9);
        } else if (!s.prim_uvw.empty()) {
            const size_t at = s.prim_uvw.size();
            s.prim_uvw.insert(s.prim_uvw.end(), 9, 0.f);
            s.prim_uvw[at] = std::nanf("");
        }
    }

    bool primitive(const Block& b) {
        const std::string type = b.str("type");
        const std::string name = b.str("name");
        auto mit = mats.find(b.str("bsdf"));
        const int32_t first = static_cast<int32_t>(s.prim_type.size());
        if (type == "sphere") {
            // createSphere (primitiveInstantiator.cpp:13-23)
            if (mit == mats.end()) return fail(MPT_ERR_PARSE, "sphere <" + name + "> names an unknown material");
            const F3 c = b.vec("center");
            const float v[9] = {c.x, c.y, c.z, b.real("radius"), 0.f, 0.f, 0.f, 0.f, 0.f};
            push_prim(PRIM_SPHERE, v, mit->second, nullptr, nullptr);
        } else if (type == "triangle" || type == "rectangle") {
            int32_t mat;
            if (mit != mats.end()) mat = mit->second;
            else if (type == "rectangle") mat = default_lambert();
            else return fail(MPT_ERR_PARSE, "triangle <" + name + "> names an unknown material");
            const F3 a = b.vec("v1"), c = b.vec("v2"), d = b.vec("v3");
            const float v[9] = {a.x, a.y, a.z, c.x, c.y, c.z, d.x, d.y, d.z};
            push_prim(type == "rectangle" ? PRIM_RECTANGLE : PRIM_TRIANGLE, v, mat, nullptr, nullptr);
        } else if (type == "triangle-mesh") {
            const int32_t mat = mit != mats.end() ? mit->second : default_lambert();
            std::vector<float> pos, nrm, uvw;
            const std::string pf = b.str("positions-file");
            if (!pf.empty()) {
                if (!read_f32(resolve(base, pf), &pos)) return fail(MPT_ERR_IO, "cannot read " + pf);
                const std::string nf = b.str("normals-file"), uf = b.str("uvws-file");
                if (!nf.empty() && !read_f32(resolve(base, nf), &nrm)) return fail(MPT_ERR_IO, "cannot read " + nf);
                if (!uf.empty() && !read_f32(resolve(base, uf), &uvw)) return fail(MPT_ERR_IO, "cannot read " + uf);
            } else {
                try {
                    pos = b.vec_array("positions");
                    nrm = b.vec_array("normals");
                    uvw = b.vec_array("uvws");
                } catch (...) {
                    return fail(MPT_ERR_PARSE, "malformed vector3r-array in <" + name + ">");
                }
            }
            // TriangleMesh::transformToTriangles: one triangle per 3 positions
            const size_t ntri = pos.size() / 9;
            if ((!nrm.empty() && nrm.size() < 9 * ntri) || (!uvw.empty() && uvw.size() < 9 * ntri))
                return fail(MPT_ERR_PARSE, "normals / uvws shorter than positions in <" + name + ">");
            for (size_t t = 0; t < ntri; ++t)
                push_prim(PRIM_TRIANGLE, &pos[9 * t], mat, nrm.empty() ? nullptr : &nrm[9 * t],
                          uvw.empty() ? nullptr : &uvw[9 * t]);
        } else {
            return fail(MPT_ERR_UNSUPPORTED, "primitive type <" + type + "> is not on the path-tracing path");
        }
        if (static_cast<int32_t>(s.prim_type.size()) > first) prims.emplace(name, first);  // first wins
        return true;
    }

    bool light(const Block& b) {
        const std::string type = b.str("type");
        if (type == "environment") return environment(b);
        if (type != "area") return fail(MPT_ERR_UNSUPPORTED, "light type <" + type + "> is not on the path-tracing path");
        const std::string pn = b.str("primitive");
        auto it = prims.find(pn);
        if (it == prims.end()) return fail(MPT_ERR_PARSE, "light names an unknown primitive <" + pn + ">");
        Light l;
        l.prim = it->second;
        l.color = b.vec("color");
        l.watt = b.real("watt");
        l.back_face = b.boolean("is-back-face-emit") ? 1 : 0;
        s.lights.push_back(l);
        return true;
    }

    // lightInstantiator.cpp:52-76: the radiance image is loaded with its rows
    // flipped (stbi_set_flip_vertically_on_load) and then flipped horizontally
    // (TImage::flipHorizontal), and sampled nearest with repeat wrap.
    bool environment(const Block& b) {
        for (const Light& l : s.lights)
            if (l.type == LIGHT_ENVIRONMENT) return fail(MPT_ERR_UNSUPPORTED, "more than one environment light");
        const std::string fn = b.str("hdr-filename");
        if (fn.size() < 4 || (fn.compare(fn.size() - 4, 4, ".hdr") != 0 && fn.compare(fn.size() - 4, 4, ".HDR") != 0))
            return fail(MPT_ERR_UNSUPPORTED, "environment image <" + fn + ">: only Radiance .hdr files are read");
        std::vector<float> raw;
        int w = 0, h = 0;
        std::string e;
        if (!read_hdr(resolve(base, fn), &raw, &w, &h, &e)) return fail(MPT_ERR_IO, e);
        Texture t;
        t.width = w;
        t.height = h;
        t.channels = 3;
        t.wrap_clamp = 0;
        t.data.resize(raw.size());
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c)
                    t.data[3 * (static_cast<size_t>(y) * w + x) + c] =
                        raw[3 * (static_cast<size_t>(h - 1 - y) * w + (w - 1 - x)) + c];
        s.textures.push_back(std::move(t));
        Light l;
        l.type = LIGHT_ENVIRONMENT;
        l.env_tex = static_cast<int32_t>(s.textures.size() - 1);
        s.lights.push_back(l);
        return true;
    }

    bool block(const Block& b) {
        if (b.cls == "film") { film = b; have_film = true; return true; }
        if (b.cls == "camera") { camera = b; have_camera = true; return true; }
        if (b.cls == "renderer") { renderer = b; have_renderer = true; return true; }
        if (b.cls == "accelerator") { accel = b; have_accel = true; return true; }
        if (b.cls == "light-cluster") { cluster = b; have_cluster = true; return true; }
        if (b.cls == "texture-real" || b.cls == "texture-spectrum") {
            const bool spec = b.cls == "texture-spectrum";
            const int32_t id = add_texture(b, spec ? 3 : 1);
            if (id < 0) return false;
            (spec ? spec_tex : real_tex).emplace(b.str("name"), id);
            return true;
        }
        if (b.cls == "material") return material(b);
        if (b.cls == "primitive") return primitive(b);
        if (b.cls == "light") return light(b);
        return true;   // unknown class: ignored, as SdData::setClassType leaves it unset
    }

    // RenderDatabase::prepareRender (renderDatabase.cpp:78-131) for this path
    bool finish() {
        if (!have_film || !have_camera || !have_renderer || !have_accel || !have_cluster)
            return fail(MPT_ERR_PARSE, "missing film / camera / renderer / accelerator / light-cluster block");
        s.width = film.i32("image-width", 1024);
        s.height = film.i32("image-height", 768);
        s.tile_x = film.i32("tile-size-x", 16);
        s.tile_y = film.i32("tile-size-y", 16);
        s.output = film.str("output-filename", "cadise.jpg");
        const std::string f = film.str("filter");
        if (f == "box") { s.filter = FILTER_BOX; s.filter_w = film.real("filter-width", 1.f); s.filter_h = film.real("filter-height", 1.f); }
        else if (f == "gaussian") {
            s.filter = FILTER_GAUSSIAN;
            s.filter_w = film.real("filter-width", 4.f); s.filter_h = film.real("filter-height", 4.f);
            s.sigma_x = film.real("sigma-x", 0.5f); s.sigma_y = film.real("sigma-y", 0.5f);
            s.amplitude = film.real("amplitude", 1.f);
        } else if (f == "cone") { s.filter = FILTER_CONE; s.filter_w = film.real("filter-width", 4.f); s.filter_h = film.real("filter-height", 4.f); }
        else if (f == "mitchell") {
            s.filter = FILTER_MITCHELL;
            s.filter_w = film.real("filter-width", 4.f); s.filter_h = film.real("filter-height", 4.f);
            s.mitchell_b = film.real("b", 1.f / 3.f); s.mitchell_c = film.real("c", 1.f / 3.f);
        } else return fail(MPT_ERR_PARSE, "film has no filter (the reference requires one, film.cpp:22)");
        if (s.width <= 0 || s.height <= 0 || s.tile_x <= 0 || s.tile_y <= 0) return fail(MPT_ERR_PARSE, "bad film size");

        const std::string ct = camera.str("type");
        if (ct != "perspective-pinhole") return fail(MPT_ERR_UNSUPPORTED, "camera type <" + ct + ">");
        s.cam_pos = camera.vec("position");
        s.cam_dir = camera.vec("direction");
        s.cam_up = camera.vec("up");
        s.fov = camera.real("fov");
        s.sensor_width_mm = camera.real("sensor-width-mm", 36.f);

        const std::string rt = renderer.str("type");
        if (rt != "sampling") return fail(MPT_ERR_UNSUPPORTED, "renderer <" + rt + "> stays on the CPU (this path is sampling + pt)");
        const std::string est = renderer.str("estimator");
        if (est == "pt") s.estimator = EST_PT;
        else if (est == "vpt") s.estimator = EST_VPT;
        else return fail(MPT_ERR_UNSUPPORTED, "estimator <" + est + "> is not on the path-tracing path");
        const std::string smp = renderer.str("sampler");
        if (smp == "random") s.sampler = 0;
        else if (smp == "stratified") s.sampler = 1;
        else return fail(MPT_ERR_PARSE, "unknown sampler <" + smp + ">");
        s.spp = renderer.i32("sample-number", 4);
        s.max_depth = renderer.i32("max-depth", 10000);
        if (s.spp <= 0 || s.max_depth < 0) return fail(MPT_ERR_PARSE, "bad sample-number / max-depth");

        // any accelerator type is served by the BVH; split-mode picks SAH or equal
        s.split_sah = accel.str("split-mode", "sah") == "equal" ? 0 : 1;

        const std::string cl = cluster.str("type");
        if (cl == "uniform") s.cluster = 0;
        else if (cl == "power") s.cluster = 1;
        else return fail(MPT_ERR_PARSE, "unknown light-cluster <" + cl + ">");
        if (s.lights.empty()) return fail(MPT_ERR_PARSE, "no light (the light cluster needs one, uniformLightCluster.cpp:12)");
        for (const Light& l : s.lights) {
            if (l.type != LIGHT_AREA) continue;
            // Sphere::evaluatePositionSample is unimplemented (sphere.cpp:132-136)
            if (s.prim_type[l.prim] == PRIM_SPHERE)
                return fail(MPT_ERR_UNSUPPORTED, "area light on a sphere (the reference cannot sample it)");
            if (!(prim_area(s, l.prim) > 0.f)) return fail(MPT_ERR_PARSE, "zero-area emitter (singleAreaLight.cpp:30)");
        }
        return true;
    }
};

}  // namespace

CrsdResult parse_crsd(const std::string& text, const std::string& base_dir, HostScene* out) {
    *out = HostScene();
    Db db{*out, base_dir};
    // SdParser::_parseSdToString: '#' starts a comment to the end of the line
    std::string body;
    body.reserve(text.size());
    for (size_t i = 0; i < text.size(); ++i) {
        if (text[i] == '#') {
            while (i < text.size() && text[i] != '\n') ++i;
            continue;
        }
        body.push_back(text[i]);
    }
    // SdParser::parseSd: one block per '}'
    std::string_view view(body);
    size_t start = 0, end;
    try {
        while ((end = view.find_first_of('}', start)) != std::string_view::npos) {
            if (!db.block(tokenize(view.substr(start, end - start + 1)))) return db.err;
            start = end + 1;
        }
    } catch (const std::exception& e) {
        db.err.code = MPT_ERR_PARSE;
        db.err.message = std::string("malformed value: ") + e.what();
        return db.err;
    }
    if (!db.finish()) return db.err;
    return db.err;
}

}  // namespace mpt
