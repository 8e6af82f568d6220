// Host-side scene model: what RenderDatabase assembles from CRSD blocks
// (reference core/renderDatabase.cpp:39-131), flattened into arrays that the
// device upload consumes directly.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mpt {

struct F3 {
    float x = 0.f, y = 0.f, z = 0.f;
};

// A sphere's 9 values are (centre, radius, 0, ...).
enum PrimType : int32_t { PRIM_TRIANGLE = 0, PRIM_RECTANGLE = 1, PRIM_SPHERE = 2 };
// bsdfInstantiator.cpp:177-221: matte-lambertian, conductor-microfacet (ggx),
// phong, plastic (MixedBsdf of Lambertian diffuse-albedo and Phong), mirror
// (SpecularReflection), glass (SpecularDielectric), transmission
// (SpecularTransmission), absorb.
enum MatType : int32_t {
    MAT_LAMBERT = 0, MAT_GGX = 1, MAT_PHONG = 2, MAT_PLASTIC = 3, MAT_MIRROR = 4, MAT_GLASS = 5,
    MAT_TRANSMISSION = 6, MAT_ABSORB = 7
};
enum Estimator : int32_t { EST_PT = 0, EST_VPT = 1 };
enum FilterType : int32_t { FILTER_BOX = 0, FILTER_GAUSSIAN = 1, FILTER_CONE = 2, FILTER_MITCHELL = 3 };

// An image (data), a constant (value), or a checkerboard of two other
// textures (odd >= 0): TCheckerboardTexture::evaluate
// (tCheckerboardTexture.ipp:28-40) takes the even texture where
// floor(u / odd_size) + floor(v / even_size) is even, else the odd one, with
// odd_size = 1 / odd-number and even_size = 1 / even-number.
struct Texture {
    std::vector<float> data;   // empty = constant (or checkerboard)
    int32_t width = 0, height = 0, channels = 3;
    int32_t wrap_clamp = 0;
    F3 value;
    int32_t odd = -1, even = -1;
    float odd_size = 0.0f, even_size = 0.0f;
    bool checker() const { return odd >= 0; }
};

// Deepest nesting of checkerboards in checkerboards.
constexpr int kMaxTextureNesting = 8;

struct Material {
    int32_t type = MAT_LAMBERT;
    int32_t albedo_tex = -1;    // index into HostScene::textures, -1 = constant
    F3      albedo{0.5f, 0.5f, 0.5f};   // lambertian / plastic diffuse / specular albedo
    F3      f0{1.f, 1.f, 1.f};          // schlick f0 (conductor-microfacet, conductor mirror)
    int32_t rough_tex = -1;
    float   roughness = 0.f;
    float   exponent = 32.f;            // phong / plastic specular exponent
    float   ratio = 0.7f;               // plastic diffuse ratio
    int32_t fresnel = 0;                // mirror: 0 schlick conductor (f0), 1 dielectric (iors)
    float   ior_outer = 1.f, ior_inner = 1.5f;
};

// Lobes with a specular part: the pt estimator does no NEE there and counts
// emission at the next hit (ptEstimator.cpp:68-93).
inline bool mat_specular(int32_t type) { return type == MAT_MIRROR || type == MAT_GLASS || type == MAT_TRANSMISSION; }

enum LightType : int32_t { LIGHT_AREA = 0, LIGHT_ENVIRONMENT = 1 };

struct Light {
    int32_t type = LIGHT_AREA;
    int32_t prim = -1;              // area light: the emitting primitive
    F3      color{1.f, 1.f, 1.f};
    float   watt = 1.f;
    int32_t back_face = 0;
    int32_t env_tex = -1;           // environment light: radiance texture (HostScene::textures)
};

struct HostScene {
    // primitives in creation order (RenderDatabase::_intersectors)
    std::vector<int32_t> prim_type;
    std::vector<float>   prim_v;     // 9 per primitive
    std::vector<float>   prim_n;     // 9 per primitive, or empty
    std::vector<float>   prim_uvw;   // 9 per primitive, or empty (NaN first = unset)
    std::vector<int32_t> prim_mat;
    std::vector<Material> materials;
    std::vector<Texture>  textures;
    std::vector<Light>    lights;
    int32_t estimator = EST_PT;      // renderer estimator: pt or vpt
    int32_t cluster = 0;             // 0 uniform, 1 power
    // camera (perspective-pinhole)
    F3 cam_pos, cam_dir{0.f, 0.f, -1.f}, cam_up{0.f, 1.f, 0.f};
    float fov = 45.f, sensor_width_mm = 36.f;
    // film
    int32_t width = 1024, height = 768, tile_x = 16, tile_y = 16;
    int32_t filter = -1;             // -1: none given
    float filter_w = 1.f, filter_h = 1.f, sigma_x = 0.5f, sigma_y = 0.5f, amplitude = 1.f;
    float mitchell_b = 1.f / 3.f, mitchell_c = 1.f / 3.f;
    std::string output = "cadise.jpg";
    // renderer
    int32_t sampler = 1;             // 0 random, 1 stratified
    int32_t spp = 4, max_depth = 10000;
    int32_t split_sah = 1;

    int64_t nprims() const { return static_cast<int64_t>(prim_type.size()); }
};

// Camera matrices exactly as PerspectivePinholeCamera::updateTransform and
// TMatrix4::lookAt compute them (perspectivePinholeCamera.cpp:32-50,
// tMatrix4.ipp:47-60), row-major float 4x4.
void camera_matrices(const HostScene& s, float film_to_camera[16], float camera_to_world[16]);

// Emitted radiance of each light (SingleAreaLight ctor, singleAreaLight.cpp:20-37).
std::vector<F3> light_radiance(const HostScene& s);
// Surface area of a primitive (triangle.cpp:199-201, rectangle.cpp:179-181).
float prim_area(const HostScene& s, int64_t i);
// Distribution1D's CDF (distribution1D.cpp:10-40) of n values: n + 1 entries.
void distribution_cdf(const float* value, size_t n, float* cdf);

// EnvironmentLight's sampling tables (environmentLight.cpp:16-56,
// distribution2D.cpp:9-28): marginal CDF over rows (height + 1), then one
// conditional CDF per row (width + 1 each), and the background flux (sum of
// luminance * sin(theta) over texels).
struct EnvTables {
    std::vector<float> cdf;
    float background_flux = 0.f;
};
EnvTables env_tables(const Texture& t);

// The reference BVH's root bound (the union of every primitive's padded bound,
// built on the reference's empty box), and the environment light's scene
// radius from it (renderDatabase.cpp:99-111): |extent| / 2, or 5000 when empty.
float env_scene_radius(const HostScene& s);

// Power light-cluster CDF (powerLightCluster.cpp:9-25, distribution1D.cpp:10-40).
std::vector<float> power_cdf(const HostScene& s, const std::vector<F3>& radiance);
// StratifiedSampler's base number (stratifiedSampler.cpp:14-20).
int32_t strata_base(int32_t spp);

}  // namespace mpt
