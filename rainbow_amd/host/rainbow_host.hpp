// Host-side mirror of rainbow's render interfaces (C++17), and the drop-in
// integrator that renders through librbw.so.
//
// The classes keep the reference's names, namespaces, constructor arguments
// and ownership (std::shared_ptr everywhere):
//   integrators::integrator::render(camera, scene)   rainbow-cpu/integrators/integrator.hpp:58-71
//   integrators::wavefront_path_integrator           in place of path_integrator (path_integrator.hpp:7-27)
//   scenes::scene / scenes::entity                   scenes/scene.hpp:18-44, scenes/entity.hpp:36-70
//   shapes::mesh / shapes::sphere                    shapes/mesh.hpp:13-38, shapes/sphere.hpp:7-39
//   materials::matte_material / metal_material       materials/matte_material.hpp, metal_material.hpp
//   emitters::surface_light                          emitters/surface_light.hpp:7-13
//   cameras::film / cameras::perspective_camera      cameras/film.hpp:52-91, cameras/perspective_camera.hpp:7-27
//   filters::box_filter, samplers::random_sampler    filters/box_filter.hpp, samplers/random_sampler.hpp
// The reference builds its matrices with glm; this mirror composes them in
// double and rounds to float.  Accessors the reference keeps private (sphere
// radius, material textures, camera matrices) are public here because the
// flattening into the C ABI needs them.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rbw.h"

namespace rainbow::core {
using real = float;
using uint32 = unsigned;
using uint64 = unsigned long long;
namespace math {
struct vector2 { real x = 0, y = 0; };
struct vector2i { int32_t x = 0, y = 0; };
struct vector3 { real x = 0, y = 0, z = 0; };
struct bound2 { vector2 min{0, 0}, max{1, 1}; };
struct bound2i { vector2i min, max; };
// column-major m[col * 4 + row], as glm
struct matrix4x4 {
    std::array<double, 16> m{};
    static matrix4x4 identity() {
        matrix4x4 r;
        for (int i = 0; i < 4; i++) r.m[i * 5] = 1.0;
        return r;
    }
    matrix4x4 operator*(const matrix4x4& b) const {
        matrix4x4 r;
        for (int c = 0; c < 4; c++)
            for (int w = 0; w < 4; w++) {
                double s = 0;
                for (int k = 0; k < 4; k++) s += m[k * 4 + w] * b.m[c * 4 + k];
                r.m[c * 4 + w] = s;
            }
        return r;
    }
    matrix4x4 inverse() const;
};
}  // namespace math
}  // namespace rainbow::core

namespace rainbow::cpus {

using core::real;
using core::math::bound2;
using core::math::bound2i;
using core::math::matrix4x4;
using core::math::vector2;
using core::math::vector2i;
using core::math::vector3;

namespace interfaces {
class noncopyable {
public:
    noncopyable() = default;
    virtual ~noncopyable() = default;
    noncopyable(const noncopyable&) = delete;
    noncopyable& operator=(const noncopyable&) = delete;
};
}  // namespace interfaces

namespace shared {
namespace spectrums {
struct spectrum {
    std::array<real, 3> c{0, 0, 0};
    spectrum() = default;
    spectrum(real v) : c{v, v, v} {}
    spectrum(real r, real g, real b) : c{r, g, b} {}
    real operator[](size_t i) const { return c[i]; }
};
}  // namespace spectrums

// shared/transform.hpp:16-48: a matrix with its inverse
class transform {
public:
    transform() : mTransform(matrix4x4::identity()), mInverse(matrix4x4::identity()) {}
    transform(const matrix4x4& m, const matrix4x4& inv) : mTransform(m), mInverse(inv) {}
    transform operator*(const transform& r) const { return transform(mTransform * r.mTransform, r.mInverse * mInverse); }
    transform inverse() const { return transform(mInverse, mTransform); }
    const matrix4x4& matrix() const noexcept { return mTransform; }
    const matrix4x4& inverse_matrix() const noexcept { return mInverse; }
    bool is_identity() const;
    bool is_translation(vector3& t) const;
private:
    matrix4x4 mTransform, mInverse;
};

transform translate(const vector3& v);
transform scale(const vector3& v);
// shared/transform.cpp:132-137: rotation by `angle` radians about `axis` (glm::rotate)
transform rotate(real angle, const vector3& axis);
transform perspective_right_hand(real fov, real width, real height, real near = 0.01f, real far = 1000.0f);
transform look_at_right_hand(const vector3& origin, const vector3& target, const vector3& up);
}  // namespace shared

namespace textures {
template <typename T>
class constant_texture2d final : public interfaces::noncopyable {
public:
    explicit constant_texture2d(const T& value) : mValue(value) {}
    const T& value() const noexcept { return mValue; }
private:
    T mValue;
};
template <typename T>
using texture2d = constant_texture2d<T>;
template <typename T>
using constant_texture3d = constant_texture2d<T>;
// textures/image_texture.hpp:58 (image_texture_t<3, T>): size.x * size.y * size.z values, x fastest
template <typename T>
class image_texture3d final : public interfaces::noncopyable {
public:
    image_texture3d(std::array<size_t, 3> size, std::vector<T> values) : mSize(size), mValues(std::move(values)) {}
    std::array<size_t, 3> size() const noexcept { return mSize; }
    const std::vector<T>& values() const noexcept { return mValues; }
private:
    std::array<size_t, 3> mSize;
    std::vector<T> mValues;
};
}  // namespace textures

namespace shapes {
class shape : public interfaces::noncopyable, public std::enable_shared_from_this<shape> {
public:
    explicit shape(bool reverse_orientation, size_t count) : mReverseOrientation(reverse_orientation), mCount(count) {}
    bool reverse_orientation() const noexcept { return mReverseOrientation; }
    size_t count() const noexcept { return mCount; }
protected:
    bool mReverseOrientation = false;
    size_t mCount = 0;
};

// shapes/mesh.hpp: per-vertex positions, optional tangents/normals/uvs, 3 indices per face
class mesh final : public shape {
public:
    mesh(const std::vector<vector3>& positions, const std::vector<vector3>& tangents,
         const std::vector<vector3>& normals, const std::vector<vector3>& uvs, const std::vector<unsigned>& indices,
         bool reverse_orientation = false)
        : shape(reverse_orientation, indices.size() / 3), positions(positions), tangents(tangents), normals(normals),
          uvs(uvs), indices(indices) {}
    // shapes/mesh.cpp:225-306: an axis-aligned box centred at the origin (24 vertices,
    // outward normals) and a quad in the z = 0 plane facing +z
    static std::shared_ptr<mesh> create_box(real width, real height, real depth, bool reverse_orientation = false);
    static std::shared_ptr<mesh> create_quad(real width, real height, bool reverse_orientation = false);
    std::vector<vector3> positions, tangents, normals, uvs;
    std::vector<unsigned> indices;
};

class sphere final : public shape {
public:
    explicit sphere(real radius, bool reverse_orientation = false) : shape(reverse_orientation, 1), mRadius(radius) {}
    real radius() const noexcept { return mRadius; }
private:
    real mRadius;
};
}  // namespace shapes

namespace materials {
using textures::texture2d;
using shared::spectrums::spectrum;
class material : public interfaces::noncopyable {};

class matte_material final : public material {
public:
    matte_material(const std::shared_ptr<texture2d<spectrum>>& diffuse, const std::shared_ptr<texture2d<real>>& sigma)
        : diffuse(diffuse), sigma(sigma) {}
    std::shared_ptr<texture2d<spectrum>> diffuse;
    std::shared_ptr<texture2d<real>> sigma;
};

class metal_material final : public material {
public:
    metal_material(const std::shared_ptr<texture2d<spectrum>>& eta, const std::shared_ptr<texture2d<spectrum>>& k,
                   const std::shared_ptr<texture2d<real>>& roughness_u,
                   const std::shared_ptr<texture2d<real>>& roughness_v, bool map_roughness_to_alpha = true)
        : eta(eta), k(k), roughness_u(roughness_u), roughness_v(roughness_v), map_roughness_to_alpha(map_roughness_to_alpha) {}
    std::shared_ptr<texture2d<spectrum>> eta, k;
    std::shared_ptr<texture2d<real>> roughness_u, roughness_v;
    bool map_roughness_to_alpha;
};

// materials/plastic_material.cpp:9-28 (eta 1.49 unless given)
class plastic_material final : public material {
public:
    plastic_material(const std::shared_ptr<texture2d<spectrum>>& specular, const std::shared_ptr<texture2d<spectrum>>& diffuse,
                     const std::shared_ptr<texture2d<real>>& roughness, bool map_roughness_to_alpha = true)
        : plastic_material(specular, diffuse, roughness, std::make_shared<texture2d<real>>(1.49f),
                           map_roughness_to_alpha) {}
    plastic_material(const std::shared_ptr<texture2d<spectrum>>& specular, const std::shared_ptr<texture2d<spectrum>>& diffuse,
                     const std::shared_ptr<texture2d<real>>& roughness, const std::shared_ptr<texture2d<real>>& eta,
                     bool map_roughness_to_alpha = true)
        : specular(specular), diffuse(diffuse), roughness(roughness), eta(eta),
          map_roughness_to_alpha(map_roughness_to_alpha) {}
    std::shared_ptr<texture2d<spectrum>> specular, diffuse;
    std::shared_ptr<texture2d<real>> roughness, eta;
    bool map_roughness_to_alpha;
};

// trowbridge_reitz_distribution.cpp:128-136
real roughness_to_alpha(real roughness);
}  // namespace materials

namespace emitters {
using shared::spectrums::spectrum;
class emitter : public interfaces::noncopyable {};
class surface_light final : public emitter {
public:
    explicit surface_light(const spectrum& radiance) : radiance(radiance) {}
    spectrum radiance;
};
}  // namespace emitters

namespace media {
using shared::spectrums::spectrum;
// media/medium.hpp:51-58
class medium : public interfaces::noncopyable {
public:
    virtual ~medium() = default;
};
// media/homogeneous_medium.hpp
class homogeneous_medium final : public medium {
public:
    homogeneous_medium(const spectrum& sigma_a, const spectrum& sigma_s, real g)
        : sigma_a(sigma_a), sigma_s(sigma_s), g(g) {}
    spectrum sigma_a, sigma_s;
    real g;
};
// media/heterogeneous_medium.hpp: sigma_t an image texture over the unit cube of the
// medium's local space (transform: local to world); albedo and g constant textures
class heterogeneous_medium final : public medium {
public:
    heterogeneous_medium(const std::shared_ptr<textures::constant_texture3d<spectrum>>& albedo,
                         const std::shared_ptr<textures::image_texture3d<real>>& sigma_t,
                         const std::shared_ptr<textures::constant_texture3d<real>>& g, const shared::transform& transform)
        : albedo(albedo), sigma_t(sigma_t), g(g), local_to_world(transform) {}
    std::shared_ptr<textures::constant_texture3d<spectrum>> albedo;
    std::shared_ptr<textures::image_texture3d<real>> sigma_t;
    std::shared_ptr<textures::constant_texture3d<real>> g;
    shared::transform local_to_world;
};
// media/medium.hpp:60-75: the media component of an entity (null = vacuum)
class media final : public interfaces::noncopyable {
public:
    media(const std::shared_ptr<medium>& outside, const std::shared_ptr<medium>& inside)
        : mOutside(outside), mInside(inside) {}
    explicit media(const std::shared_ptr<medium>& two_side) : media(two_side, two_side) {}
    std::shared_ptr<medium> outside() const noexcept { return mOutside; }
    std::shared_ptr<medium> inside() const noexcept { return mInside; }
    bool different_sides() const noexcept { return mOutside != mInside; }
private:
    std::shared_ptr<medium> mOutside, mInside;
};
}  // namespace media

namespace scenes {
class entity final : public interfaces::noncopyable, public std::enable_shared_from_this<entity> {
public:
    entity(const std::shared_ptr<materials::material>& material, const std::shared_ptr<emitters::emitter>& emitter,
           const std::shared_ptr<shapes::shape>& shape, const std::shared_ptr<media::media>& media,
           const shared::transform& transform)
        : mMaterial(material), mEmitter(emitter), mShape(shape), mMedia(media), mTransform(transform) {}
    std::shared_ptr<materials::material> material() const noexcept { return mMaterial; }
    std::shared_ptr<emitters::emitter> emitter() const noexcept { return mEmitter; }
    std::shared_ptr<shapes::shape> shape() const noexcept { return mShape; }
    std::shared_ptr<media::media> media() const noexcept { return mMedia; }
    shared::transform transform() const noexcept { return mTransform; }
private:
    std::shared_ptr<materials::material> mMaterial;
    std::shared_ptr<emitters::emitter> mEmitter;
    std::shared_ptr<shapes::shape> mShape;
    std::shared_ptr<media::media> mMedia;
    shared::transform mTransform;
};

class scene final : public interfaces::noncopyable {
public:
    void add_entity(const std::shared_ptr<entity>& e) {
        if (e->emitter()) mEmitters.push_back(e);
        mEntities.push_back(e);
        mGeneration++;
    }
    void build_accelerator() { mGeneration++; }
    const std::vector<std::shared_ptr<entity>>& entities() const noexcept { return mEntities; }
    const std::vector<std::shared_ptr<entity>>& emitters() const noexcept { return mEmitters; }
    uint64_t generation() const noexcept { return mGeneration; }
private:
    std::vector<std::shared_ptr<entity>> mEntities, mEmitters;
    uint64_t mGeneration = 0;
};
}  // namespace scenes

namespace filters {
class filter : public interfaces::noncopyable {
public:
    explicit filter(const vector2& radius) : mRadius(radius) {}
    vector2 radius() const noexcept { return mRadius; }
private:
    vector2 mRadius;
};
class box_filter final : public filter {
public:
    explicit box_filter(const vector2& radius = vector2{0.5f, 0.5f}) : filter(radius) {}
};
}  // namespace filters

namespace cameras {
using shared::spectrums::spectrum;
// cameras/film.hpp:52-91 (pixel values as the render leaves them)
class film final : public interfaces::noncopyable {
public:
    film(const std::shared_ptr<filters::filter>& filter, const vector2i& resolution, const bound2& crop_window,
         real scale = 1);
    void set_pixel(size_t index, const spectrum& value) { mPixels[index] = value; }
    spectrum pixel(size_t index) const { return mPixels[index]; }
    // cameras/film.cpp:114-153: the 8-bit image (scale, sRGB gamma, crop bound) as file_name + ".png"
    void write(const std::string& file_name) const;
    vector2i resolution() const noexcept { return mResolution; }
    bound2i pixels_bound() const noexcept { return mPixelsBound; }
    std::shared_ptr<filters::filter> filter() const noexcept { return mFilter; }
private:
    std::shared_ptr<filters::filter> mFilter;
    vector2i mResolution;
    bound2i mPixelsBound;
    std::vector<spectrum> mPixels;
    real mScale = 1;
};

enum class camera_system : uint32_t { right_hand = 0, left_hand = 1 };

class camera : public interfaces::noncopyable {
public:
    camera(const std::shared_ptr<cameras::film>& film, const camera_system& system, const shared::transform& transform)
        : mFilm(film), mCameraSystem(system), mCameraToWorld(transform) {}
    std::shared_ptr<cameras::film> film() const noexcept { return mFilm; }
    const shared::transform& camera_to_world() const noexcept { return mCameraToWorld; }
protected:
    std::shared_ptr<cameras::film> mFilm;
    camera_system mCameraSystem;
    shared::transform mCameraToWorld;
};

// cameras/camera.cpp:21-52 + perspective_camera.cpp:1-20
class perspective_camera final : public camera {
public:
    perspective_camera(const std::shared_ptr<cameras::film>& film, const camera_system& system,
                       const bound2& screen_window, const shared::transform& projective,
                       const shared::transform& transform, real focus = 1e6, real lens = 0);
    const shared::transform& raster_to_camera() const noexcept { return mRasterToCamera; }
    real focus() const noexcept { return mFocus; }
    real lens() const noexcept { return mLens; }
private:
    shared::transform mCameraToScreen, mScreenToRaster, mRasterToCamera;
    real mFocus, mLens;
};
}  // namespace cameras

namespace samplers {
// samplers/samplers.hpp:13-45: what the GPU tracer needs of a sampler is its samples
// per pixel and, for the stratified sampler, its strata and stratified dimensions
template <size_t Dimension>
class sampler_t : public interfaces::noncopyable {
public:
    explicit sampler_t(size_t samples_per_pixel) : mSamplesPerPixel(samples_per_pixel) {}
    virtual ~sampler_t() = default;
    size_t samples_per_pixel() const noexcept { return mSamplesPerPixel; }
    virtual bool stratified() const noexcept { return false; }
    virtual size_t strata_x() const noexcept { return mSamplesPerPixel; }
    virtual size_t strata_y() const noexcept { return 1; }
    virtual size_t dimension() const noexcept { return 0; }
protected:
    size_t mSamplesPerPixel;
};

// samplers/random_sampler.hpp
template <size_t Dimension>
class random_sampler_t final : public sampler_t<Dimension> {
public:
    explicit random_sampler_t(size_t samples_per_pixel) : sampler_t<Dimension>(samples_per_pixel) {}
};

// samplers/stratified_sampler.hpp:8-38 (constructor arguments as the reference's)
template <size_t Dimension>
class stratified_sampler_t final : public sampler_t<Dimension> {
public:
    stratified_sampler_t(size_t samples_per_pixel_x, size_t samples_per_pixel_y, size_t dimension)
        : sampler_t<Dimension>(samples_per_pixel_x * samples_per_pixel_y), mX(samples_per_pixel_x),
          mY(samples_per_pixel_y), mDimension(dimension) {}
    bool stratified() const noexcept override { return true; }
    size_t strata_x() const noexcept override { return mX; }
    size_t strata_y() const noexcept override { return mY; }
    size_t dimension() const noexcept override { return mDimension; }
private:
    size_t mX, mY, mDimension;
};

using sampler1d = sampler_t<1>;
using sampler2d = sampler_t<2>;
using random_sampler1d = random_sampler_t<1>;
using random_sampler2d = random_sampler_t<2>;
using stratified_sampler1d = stratified_sampler_t<1>;
using stratified_sampler2d = stratified_sampler_t<2>;
}  // namespace samplers

namespace integrators {
using cameras::camera;
using scenes::scene;

// integrators/integrator.hpp:58-71
class integrator : public interfaces::noncopyable {
public:
    virtual void render(const std::shared_ptr<camera>& camera, const std::shared_ptr<scene>& scene) = 0;
};

// Flat copy of a rainbow scene in the C ABI's form, owning its arrays.
struct flat_scene {
    std::vector<float> positions, normals, uvs;  // uvs: 2 per vertex (zero where a mesh has none)
    std::vector<uint32_t> indices, triangle_entity, triangle_material;
    std::vector<float> spheres;
    std::vector<uint32_t> sphere_entity;
    std::vector<rbw_entity> entities;
    std::vector<rbw_material> materials;
    std::vector<rbw_medium> media;        // empty: no entity has media
    std::vector<uint32_t> entity_media;   // (outside, inside) per entity when media is not empty
    std::vector<float> medium_grid;
    bool has_normals = false, has_uvs = false;
    rbw_scene_desc desc() const;
};

// Flattens entities in scene order (scenes/scene.cpp:15-28).  Meshes with a
// non-identity transform are baked to world space; spheres need a translation.
flat_scene flatten(const scene& s);
rbw_camera_desc flatten(const cameras::perspective_camera& c);

// The drop-in: path_integrator's constructor arguments (path_integrator.hpp:9-12),
// rendered by the MI355X wavefront tracer.
class wavefront_path_integrator final : public integrator {
public:
    explicit wavefront_path_integrator(const std::shared_ptr<samplers::sampler2d>& sampler2d,
                                       const std::shared_ptr<samplers::sampler1d>& sampler1d, size_t max_depth = 5,
                                       real threshold = 1, int device = 0);
    ~wavefront_path_integrator() override;
    void render(const std::shared_ptr<camera>& camera, const std::shared_ptr<scene>& scene) override;

    uint32_t seed = 0;
    uint32_t triangle_test = RBW_TRI_MOLLER_TRUMBORE;
    uint32_t shadow_mode = RBW_SHADOW_CLOSEST_VISIBLE;
    // RBW_INTEGRATOR_DIRECT makes this wavefront_direct_integrator's engine
    uint32_t integrator_kind = RBW_INTEGRATOR_PATH;
    uint32_t emitter_samples = 1, bsdf_samples = 1;
    uint32_t medium_mis = 0;  // volume path: 1 = phase-sampled emitter hits count (see rbw_render_params)
    rbw_stats last_stats{};
private:
    std::shared_ptr<samplers::sampler2d> mSampler2D;
    std::shared_ptr<samplers::sampler1d> mSampler1D;
    size_t mMaxDepth;
    real mThreshold;
    int mDevice;
    rbw_ctx* mCtx = nullptr;
    const scene* mUploaded = nullptr;
    uint64_t mUploadedGeneration = 0;
};

// integrators/direct_integrator.hpp: direct lighting with emitter_samples light samples and
// bsdf_samples BSDF samples per camera sample, on the same wavefront engine
class wavefront_direct_integrator final : public integrator {
public:
    wavefront_direct_integrator(const std::shared_ptr<samplers::sampler2d>& sampler2d,
                                const std::shared_ptr<samplers::sampler1d>& sampler1d, size_t emitter_samples = 1,
                                size_t bsdf_samples = 1, int device = 0)
        : mEngine(sampler2d, sampler1d, 1, 1, device) {
        mEngine.integrator_kind = RBW_INTEGRATOR_DIRECT;
        mEngine.emitter_samples = uint32_t(emitter_samples);
        mEngine.bsdf_samples = uint32_t(bsdf_samples);
    }
    void render(const std::shared_ptr<camera>& camera, const std::shared_ptr<scene>& scene) override {
        mEngine.render(camera, scene);
    }
    wavefront_path_integrator& engine() noexcept { return mEngine; }  // seed, options, last_stats
private:
    wavefront_path_integrator mEngine;
};

// integrators/volume_path_integrator.hpp: the path integrator through the scene's media
// (entities' media components), same constructor arguments as path_integrator
class wavefront_volume_path_integrator final : public integrator {
public:
    wavefront_volume_path_integrator(const std::shared_ptr<samplers::sampler2d>& sampler2d,
                                     const std::shared_ptr<samplers::sampler1d>& sampler1d, size_t max_depth = 5,
                                     real threshold = 1, int device = 0)
        : mEngine(sampler2d, sampler1d, max_depth, threshold, device) {
        mEngine.integrator_kind = RBW_INTEGRATOR_VOLUME_PATH;
    }
    void render(const std::shared_ptr<camera>& camera, const std::shared_ptr<scene>& scene) override {
        mEngine.render(camera, scene);
    }
    wavefront_path_integrator& engine() noexcept { return mEngine; }  // seed, medium_mis, last_stats
private:
    wavefront_path_integrator mEngine;
};
}  // namespace integrators

namespace scenes {
// A scene description file (JSON): materials, entities, camera and render settings.
//   {"materials": {"white": {"matte": {"diffuse": [0.73, 0.73, 0.73], "sigma": 0}},
//                  "red": {"plastic": {"diffuse": [...], "specular": [...], "roughness": 0.1, "eta": 1.49}},
//                  "gold": {"metal": {"eta": [...], "k": [...], "roughness": 0.1, "remap_roughness": false}}},
//    "media": {"fog": {"homogeneous": {"sigma_a": [...], "sigma_s": [...], "g": 0.5}},
//              "smoke": {"grid": {"size": [nx, ny, nz], "density": [...], "albedo": [...], "g": 0,
//                                 "min": [x, y, z], "max": [x, y, z]}}},
//    "entities": [{"shape": {"box": [w, h, d]} | {"quad": [w, h]} | {"sphere": r} |
//                           {"mesh": {"positions": [...], "indices": [...], "normals": [...]}},
//                  "material": "white" | {...}, "emitter": {"radiance": [r, g, b]},
//                  "media": "fog" | {"outside": null, "inside": "fog"},  (no material: invisible boundary)
//                  "transform": [{"translate": [x, y, z]}, {"scale": s | [x, y, z]},
//                                {"rotate": {"axis": [x, y, z], "degrees": a}}],
//                  "reverse_orientation": false}, ...],
//    "camera": {"resolution": [w, h], "fov": 40, "eye": [...], "target": [...], "up": [0, 1, 0],
//               "lens_radius": 0, "focus": 1e6, "crop": [x0, y0, x1, y1], "scale": 1},
//    "render": {"spp": 16, "max_depth": 5, "threshold": 1, "seed": 0,
//               "integrator": "path" | "direct" | "volume_path", "medium_mis": false,
//               "emitter_samples": 1, "bsdf_samples": 1,
//               "stratified": {"strata": [4, 4], "dims_1d": 8, "dims_2d": 12}, "output": "name"}}
// Transforms compose left to right; the screen window is [-1, 1]^2 (square pixels: the projection
// carries the aspect), "fov" the vertical field of view in degrees.
struct scene_file {
    std::shared_ptr<scenes::scene> scene;
    std::shared_ptr<cameras::camera> camera;
    size_t spp = 16, max_depth = 5;
    real threshold = 1;
    uint32_t seed = 0;
    bool stratified = false;
    size_t strata_x = 1, strata_y = 1, dims_1d = 0, dims_2d = 0;
    bool direct = false;  // "integrator": "direct"
    bool volume = false;  // "integrator": "volume_path"
    bool medium_mis = false;
    size_t emitter_samples = 1, bsdf_samples = 1;
    std::string output;  // film::write name ("" = none)
    // the integrator the render settings describe (path, or direct with its sample counts)
    std::shared_ptr<integrators::integrator> integrator() const;
};
scene_file load_scene(const std::string& json_text);
scene_file load_scene_file(const std::string& path);
}  // namespace scenes

}  // namespace rainbow::cpus
