// Host-side mirror of rainbow's render interfaces and the wavefront path
// integrator that renders through the C ABI (include/rbw.h).
#include "rainbow_host.hpp"

#include <algorithm>
#include <cstring>
#include <map>

namespace rainbow::core::math {

matrix4x4 matrix4x4::inverse() const {
    // Gauss-Jordan with partial pivoting on the row-major view (double)
    double a[4][8];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            a[r][c] = m[c * 4 + r];
            a[r][c + 4] = (r == c) ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; col++) {
        int piv = col;
        for (int r = col + 1; r < 4; r++)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
        if (a[piv][col] == 0.0) throw std::runtime_error("singular matrix");
        if (piv != col)
            for (int c = 0; c < 8; c++) std::swap(a[piv][c], a[col][c]);
        const double inv = 1.0 / a[col][col];
        for (int c = 0; c < 8; c++) a[col][c] *= inv;
        for (int r = 0; r < 4; r++) {
            if (r == col) continue;
            const double f = a[r][col];
            if (f == 0.0) continue;
            for (int c = 0; c < 8; c++) a[r][c] -= f * a[col][c];
        }
    }
    matrix4x4 out;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) out.m[c * 4 + r] = a[r][c + 4];
    return out;
}

}  // namespace rainbow::core::math

namespace rainbow::cpus {

namespace shared {

bool transform::is_identity() const {
    const matrix4x4 I = matrix4x4::identity();
    for (int i = 0; i < 16; i++)
        if ((float)mTransform.m[i] != (float)I.m[i]) return false;
    return true;
}

bool transform::is_translation(vector3& t) const {
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 4; r++)
            if ((float)mTransform.m[c * 4 + r] != (c == r ? 1.0f : 0.0f)) return false;
    if ((float)mTransform.m[15] != 1.0f) return false;
    t = vector3{(float)mTransform.m[12], (float)mTransform.m[13], (float)mTransform.m[14]};
    return true;
}

transform translate(const vector3& v) {
    matrix4x4 m = matrix4x4::identity(), i = matrix4x4::identity();
    m.m[12] = v.x; m.m[13] = v.y; m.m[14] = v.z;
    i.m[12] = -v.x; i.m[13] = -v.y; i.m[14] = -v.z;
    return transform(m, i);
}

transform rotate(real angle, const vector3& axis) {
    // glm::rotate(mat4(1), angle, axis) in double, rounded; the inverse of a rotation is its transpose
    const double len = std::sqrt(double(axis.x) * axis.x + double(axis.y) * axis.y + double(axis.z) * axis.z);
    const double x = axis.x / len, y = axis.y / len, z = axis.z / len;
    const double c = std::cos(double(angle)), s = std::sin(double(angle)), t = 1.0 - c;
    const double r[3][3] = {{c + t * x * x, t * x * y + s * z, t * x * z - s * y},   // column 0
                            {t * y * x - s * z, c + t * y * y, t * y * z + s * x},   // column 1
                            {t * z * x + s * y, t * z * y - s * x, c + t * z * z}};  // column 2
    matrix4x4 m = matrix4x4::identity(), i = matrix4x4::identity();
    for (int col = 0; col < 3; col++)
        for (int row = 0; row < 3; row++) {
            m.m[col * 4 + row] = real(r[col][row]);
            i.m[row * 4 + col] = real(r[col][row]);
        }
    return transform(m, i);
}

transform scale(const vector3& v) {
    matrix4x4 m = matrix4x4::identity(), i = matrix4x4::identity();
    m.m[0] = v.x; m.m[5] = v.y; m.m[10] = v.z;
    i.m[0] = 1.0 / v.x; i.m[5] = 1.0 / v.y; i.m[10] = 1.0 / v.z;
    return transform(m, i);
}

// glm::perspectiveFovRH_NO
transform perspective_right_hand(real fov, real width, real height, real near, real far) {
    const double h = std::cos(0.5 * fov) / std::sin(0.5 * fov);
    const double w = h * height / width;
    matrix4x4 m;
    m.m[0] = w;
    m.m[5] = h;
    m.m[10] = -(double(far) + near) / (double(far) - near);
    m.m[11] = -1.0;
    m.m[14] = -(2.0 * far * near) / (double(far) - near);
    return transform(m, m.inverse());
}

// glm::lookAtRH
transform look_at_right_hand(const vector3& origin, const vector3& target, const vector3& up) {
    double f[3] = {double(target.x) - origin.x, double(target.y) - origin.y, double(target.z) - origin.z};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    const double u0[3] = {up.x, up.y, up.z};
    double s[3] = {f[1] * u0[2] - f[2] * u0[1], f[2] * u0[0] - f[0] * u0[2], f[0] * u0[1] - f[1] * u0[0]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& x : s) x /= sl;
    const double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    const double e[3] = {origin.x, origin.y, origin.z};
    matrix4x4 m = matrix4x4::identity();
    for (int k = 0; k < 3; k++) {
        m.m[k * 4 + 0] = s[k];
        m.m[k * 4 + 1] = u[k];
        m.m[k * 4 + 2] = -f[k];
    }
    m.m[12] = -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]);
    m.m[13] = -(u[0] * e[0] + u[1] * e[1] + u[2] * e[2]);
    m.m[14] = (f[0] * e[0] + f[1] * e[1] + f[2] * e[2]);
    return transform(m, m.inverse());
}

}  // namespace shared

namespace materials {
real roughness_to_alpha(real roughness) {
    roughness = std::max(roughness, static_cast<real>(1e-3));
    const real x = std::log(roughness);
    return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}
}  // namespace materials

namespace cameras {

film::film(const std::shared_ptr<filters::filter>& filter, const vector2i& resolution, const bound2& crop_window,
           real scale)
    : mFilter(filter), mResolution(resolution), mScale(scale) {
    // cameras/film.cpp:186-204
    mPixelsBound.min = vector2i{(int32_t)std::ceil(mResolution.x * crop_window.min.x),
                                (int32_t)std::ceil(mResolution.y * crop_window.min.y)};
    mPixelsBound.max = vector2i{(int32_t)std::ceil(mResolution.x * crop_window.max.x),
                                (int32_t)std::ceil(mResolution.y * crop_window.max.y)};
    mPixels.assign(size_t(resolution.x) * resolution.y, spectrum(0));
}

void film::write(const std::string& file_name) const {
    std::vector<float> rgb(mPixels.size() * 3);
    for (size_t i = 0; i < mPixels.size(); i++)
        for (int c = 0; c < 3; c++) rgb[i * 3 + c] = mPixels[i][c];
    const int32_t bmin[2] = {mPixelsBound.min.x, mPixelsBound.min.y};
    const int32_t bmax[2] = {mPixelsBound.max.x, mPixelsBound.max.y};
    if (rbw_write_png((file_name + ".png").c_str(), rgb.data(), (uint32_t)mResolution.x, (uint32_t)mResolution.y, bmin,
                      bmax, mScale) != RBW_OK)
        throw std::runtime_error(rbw_last_error());
}

perspective_camera::perspective_camera(const std::shared_ptr<cameras::film>& film, const camera_system& system,
                                       const bound2& sw, const shared::transform& projective,
                                       const shared::transform& transform, real focus, real lens)
    : camera(film, system, transform), mCameraToScreen(projective), mFocus(focus), mLens(lens) {
    // cameras/camera.cpp:30-51: ScreenToRaster = scale(res) * scale(1/w, 1/-h) * translate(-min.x, -max.y)
    const vector2i res = mFilm->resolution();
    mScreenToRaster = shared::scale(vector3{(float)res.x, (float)res.y, 1.0f});
    matrix4x4 s2 = matrix4x4::identity();
    s2.m[0] = 1.0 / (double(sw.max.x) - sw.min.x);
    s2.m[5] = 1.0 / (double(sw.min.y) - sw.max.y);
    mScreenToRaster = mScreenToRaster * shared::transform(s2, s2.inverse());
    matrix4x4 t = matrix4x4::identity();
    t.m[12] = -double(sw.min.x);
    t.m[13] = -double(sw.max.y);
    mScreenToRaster = mScreenToRaster * shared::transform(t, t.inverse());
    mRasterToCamera = shared::transform(mCameraToScreen.matrix().inverse() * mScreenToRaster.matrix().inverse(),
                                        matrix4x4::identity());
}

}  // namespace cameras

namespace integrators {

namespace {

void require(bool c, const char* msg) {
    if (!c) throw std::runtime_error(msg);
}

// glm mat4 * vec4 in float: (m0*x + m1*y) + (m2*z + m3*w)
void mat_mul_f(const float* m, const float* v, float w, float* out) {
    for (int r = 0; r < 3; r++) out[r] = (m[r] * v[0] + m[4 + r] * v[1]) + (m[8 + r] * v[2] + m[12 + r] * w);
}

void to_float(const matrix4x4& m, float* out) {
    for (int i = 0; i < 16; i++) out[i] = (float)m.m[i];
}

}  // namespace

rbw_scene_desc flat_scene::desc() const {
    rbw_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.num_vertices = uint32_t(positions.size() / 3);
    d.positions = positions.data();
    d.normals = has_normals ? normals.data() : nullptr;
    d.uvs = has_uvs ? uvs.data() : nullptr;
    d.num_triangles = uint32_t(indices.size() / 3);
    d.indices = indices.data();
    d.triangle_entity = triangle_entity.data();
    d.triangle_material = nullptr;
    d.num_spheres = uint32_t(spheres.size() / 4);
    d.spheres = spheres.empty() ? nullptr : spheres.data();
    d.sphere_entity = sphere_entity.empty() ? nullptr : sphere_entity.data();
    d.num_entities = uint32_t(entities.size());
    d.entities = entities.data();
    d.num_materials = uint32_t(materials.size());
    d.materials = materials.data();
    if (!media.empty()) {
        d.num_media = uint32_t(media.size());
        d.media = media.data();
        d.entity_media = entity_media.data();
        d.medium_grid_count = uint32_t(medium_grid.size());
        d.medium_grid = medium_grid.empty() ? nullptr : medium_grid.data();
    }
    return d;
}

namespace {
// one rbw_medium per distinct medium object: homogeneous (homogeneous_medium.cpp:7-10) or a
// grid (heterogeneous_medium.cpp:5-22) with world_to_local = the inverse of its transform
uint32_t flatten_medium(flat_scene& f, std::map<const media::medium*, uint32_t>& index,
                        const std::shared_ptr<media::medium>& m) {
    if (!m) return RBW_NONE;
    auto it = index.find(m.get());
    if (it != index.end()) return it->second;
    rbw_medium out;
    std::memset(&out, 0, sizeof(out));
    if (auto h = std::dynamic_pointer_cast<media::homogeneous_medium>(m)) {
        out.type = RBW_MEDIUM_HOMOGENEOUS;
        for (int c = 0; c < 3; c++) {
            out.sigma_a[c] = h->sigma_a[c];
            out.sigma_s[c] = h->sigma_s[c];
        }
        out.g = h->g;
    } else if (auto g = std::dynamic_pointer_cast<media::heterogeneous_medium>(m)) {
        require(g->sigma_t && g->albedo && g->g, "heterogeneous_medium needs sigma_t, albedo and g");
        const auto size = g->sigma_t->size();
        const auto& vals = g->sigma_t->values();
        require(vals.size() >= size[0] * size[1] * size[2] && size[0] * size[1] * size[2] > 0,
                "sigma_t grid is smaller than its size");
        out.type = RBW_MEDIUM_GRID;
        out.grid_offset = uint32_t(f.medium_grid.size());
        for (int k = 0; k < 3; k++) out.grid_size[k] = uint32_t(size[k]);
        f.medium_grid.insert(f.medium_grid.end(), vals.begin(), vals.begin() + size[0] * size[1] * size[2]);
        for (int c = 0; c < 3; c++) out.albedo[c] = g->albedo->value()[c];
        out.g = g->g->value();
        float Minv[16];
        to_float(g->local_to_world.inverse_matrix(), Minv);  // column-major
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) out.world_to_local[4 * r + c] = Minv[c * 4 + r];
    } else {
        throw std::runtime_error("unsupported medium");
    }
    const uint32_t id = uint32_t(f.media.size());
    f.media.push_back(out);
    index.emplace(m.get(), id);
    return id;
}
}  // namespace

flat_scene flatten(const scene& s) {
    flat_scene f;
    std::map<const materials::material*, uint32_t> mat_index;
    std::map<const media::medium*, uint32_t> med_index;
    std::vector<uint32_t> ent_media;
    for (const auto& ep : s.entities()) {
        const scenes::entity& e = *ep;
        rbw_entity out;
        std::memset(&out, 0, sizeof(out));
        out.material = RBW_NONE;
        if (auto m = e.material()) {
            auto it = mat_index.find(m.get());
            if (it == mat_index.end()) {
                rbw_material rm;
                std::memset(&rm, 0, sizeof(rm));
                if (auto mm = std::dynamic_pointer_cast<materials::matte_material>(m)) {
                    rm.type = RBW_MATERIAL_MATTE;
                    for (int c = 0; c < 3; c++) rm.diffuse[c] = mm->diffuse->value()[c];
                    rm.sigma = std::min(std::max(mm->sigma->value(), 0.0f), 90.0f);
                } else if (auto mt = std::dynamic_pointer_cast<materials::metal_material>(m)) {
                    rm.type = RBW_MATERIAL_METAL;
                    for (int c = 0; c < 3; c++) {
                        rm.eta[c] = mt->eta->value()[c];
                        rm.k[c] = mt->k->value()[c];
                    }
                    const real ru = mt->roughness_u->value(), rv = mt->roughness_v->value();
                    rm.alpha_u = mt->map_roughness_to_alpha ? materials::roughness_to_alpha(ru) : ru;
                    rm.alpha_v = mt->map_roughness_to_alpha ? materials::roughness_to_alpha(rv) : rv;
                } else if (auto pm = std::dynamic_pointer_cast<materials::plastic_material>(m)) {
                    // plastic_material.cpp:30-56: Lambert + GGX (isotropic) with dielectric Fresnel
                    rm.type = RBW_MATERIAL_PLASTIC;
                    for (int c = 0; c < 3; c++) {
                        rm.diffuse[c] = pm->diffuse->value()[c];
                        rm.specular[c] = pm->specular->value()[c];
                    }
                    const real r = pm->roughness->value();
                    rm.alpha_u = rm.alpha_v = pm->map_roughness_to_alpha ? materials::roughness_to_alpha(r) : r;
                    rm.ior = pm->eta->value();
                } else {
                    throw std::runtime_error("unsupported material");
                }
                it = mat_index.emplace(m.get(), uint32_t(f.materials.size())).first;
                f.materials.push_back(rm);
            }
            out.material = it->second;
        }
        if (auto em = e.emitter()) {
            auto sl = std::dynamic_pointer_cast<emitters::surface_light>(em);
            require(sl != nullptr, "only surface_light emitters are supported");
            out.flags |= RBW_ENTITY_EMITTER;
            for (int c = 0; c < 3; c++) out.radiance[c] = sl->radiance[c];
        }
        const uint32_t eid = uint32_t(f.entities.size());
        auto sh = e.shape();
        if (!sh) {
            out.shape = RBW_SHAPE_NONE;
        } else if (auto me = std::dynamic_pointer_cast<shapes::mesh>(sh)) {
            require(me->tangents.empty(), "mesh tangents are not supported");
            const uint32_t base = uint32_t(f.positions.size() / 3);
            const bool ident = e.transform().is_identity();
            float M[16], Minv[16];
            to_float(e.transform().matrix(), M);
            to_float(e.transform().inverse_matrix(), Minv);
            for (const vector3& p : me->positions) {
                float q[3] = {p.x, p.y, p.z}, r[3];
                if (ident) {
                    f.positions.insert(f.positions.end(), {p.x, p.y, p.z});
                } else {
                    mat_mul_f(M, q, 1.0f, r);  // transform_point (shared/transform.cpp:172-177)
                    f.positions.insert(f.positions.end(), {r[0], r[1], r[2]});
                }
            }
            if (!me->normals.empty()) {
                require(me->normals.size() == me->positions.size(), "normals must match positions");
                f.has_normals = true;
                out.flags |= RBW_ENTITY_HAS_NORMALS;
                for (const vector3& n : me->normals) {
                    if (ident) {
                        f.normals.insert(f.normals.end(), {n.x, n.y, n.z});
                    } else {
                        // transform_normal: normalize(transpose(inverse) * n)
                        float T[16];
                        for (int c = 0; c < 4; c++)
                            for (int r = 0; r < 4; r++) T[c * 4 + r] = Minv[r * 4 + c];
                        float q[3] = {n.x, n.y, n.z}, r[3];
                        mat_mul_f(T, q, 0.0f, r);
                        const float l = 1.0f / std::sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
                        f.normals.insert(f.normals.end(), {r[0] * l, r[1] * l, r[2] * l});
                    }
                }
            } else {
                f.normals.insert(f.normals.end(), me->positions.size() * 3, 0.0f);
            }
            if (!me->uvs.empty()) {
                // per-vertex (u, v) (the reference's uvs are vector3 with z = 0); they set the
                // triangle's dp/du, dp/dv and so its shading frame (mesh.cpp:328-442)
                require(me->uvs.size() == me->positions.size(), "uvs must match positions");
                f.has_uvs = true;
                out.flags |= RBW_ENTITY_HAS_UVS;
                for (const vector3& t : me->uvs) f.uvs.insert(f.uvs.end(), {t.x, t.y});
            } else {
                f.uvs.insert(f.uvs.end(), me->positions.size() * 2, 0.0f);
            }
            if (sh->reverse_orientation()) out.flags |= RBW_ENTITY_REVERSE_ORIENTATION;
            out.shape = RBW_SHAPE_MESH;
            out.first = uint32_t(f.indices.size() / 3);
            out.count = uint32_t(me->indices.size() / 3);
            for (unsigned i : me->indices) f.indices.push_back(base + i);
            f.triangle_entity.insert(f.triangle_entity.end(), out.count, eid);
        } else if (auto sp = std::dynamic_pointer_cast<shapes::sphere>(sh)) {
            vector3 t;
            require(e.transform().is_translation(t), "sphere entities need a translation transform");
            out.shape = RBW_SHAPE_SPHERE;
            out.first = uint32_t(f.spheres.size() / 4);
            out.count = 1;
            if (sh->reverse_orientation()) out.flags |= RBW_ENTITY_REVERSE_ORIENTATION;
            f.spheres.insert(f.spheres.end(), {t.x, t.y, t.z, sp->radius()});
            f.sphere_entity.push_back(eid);
        } else {
            throw std::runtime_error("unsupported shape");
        }
        uint32_t mo = RBW_NONE, mi = RBW_NONE;
        if (auto md = e.media()) {
            out.flags |= RBW_ENTITY_HAS_MEDIA;
            mo = flatten_medium(f, med_index, md->outside());
            mi = flatten_medium(f, med_index, md->inside());
        }
        ent_media.insert(ent_media.end(), {mo, mi});
        f.entities.push_back(out);
    }
    f.triangle_material.assign(f.triangle_entity.size(), RBW_NONE);
    bool any_media = false;
    for (const auto& en : f.entities) any_media = any_media || (en.flags & RBW_ENTITY_HAS_MEDIA);
    if (any_media) {
        f.entity_media = std::move(ent_media);
        if (f.media.empty()) {
            // media components that are vacuum on both sides still need a (unused) table
            rbw_medium none;
            std::memset(&none, 0, sizeof(none));
            f.media.push_back(none);
        }
    }
    return f;
}

rbw_camera_desc flatten(const cameras::perspective_camera& c) {
    rbw_camera_desc d;
    std::memset(&d, 0, sizeof(d));
    to_float(c.raster_to_camera().matrix(), d.raster_to_camera);
    to_float(c.camera_to_world().matrix(), d.camera_to_world);
    const auto film = c.film();
    d.width = uint32_t(film->resolution().x);
    d.height = uint32_t(film->resolution().y);
    const bound2i b = film->pixels_bound();
    d.bound_min[0] = b.min.x;
    d.bound_min[1] = b.min.y;
    d.bound_max[0] = b.max.x;
    d.bound_max[1] = b.max.y;
    d.lens_radius = c.lens();
    d.focus_distance = c.focus();
    return d;
}

wavefront_path_integrator::wavefront_path_integrator(const std::shared_ptr<samplers::sampler2d>& sampler2d,
                                                     const std::shared_ptr<samplers::sampler1d>& sampler1d,
                                                     size_t max_depth, real threshold, int device)
    : mSampler2D(sampler2d), mSampler1D(sampler1d), mMaxDepth(max_depth), mThreshold(threshold), mDevice(device) {}

wavefront_path_integrator::~wavefront_path_integrator() {
    if (mCtx) rbw_destroy(mCtx);
}

void wavefront_path_integrator::render(const std::shared_ptr<camera>& cam, const std::shared_ptr<scene>& sc) {
    auto pc = std::dynamic_pointer_cast<cameras::perspective_camera>(cam);
    require(pc != nullptr, "wavefront_path_integrator renders perspective cameras");
    require(sc != nullptr, "scene is null");
    auto check = [](rbw_status s) {
        if (s != RBW_OK) throw std::runtime_error(std::string("rbw: ") + rbw_last_error());
    };
    if (!mCtx) {
        rbw_device_opts o{mDevice};
        check(rbw_create(&o, &mCtx));
    }
    if (mUploaded != sc.get() || mUploadedGeneration != sc->generation()) {
        const flat_scene fs = flatten(*sc);
        const rbw_scene_desc d = fs.desc();
        check(rbw_upload_scene(mCtx, &d));
        mUploaded = sc.get();
        mUploadedGeneration = sc->generation();
    }
    const rbw_camera_desc cd = flatten(*pc);
    rbw_render_params p;
    std::memset(&p, 0, sizeof(p));
    p.spp = uint32_t(mSampler2D->samples_per_pixel());
    p.max_depth = uint32_t(mMaxDepth);
    p.rr_threshold = mThreshold;
    p.seed = seed;
    p.sampler = RBW_SAMPLER_PER_SAMPLE_PCG;
    if (mSampler2D->stratified()) {
        // the stratified samplers' contract on the GPU: the 2D sampler's strata and
        // dimensions, the 1D sampler's dimensions when it is stratified too
        p.sampler = RBW_SAMPLER_STRATIFIED;
        p.strata_x = uint32_t(mSampler2D->strata_x());
        p.strata_y = uint32_t(mSampler2D->strata_y());
        p.dims_2d = uint32_t(std::min<size_t>(mSampler2D->dimension(), 0xFFFF));
        p.dims_1d = mSampler1D && mSampler1D->stratified() ? uint32_t(std::min<size_t>(mSampler1D->dimension(), 0xFFFF))
                                                           : 0u;
    }
    p.triangle_test = triangle_test;
    p.shadow_mode = shadow_mode;
    p.integrator = integrator_kind;
    p.emitter_samples = emitter_samples;
    p.medium_mis = medium_mis;
    p.bsdf_samples = bsdf_samples;
    std::vector<float> rgb(size_t(cd.width) * cd.height * 3);
    check(rbw_render(mCtx, &cd, &p, rgb.data(), &last_stats));
    // film::set_pixel (cameras/film.cpp:202-205) for every pixel in bounds
    auto film = pc->film();
    for (int32_t y = cd.bound_min[1]; y < cd.bound_max[1]; y++)
        for (int32_t x = cd.bound_min[0]; x < cd.bound_max[0]; x++) {
            const size_t i = size_t(y) * cd.width + x;
            film->set_pixel(i, cameras::spectrum(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]));
        }
}

}  // namespace integrators
}  // namespace rainbow::cpus

namespace rainbow::cpus::shapes {

// shapes/mesh.cpp:225-278: four vertices per face (-z, +z, +y, -y, -x, +x), two triangles each
std::shared_ptr<mesh> mesh::create_box(real width, real height, real depth, bool reverse_orientation) {
    const real w = 0.5f * width, h = 0.5f * height, d = 0.5f * depth;
    // corners of each face in the reference's order, with the face normal
    const struct {
        real n[3];
        real c[4][3];
    } faces[6] = {
        {{0, 0, -1}, {{-w, -h, -d}, {-w, +h, -d}, {+w, +h, -d}, {+w, -h, -d}}},
        {{0, 0, +1}, {{-w, -h, +d}, {+w, -h, +d}, {+w, +h, +d}, {-w, +h, +d}}},
        {{0, +1, 0}, {{-w, +h, -d}, {-w, +h, +d}, {+w, +h, +d}, {+w, +h, -d}}},
        {{0, -1, 0}, {{-w, -h, -d}, {+w, -h, -d}, {+w, -h, +d}, {-w, -h, +d}}},
        {{-1, 0, 0}, {{-w, -h, +d}, {-w, +h, +d}, {-w, +h, -d}, {-w, -h, -d}}},
        {{+1, 0, 0}, {{+w, -h, -d}, {+w, +h, -d}, {+w, +h, +d}, {+w, -h, +d}}},
    };
    // the reference's per-face uv corners
    const real uv[6][4][2] = {{{0, 1}, {0, 0}, {1, 0}, {1, 1}}, {{1, 1}, {0, 1}, {0, 0}, {1, 0}},
                              {{0, 1}, {0, 0}, {1, 0}, {1, 1}}, {{1, 1}, {0, 1}, {0, 0}, {1, 0}},
                              {{0, 1}, {0, 0}, {1, 0}, {1, 1}}, {{0, 1}, {0, 0}, {1, 0}, {1, 1}}};
    std::vector<vector3> P, N, UV;
    std::vector<unsigned> I;
    for (unsigned f = 0; f < 6; f++) {
        for (int k = 0; k < 4; k++) {
            P.push_back(vector3{faces[f].c[k][0], faces[f].c[k][1], faces[f].c[k][2]});
            N.push_back(vector3{faces[f].n[0], faces[f].n[1], faces[f].n[2]});
            UV.push_back(vector3{uv[f][k][0], uv[f][k][1], 0});
        }
        const unsigned b = 4 * f;
        for (unsigned k : {b, b + 1, b + 2, b, b + 2, b + 3}) I.push_back(k);
    }
    return std::make_shared<mesh>(P, std::vector<vector3>{}, N, UV, I, reverse_orientation);
}

// shapes/mesh.cpp:280-306
std::shared_ptr<mesh> mesh::create_quad(real width, real height, bool reverse_orientation) {
    const real w = 0.5f * width, h = 0.5f * height;
    std::vector<vector3> P = {{-w, +h, 0}, {-w, -h, 0}, {+w, -h, 0}, {+w, +h, 0}};
    std::vector<vector3> N(4, vector3{0, 0, 1});
    std::vector<vector3> UV = {{0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}};
    return std::make_shared<mesh>(P, std::vector<vector3>{}, N, UV, std::vector<unsigned>{0, 1, 2, 2, 3, 0},
                                  reverse_orientation);
}

}  // namespace rainbow::cpus::shapes
