// Builds the Cornell-box workload through the rainbow-style host interfaces
// (rainbow_amd/host/rainbow_host.hpp) the way a rainbow application would,
// then either writes the flattened C-ABI descriptors ("flatten <file>") or
// renders with wavefront_path_integrator and writes the film ("render <file>").
#include <cstdio>
#include <cstring>
#include <string>

#include "../../rainbow_amd/host/rainbow_host.hpp"

using namespace rainbow::cpus;
using shared::spectrums::spectrum;

static std::shared_ptr<shapes::mesh> quad(vector3 a, vector3 b, vector3 c, vector3 d, vector3 n) {
    return std::make_shared<shapes::mesh>(std::vector<vector3>{a, b, c, d}, std::vector<vector3>{},
                                          std::vector<vector3>{n, n, n, n}, std::vector<vector3>{},
                                          std::vector<unsigned>{0, 1, 2, 0, 2, 3});
}

static std::shared_ptr<materials::matte_material> matte(float r, float g, float b) {
    return std::make_shared<materials::matte_material>(
        std::make_shared<textures::constant_texture2d<spectrum>>(spectrum(r, g, b)),
        std::make_shared<textures::constant_texture2d<real>>(0.0f));
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s flatten|render|render_stratified <out> [spp depth seed [png name]] | load|render_file <json> <out>\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    auto sc = std::make_shared<scenes::scene>();
    const shared::transform I;
    auto white = matte(0.73f, 0.73f, 0.73f), red = matte(0.65f, 0.05f, 0.05f), green = matte(0.12f, 0.45f, 0.15f),
         black = matte(0, 0, 0);
    auto add = [&](std::shared_ptr<shapes::shape> s, std::shared_ptr<materials::material> m,
                   std::shared_ptr<emitters::emitter> e = nullptr, shared::transform t = shared::transform()) {
        sc->add_entity(std::make_shared<scenes::entity>(m, e, s, nullptr, t));
    };
    add(quad({-1, -1, 1}, {1, -1, 1}, {1, -1, -1}, {-1, -1, -1}, {0, 1, 0}), white);
    add(quad({-1, 1, -1}, {1, 1, -1}, {1, 1, 1}, {-1, 1, 1}, {0, -1, 0}), white);
    add(quad({-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {0, 0, 1}), white);
    add(quad({-1, -1, 1}, {-1, -1, -1}, {-1, 1, -1}, {-1, 1, 1}, {1, 0, 0}), red);
    add(quad({1, -1, -1}, {1, -1, 1}, {1, 1, 1}, {1, 1, -1}, {-1, 0, 0}), green);
    const float h = 0.125f, y = 0.999f;
    add(quad({-h, y, -h}, {h, y, -h}, {h, y, h}, {-h, y, h}, {0, -1, 0}), black,
        std::make_shared<emitters::surface_light>(spectrum(17, 12, 4)));
    add(std::make_shared<shapes::sphere>(0.3f), matte(0.8f, 0.8f, 0.8f), nullptr,
        shared::translate({-0.45f, -0.7f, -0.35f}));
    auto gold = std::make_shared<materials::metal_material>(
        std::make_shared<textures::constant_texture2d<spectrum>>(spectrum(0.143119f, 0.374957f, 1.442479f)),
        std::make_shared<textures::constant_texture2d<spectrum>>(spectrum(3.983160f, 2.385721f, 1.603215f)),
        std::make_shared<textures::constant_texture2d<real>>(0.1f),
        std::make_shared<textures::constant_texture2d<real>>(0.1f), false);
    add(std::make_shared<shapes::sphere>(0.2f), gold, nullptr, shared::translate({0.5f, -0.8f, 0.3f}));
    sc->build_accelerator();

    auto film = std::make_shared<cameras::film>(std::make_shared<filters::box_filter>(), vector2i{256, 256},
                                                bound2{{0, 0}, {1, 1}});
    const float fov = static_cast<float>(40.0 * 3.14159265358979323846 / 180.0);
    auto cam = std::make_shared<cameras::perspective_camera>(
        film, cameras::camera_system::right_hand, bound2{{-1, -1}, {1, 1}},
        shared::perspective_right_hand(fov, 256, 256),
        shared::look_at_right_hand({0, 0, 3.5f}, {0, 0, 0}, {0, 1, 0}).inverse());

    // a scene description file instead: "load <json> <out>" flattens it, "render_file <json> <out>"
    // renders it with its own settings (and writes its "output" PNG, if it names one)
    std::shared_ptr<cameras::camera> view = cam;
    std::shared_ptr<scenes::scene> world = sc;
    scenes::scene_file file;
    const char* out_path = argv[2];
    if (mode == "load" || mode == "render_file") {
        if (argc < 4) return 2;
        try {
            file = scenes::load_scene_file(argv[2]);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 4;
        }
        view = file.camera;
        world = file.scene;
        out_path = argv[3];
    }
    FILE* out = std::fopen(out_path, "wb");
    if (!out) return 3;
    if (mode == "render_file") {
        auto integ = file.integrator();
        integ->render(view, world);
        auto f = view->film();
        const size_t n = size_t(f->resolution().x) * f->resolution().y;
        for (size_t i = 0; i < n; i++) std::fwrite(f->pixel(i).c.data(), 4, 3, out);
        if (!file.output.empty()) f->write(file.output);
        std::fclose(out);
        return 0;
    }
    if (mode == "flatten" || mode == "load") {
        auto* pc = dynamic_cast<cameras::perspective_camera*>(view.get());
        if (!pc) return 5;
        const integrators::flat_scene fs = integrators::flatten(*world);
        const rbw_camera_desc cd = integrators::flatten(*pc);
        const uint32_t counts[5] = {uint32_t(fs.positions.size() / 3), uint32_t(fs.indices.size() / 3),
                                    uint32_t(fs.spheres.size() / 4), uint32_t(fs.entities.size()),
                                    uint32_t(fs.materials.size())};
        std::fwrite(counts, 4, 5, out);
        std::fwrite(fs.positions.data(), 4, fs.positions.size(), out);
        std::fwrite(fs.normals.data(), 4, fs.normals.size(), out);
        std::fwrite(fs.indices.data(), 4, fs.indices.size(), out);
        std::fwrite(fs.triangle_entity.data(), 4, fs.triangle_entity.size(), out);
        std::fwrite(fs.spheres.data(), 4, fs.spheres.size(), out);
        std::fwrite(fs.sphere_entity.data(), 4, fs.sphere_entity.size(), out);
        std::fwrite(fs.entities.data(), sizeof(rbw_entity), fs.entities.size(), out);
        std::fwrite(fs.materials.data(), sizeof(rbw_material), fs.materials.size(), out);
        std::fwrite(&cd, sizeof(cd), 1, out);
        const uint32_t has_uvs = fs.has_uvs ? 1u : 0u;
        std::fwrite(&has_uvs, 4, 1, out);
        if (fs.has_uvs) std::fwrite(fs.uvs.data(), 4, fs.uvs.size(), out);  // (u, v) per vertex, when any
        // media: count, table, (outside, inside) per entity, grid densities
        const uint32_t nmed = uint32_t(fs.media.size()), ngrid = uint32_t(fs.medium_grid.size());
        std::fwrite(&nmed, 4, 1, out);
        if (nmed) {
            std::fwrite(fs.media.data(), sizeof(rbw_medium), nmed, out);
            std::fwrite(fs.entity_media.data(), 4, fs.entity_media.size(), out);
        }
        std::fwrite(&ngrid, 4, 1, out);
        if (ngrid) std::fwrite(fs.medium_grid.data(), 4, ngrid, out);
    } else {
        const size_t spp = argc > 3 ? std::stoul(argv[3]) : 16;
        const size_t depth = argc > 4 ? std::stoul(argv[4]) : 5;
        // mode "render_stratified": stratified samplers of 4 x (spp / 4) strata, 8 1D and 12 2D dimensions
        std::shared_ptr<samplers::sampler2d> s2 = std::make_shared<samplers::random_sampler2d>(spp);
        std::shared_ptr<samplers::sampler1d> s1 = std::make_shared<samplers::random_sampler1d>(spp);
        if (mode == "render_stratified") {
            s2 = std::make_shared<samplers::stratified_sampler2d>(4, spp / 4, 12);
            s1 = std::make_shared<samplers::stratified_sampler1d>(4, spp / 4, 8);
        }
        integrators::wavefront_path_integrator integ(s2, s1, depth, 1.0f);
        integ.seed = argc > 5 ? uint32_t(std::stoul(argv[5])) : 0;
        integ.render(cam, sc);
        for (size_t i = 0; i < 256 * 256; i++) {
            const spectrum p = film->pixel(i);
            std::fwrite(p.c.data(), 4, 3, out);
        }
        std::fprintf(stderr, "rendered %.1f Msamples/s\n", integ.last_stats.msamples_per_s);
        if (argc > 6) film->write(argv[6]);  // argv[6] + ".png"
    }
    std::fclose(out);
    return 0;
}
