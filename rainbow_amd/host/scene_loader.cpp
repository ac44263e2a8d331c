// Scene description loader for the rainbow-style host interface: a JSON file
// names materials, entities (boxes and quads with the reference's
// mesh::create_box / create_quad conventions, spheres, triangle meshes),
// transforms, emitters, the camera and the render settings, and is turned into
// the mirror's scene and camera objects (rainbow has no loader of its own; its
// scenes are built in code with these constructors: shapes/mesh.cpp:225-306,
// scenes/scene.cpp:15-28).
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

#include "rainbow_host.hpp"

namespace rainbow::cpus {

namespace {

using shared::spectrums::spectrum;

// ------------------------------------------------------------------ JSON

struct json {
    enum kind_t { null_v, bool_v, number_v, string_v, array_v, object_v } kind = null_v;
    bool b = false;
    double n = 0;
    std::string s;
    std::vector<json> a;
    std::vector<std::pair<std::string, json>> o;

    const json* find(const std::string& key) const {
        for (const auto& kv : o)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

class json_parser {
public:
    explicit json_parser(const std::string& text) : t(text) {}
    json parse() {
        json v = value();
        ws();
        if (p != t.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& t;
    size_t p = 0;

    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error("scene file: " + what + " at offset " + std::to_string(p));
    }
    void ws() {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r')) p++;
    }
    bool eat(char c) {
        ws();
        if (p < t.size() && t[p] == c) {
            p++;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail(std::string("expected '") + c + "'");
    }
    json value() {
        ws();
        if (p >= t.size()) fail("unexpected end");
        const char c = t[p];
        json v;
        if (c == '{') {
            p++;
            v.kind = json::object_v;
            if (eat('}')) return v;
            do {
                ws();
                json key = value();
                if (key.kind != json::string_v) fail("object keys must be strings");
                expect(':');
                v.o.emplace_back(key.s, value());
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            p++;
            v.kind = json::array_v;
            if (eat(']')) return v;
            do v.a.push_back(value());
            while (eat(','));
            expect(']');
        } else if (c == '"') {
            p++;
            v.kind = json::string_v;
            while (p < t.size() && t[p] != '"') {
                if (t[p] == '\\') {
                    p++;
                    if (p >= t.size()) fail("unterminated string");
                    const char e = t[p];
                    v.s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
                } else {
                    v.s += t[p];
                }
                p++;
            }
            if (p >= t.size()) fail("unterminated string");
            p++;
        } else if (t.compare(p, 4, "true") == 0) {
            p += 4;
            v.kind = json::bool_v;
            v.b = true;
        } else if (t.compare(p, 5, "false") == 0) {
            p += 5;
            v.kind = json::bool_v;
        } else if (t.compare(p, 4, "null") == 0) {
            p += 4;
        } else {
            const char* begin = t.c_str() + p;
            char* end = nullptr;
            v.n = std::strtod(begin, &end);
            if (end == begin) fail("unexpected character");
            p += size_t(end - begin);
            v.kind = json::number_v;
        }
        return v;
    }
};

// ------------------------------------------------------------------ field access

[[noreturn]] void bad(const std::string& where, const std::string& what) {
    throw std::runtime_error("scene file: " + where + ": " + what);
}

double number(const json& v, const std::string& where) {
    if (v.kind != json::number_v) bad(where, "expected a number");
    return v.n;
}

double number_or(const json& obj, const char* key, double def, const std::string& where) {
    const json* v = obj.find(key);
    return v ? number(*v, where + "." + key) : def;
}

std::vector<double> numbers(const json& v, size_t count, const std::string& where) {
    if (v.kind != json::array_v || (count && v.a.size() != count))
        bad(where, count ? "expected an array of " + std::to_string(count) + " numbers" : "expected an array");
    std::vector<double> r;
    for (size_t i = 0; i < v.a.size(); i++) r.push_back(number(v.a[i], where + "[" + std::to_string(i) + "]"));
    return r;
}

vector3 vec3(const json& v, const std::string& where) {
    const auto n = numbers(v, 3, where);
    return vector3{real(n[0]), real(n[1]), real(n[2])};
}

spectrum rgb(const json& v, const std::string& where) {
    if (v.kind == json::number_v) return spectrum(real(v.n));
    const auto n = numbers(v, 3, where);
    return spectrum(real(n[0]), real(n[1]), real(n[2]));
}

const json& member(const json& obj, const char* key, const std::string& where) {
    if (obj.kind != json::object_v) bad(where, "expected an object");
    const json* v = obj.find(key);
    if (!v) bad(where, std::string("missing \"") + key + "\"");
    return *v;
}

template <typename T>
std::shared_ptr<textures::constant_texture2d<T>> constant(const T& v) {
    return std::make_shared<textures::constant_texture2d<T>>(v);
}

std::shared_ptr<materials::material> make_material(const json& m, const std::string& where) {
    if (m.kind != json::object_v || m.o.size() != 1)
        bad(where, "a material is {\"matte\": {...}}, {\"metal\": {...}} or {\"plastic\": {...}}");
    const std::string& type = m.o[0].first;
    const json& a = m.o[0].second;
    if (type == "matte")
        return std::make_shared<materials::matte_material>(constant(rgb(member(a, "diffuse", where), where + ".diffuse")),
                                                           constant(real(number_or(a, "sigma", 0.0, where))));
    if (type == "metal") {
        real ru = 0.1f, rv = 0.1f;
        if (const json* r = a.find("roughness")) {
            if (r->kind == json::number_v) {
                ru = rv = real(r->n);
            } else {
                const auto n = numbers(*r, 2, where + ".roughness");
                ru = real(n[0]);
                rv = real(n[1]);
            }
        }
        const json* remap = a.find("remap_roughness");
        return std::make_shared<materials::metal_material>(constant(rgb(member(a, "eta", where), where + ".eta")),
                                                           constant(rgb(member(a, "k", where), where + ".k")),
                                                           constant(ru), constant(rv),
                                                           remap ? (remap->kind == json::bool_v && remap->b) : true);
    }
    if (type == "plastic") {
        const json* remap = a.find("remap_roughness");
        return std::make_shared<materials::plastic_material>(
            constant(rgb(member(a, "specular", where), where + ".specular")),
            constant(rgb(member(a, "diffuse", where), where + ".diffuse")), constant(real(number_or(a, "roughness", 0.1, where))),
            constant(real(number_or(a, "eta", 1.49, where))), remap ? (remap->kind == json::bool_v && remap->b) : true);
    }
    bad(where, "unknown material type \"" + type + "\"");
}

// transforms compose left to right: [A, B] places a shape by A * B
shared::transform make_transform(const json* list, const std::string& where) {
    shared::transform t;
    if (!list) return t;
    if (list->kind != json::array_v) bad(where, "expected an array of transforms");
    for (size_t i = 0; i < list->a.size(); i++) {
        const json& op = list->a[i];
        const std::string w = where + "[" + std::to_string(i) + "]";
        if (op.kind != json::object_v || op.o.size() != 1) bad(w, "a transform is {\"translate\"|\"scale\"|\"rotate\": ...}");
        const std::string& type = op.o[0].first;
        const json& arg = op.o[0].second;
        if (type == "translate") {
            t = t * shared::translate(vec3(arg, w));
        } else if (type == "scale") {
            t = t * shared::scale(arg.kind == json::number_v ? vector3{real(arg.n), real(arg.n), real(arg.n)} : vec3(arg, w));
        } else if (type == "rotate") {
            const double deg = number(member(arg, "degrees", w), w + ".degrees");
            t = t * shared::rotate(real(deg * 3.14159265358979323846 / 180.0), vec3(member(arg, "axis", w), w + ".axis"));
        } else {
            bad(w, "unknown transform \"" + type + "\"");
        }
    }
    return t;
}

std::shared_ptr<shapes::shape> make_shape(const json& s, bool reverse, const std::string& where) {
    if (s.kind != json::object_v || s.o.size() != 1) bad(where, "a shape is {\"box\"|\"quad\"|\"sphere\"|\"mesh\": ...}");
    const std::string& type = s.o[0].first;
    const json& a = s.o[0].second;
    if (type == "box") {
        const auto n = numbers(a, 3, where + ".box");
        return shapes::mesh::create_box(real(n[0]), real(n[1]), real(n[2]), reverse);
    }
    if (type == "quad") {
        const auto n = numbers(a, 2, where + ".quad");
        return shapes::mesh::create_quad(real(n[0]), real(n[1]), reverse);
    }
    if (type == "sphere") return std::make_shared<shapes::sphere>(real(number(a, where + ".sphere")), reverse);
    if (type == "mesh") {
        const auto P = numbers(member(a, "positions", where), 0, where + ".positions");
        const auto I = numbers(member(a, "indices", where), 0, where + ".indices");
        if (P.empty() || P.size() % 3 || I.empty() || I.size() % 3) bad(where, "positions and indices come in threes");
        std::vector<vector3> pos, nrm;
        for (size_t i = 0; i < P.size(); i += 3) pos.push_back(vector3{real(P[i]), real(P[i + 1]), real(P[i + 2])});
        if (const json* N = a.find("normals")) {
            const auto nn = numbers(*N, P.size(), where + ".normals");
            for (size_t i = 0; i < nn.size(); i += 3) nrm.push_back(vector3{real(nn[i]), real(nn[i + 1]), real(nn[i + 2])});
        }
        std::vector<unsigned> idx;
        for (double x : I) {
            if (x < 0 || x >= double(pos.size()) || x != std::floor(x)) bad(where, "index out of range");
            idx.push_back(unsigned(x));
        }
        return std::make_shared<shapes::mesh>(pos, std::vector<vector3>{}, nrm, std::vector<vector3>{}, idx, reverse);
    }
    bad(where, "unknown shape \"" + type + "\"");
}

// a named medium: {"homogeneous": {"sigma_a": rgb, "sigma_s": rgb, "g": g}} or
// {"grid": {"size": [nx, ny, nz], "density": [...], "albedo": rgb, "g": g, "min": xyz, "max": xyz}}
// (the grid spans the box [min, max]; density x fastest)
std::shared_ptr<media::medium> make_medium(const json& m, const std::string& where) {
    if (m.kind != json::object_v || m.o.size() != 1)
        bad(where, "a medium is {\"homogeneous\": {...}} or {\"grid\": {...}}");
    const std::string& type = m.o.begin()->first;
    const json& a = m.o.begin()->second;
    const real g = real(number_or(a, "g", 0.0, where));
    if (!(std::fabs(g) < 1)) bad(where + ".g", "must lie in (-1, 1)");
    if (type == "homogeneous")
        return std::make_shared<media::homogeneous_medium>(rgb(member(a, "sigma_a", where), where + ".sigma_a"),
                                                           rgb(member(a, "sigma_s", where), where + ".sigma_s"), g);
    if (type == "grid") {
        const auto n = numbers(member(a, "size", where), 3, where + ".size");
        if (n[0] < 1 || n[1] < 1 || n[2] < 1) bad(where + ".size", "must be positive");
        const std::array<size_t, 3> size{size_t(n[0]), size_t(n[1]), size_t(n[2])};
        const json& d = member(a, "density", where);
        const auto dens = numbers(d, size[0] * size[1] * size[2], where + ".density");
        std::vector<real> vals(dens.begin(), dens.end());
        const vector3 lo = vec3(member(a, "min", where), where + ".min");
        const vector3 hi = vec3(member(a, "max", where), where + ".max");
        if (!(hi.x > lo.x && hi.y > lo.y && hi.z > lo.z)) bad(where, "max must exceed min");
        const shared::transform t = shared::translate(lo) * shared::scale({hi.x - lo.x, hi.y - lo.y, hi.z - lo.z});
        return std::make_shared<media::heterogeneous_medium>(
            std::make_shared<textures::constant_texture3d<spectrum>>(rgb(member(a, "albedo", where), where + ".albedo")),
            std::make_shared<textures::image_texture3d<real>>(size, std::move(vals)),
            std::make_shared<textures::constant_texture3d<real>>(g), t);
    }
    bad(where, "unknown medium type \"" + type + "\"");
    return nullptr;
}

}  // namespace

namespace scenes {

scene_file load_scene(const std::string& text) {
    const json root = json_parser(text).parse();
    if (root.kind != json::object_v) bad("root", "expected an object");
    scene_file out;
    out.scene = std::make_shared<scenes::scene>();
    std::map<std::string, std::shared_ptr<materials::material>> mats;
    if (const json* m = root.find("materials")) {
        if (m->kind != json::object_v) bad("materials", "expected an object of named materials");
        for (const auto& kv : m->o) mats[kv.first] = make_material(kv.second, "materials." + kv.first);
    }
    std::map<std::string, std::shared_ptr<media::medium>> meds;
    if (const json* m = root.find("media")) {
        if (m->kind != json::object_v) bad("media", "expected an object of named media");
        for (const auto& kv : m->o) meds[kv.first] = make_medium(kv.second, "media." + kv.first);
    }
    auto medium_ref = [&](const json* v, const std::string& w) -> std::shared_ptr<media::medium> {
        if (!v || v->kind == json::null_v) return nullptr;
        if (v->kind != json::string_v) bad(w, "expected the name of a medium or null");
        auto it = meds.find(v->s);
        if (it == meds.end()) bad(w, "unknown medium \"" + v->s + "\"");
        return it->second;
    };
    const json& ents = member(root, "entities", "root");
    if (ents.kind != json::array_v) bad("entities", "expected an array");
    for (size_t i = 0; i < ents.a.size(); i++) {
        const json& e = ents.a[i];
        const std::string w = "entities[" + std::to_string(i) + "]";
        const json* rev = e.find("reverse_orientation");
        const bool reverse = rev && rev->kind == json::bool_v && rev->b;
        auto shape = make_shape(member(e, "shape", w), reverse, w + ".shape");
        std::shared_ptr<materials::material> mat;
        if (const json* mn = e.find("material")) {
            if (mn->kind == json::string_v) {
                auto it = mats.find(mn->s);
                if (it == mats.end()) bad(w, "unknown material \"" + mn->s + "\"");
                mat = it->second;
            } else {
                mat = make_material(*mn, w + ".material");
            }
        }
        std::shared_ptr<emitters::emitter> em;
        if (const json* el = e.find("emitter"))
            em = std::make_shared<emitters::surface_light>(rgb(member(*el, "radiance", w + ".emitter"), w + ".emitter"));
        // "media": "name" (both sides) or {"outside": name | null, "inside": name | null}
        std::shared_ptr<media::media> md;
        if (const json* mm = e.find("media")) {
            if (mm->kind == json::string_v || mm->kind == json::null_v) {
                md = std::make_shared<media::media>(medium_ref(mm, w + ".media"));
            } else if (mm->kind == json::object_v) {
                md = std::make_shared<media::media>(medium_ref(mm->find("outside"), w + ".media.outside"),
                                                    medium_ref(mm->find("inside"), w + ".media.inside"));
            } else {
                bad(w + ".media", "expected a medium name or {\"outside\", \"inside\"}");
            }
        }
        if (!mat && !em && !md) bad(w, "an entity needs a material, an emitter or media");
        if (!mat && em)  // an emitter without a material reflects nothing
            mat = std::make_shared<materials::matte_material>(constant(spectrum(0)), constant(real(0)));
        // (a media boundary without material is invisible: rays pass through it)
        out.scene->add_entity(std::make_shared<entity>(mat, em, shape, md, make_transform(e.find("transform"), w + ".transform")));
    }
    out.scene->build_accelerator();

    const json& c = member(root, "camera", "root");
    const auto res = numbers(member(c, "resolution", "camera"), 2, "camera.resolution");
    if (res[0] < 1 || res[1] < 1) bad("camera.resolution", "must be positive");
    bound2 crop{{0, 0}, {1, 1}};
    if (const json* cw = c.find("crop")) {
        const auto n = numbers(*cw, 4, "camera.crop");
        crop = bound2{{real(n[0]), real(n[1])}, {real(n[2]), real(n[3])}};
    }
    auto film = std::make_shared<cameras::film>(std::make_shared<filters::box_filter>(),
                                                vector2i{int32_t(res[0]), int32_t(res[1])}, crop,
                                                real(number_or(c, "scale", 1.0, "camera")));
    const real fov = real(number_or(c, "fov", 40.0, "camera") * 3.14159265358979323846 / 180.0);
    const vector3 eye = vec3(member(c, "eye", "camera"), "camera.eye");
    const vector3 target = vec3(member(c, "target", "camera"), "camera.target");
    const json* up = c.find("up");
    const vector3 upv = up ? vec3(*up, "camera.up") : vector3{0, 1, 0};
    // perspective_right_hand already scales x by height / width: the [-1, 1]^2 window gives square pixels
    const bound2 screen{{-1, -1}, {1, 1}};
    out.camera = std::make_shared<cameras::perspective_camera>(
        film, cameras::camera_system::right_hand, screen, shared::perspective_right_hand(fov, real(res[0]), real(res[1])),
        shared::look_at_right_hand(eye, target, upv).inverse(), real(number_or(c, "focus", 1e6, "camera")),
        real(number_or(c, "lens_radius", 0.0, "camera")));

    if (const json* r = root.find("render")) {
        out.spp = size_t(number_or(*r, "spp", 16, "render"));
        out.max_depth = size_t(number_or(*r, "max_depth", 5, "render"));
        out.threshold = real(number_or(*r, "threshold", 1.0, "render"));
        out.seed = uint32_t(number_or(*r, "seed", 0, "render"));
        if (const json* st = r->find("stratified")) {
            const auto n = numbers(member(*st, "strata", "render.stratified"), 2, "render.stratified.strata");
            out.stratified = true;
            out.strata_x = size_t(n[0]);
            out.strata_y = size_t(n[1]);
            out.dims_1d = size_t(number_or(*st, "dims_1d", 0, "render.stratified"));
            out.dims_2d = size_t(number_or(*st, "dims_2d", 0, "render.stratified"));
            out.spp = out.strata_x * out.strata_y;
        }
        if (const json* it = r->find("integrator")) {
            if (it->kind != json::string_v || (it->s != "path" && it->s != "direct" && it->s != "volume_path"))
                bad("render.integrator", "expected \"path\", \"direct\" or \"volume_path\"");
            out.direct = it->s == "direct";
            out.volume = it->s == "volume_path";
        }
        if (const json* mm = r->find("medium_mis")) {
            if (mm->kind != json::bool_v) bad("render.medium_mis", "expected true or false");
            out.medium_mis = mm->b;
        }
        out.emitter_samples = size_t(number_or(*r, "emitter_samples", 1, "render"));
        out.bsdf_samples = size_t(number_or(*r, "bsdf_samples", 1, "render"));
        if (out.emitter_samples == 0 || out.bsdf_samples == 0) bad("render", "sample counts must be positive");
        if (const json* o = r->find("output")) {
            if (o->kind != json::string_v) bad("render.output", "expected a file name");
            out.output = o->s;
        }
    }
    if (out.spp == 0 || out.max_depth == 0) bad("render", "spp and max_depth must be positive");
    return out;
}

scene_file load_scene_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("scene file: cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return load_scene(ss.str());
}

std::shared_ptr<integrators::integrator> scene_file::integrator() const {
    std::shared_ptr<samplers::sampler2d> s2 = std::make_shared<samplers::random_sampler2d>(spp);
    std::shared_ptr<samplers::sampler1d> s1 = std::make_shared<samplers::random_sampler1d>(spp);
    if (stratified) {
        s2 = std::make_shared<samplers::stratified_sampler2d>(strata_x, strata_y, dims_2d);
        s1 = std::make_shared<samplers::stratified_sampler1d>(strata_x, strata_y, dims_1d);
    }
    if (direct) {
        auto integ = std::make_shared<integrators::wavefront_direct_integrator>(s2, s1, emitter_samples, bsdf_samples);
        integ->engine().seed = seed;
        return integ;
    }
    if (volume) {
        auto integ = std::make_shared<integrators::wavefront_volume_path_integrator>(s2, s1, max_depth, threshold);
        integ->engine().seed = seed;
        integ->engine().medium_mis = medium_mis ? 1u : 0u;
        return integ;
    }
    auto integ = std::make_shared<integrators::wavefront_path_integrator>(s2, s1, max_depth, threshold);
    integ->seed = seed;
    return integ;
}

}  // namespace scenes
}  // namespace rainbow::cpus
