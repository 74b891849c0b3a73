// C ABI of the host front end (lambda_host.h): parse, flatten, write PPM.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "lambda_host.h"
#include "scene_model.h"

using namespace lambda_host;

struct lambda_scene {
    Scene s;
    std::vector<float> spectra;
    std::vector<double> complex_spectra;
    std::vector<int32_t> st_type, sp_type, sp_index, mat_type, mat_params, props;
    std::vector<float> st_params, sp_scale;
    std::vector<int32_t> obj_type, obj_mat, inst_ints, node_ints, light_inst, light_obj;
    std::vector<float> obj_geom, obj_shade, inst_mat, node_bounds, light_pow;
    lambda_scene_view view{};
};

namespace {

void set_err(char *err, size_t len, const std::string &msg) {
    if (!err || len == 0) return;
    size_t n = msg.size() < len - 1 ? msg.size() : len - 1;
    std::memcpy(err, msg.data(), n);
    err[n] = 0;
}

void push3(std::vector<float> &v, const Vec3 &a) {
    v.push_back(a.x);
    v.push_back(a.y);
    v.push_back(a.z);
}

void flatten(lambda_scene &o) {
    const Scene &s = o.s;
    for (const Spectrum &sp : s.spectra) o.spectra.insert(o.spectra.end(), sp.begin(), sp.end());
    for (const ComplexSpectrum &cs : s.complex_spectra)
        for (const auto &c : cs) {
            o.complex_spectra.push_back(c.real());
            o.complex_spectra.push_back(c.imag());
        }
    for (const ScalarTexture &t : s.scalar_textures) {
        o.st_type.push_back(t.type);
        o.st_params.insert(o.st_params.end(), t.p, t.p + 3);
    }
    for (const SpectrumTexture &t : s.spectrum_textures) {
        o.sp_type.push_back(t.type);
        o.sp_index.push_back(t.a);
        o.sp_index.push_back(t.b);
        o.sp_scale.push_back(t.scale);
    }
    for (const Material &m : s.materials) {
        o.mat_type.push_back(m.type);
        o.mat_params.insert(o.mat_params.end(), m.p, m.p + 3);
    }
    o.props = std::vector<int32_t>(s.material_properties.begin(), s.material_properties.end());
    for (const Object &ob : s.objects) {
        o.obj_type.push_back(ob.type);
        o.obj_mat.push_back(ob.material);
        if (ob.type == LAMBDA_OBJECT_SPHERE) {
            push3(o.obj_geom, ob.a);
            o.obj_geom.push_back(ob.radius);
            for (int i = 0; i < 8; i++) o.obj_geom.push_back(0);
        } else {
            push3(o.obj_geom, ob.a);
            push3(o.obj_geom, ob.h);
            push3(o.obj_geom, ob.v);
            push3(o.obj_geom, ob.n);
        }
        if (ob.type == LAMBDA_OBJECT_TRI) {
            push3(o.obj_shade, ob.n0);
            push3(o.obj_shade, ob.n1);
            push3(o.obj_shade, ob.n2);
            for (float x : {ob.u0, ob.u1, ob.u2, ob.v0, ob.v1, ob.v2}) o.obj_shade.push_back(x);
        } else {
            for (int i = 0; i < 15; i++) o.obj_shade.push_back(0);
        }
    }
    for (const Instance &in : s.instances) {
        for (int x : {in.object, in.count, in.node, in.material}) o.inst_ints.push_back(x);
        o.inst_mat.insert(o.inst_mat.end(), in.M.m, in.M.m + 16);
        o.inst_mat.insert(o.inst_mat.end(), in.Minv.m, in.Minv.m + 16);
    }
    for (const Node &n : s.nodes) {
        push3(o.node_bounds, n.min);
        push3(o.node_bounds, n.max);
        o.node_ints.push_back(n.index);
        o.node_ints.push_back(n.count);
    }
    o.light_inst = s.light_instances;
    o.light_obj = s.light_objects;
    o.light_pow = s.light_powers;

    lambda_scene_view &v = o.view;
    v.num_spectra = int32_t(s.spectra.size());
    v.spectra = o.spectra.data();
    v.num_complex_spectra = int32_t(s.complex_spectra.size());
    v.complex_spectra = o.complex_spectra.data();
    v.num_scalar_textures = int32_t(s.scalar_textures.size());
    v.scalar_texture_type = o.st_type.data();
    v.scalar_texture_params = o.st_params.data();
    v.num_spectrum_textures = int32_t(s.spectrum_textures.size());
    v.spectrum_texture_type = o.sp_type.data();
    v.spectrum_texture_index = o.sp_index.data();
    v.spectrum_texture_scale = o.sp_scale.data();
    v.num_materials = int32_t(s.materials.size());
    v.material_type = o.mat_type.data();
    v.material_params = o.mat_params.data();
    v.num_material_properties = int32_t(o.props.size());
    v.material_properties = o.props.data();
    v.num_objects = int32_t(s.objects.size());
    v.object_type = o.obj_type.data();
    v.object_material = o.obj_mat.data();
    v.object_geometry = o.obj_geom.data();
    v.object_shading = o.obj_shade.data();
    v.num_instances = int32_t(s.instances.size());
    v.instance_ints = o.inst_ints.data();
    v.instance_matrices = o.inst_mat.data();
    v.num_nodes = int32_t(s.nodes.size());
    v.node_bounds = o.node_bounds.data();
    v.node_ints = o.node_ints.data();
    v.num_lights = int32_t(s.light_instances.size());
    v.light_instances = o.light_inst.data();
    v.light_objects = o.light_obj.data();
    v.light_powers = o.light_pow.data();
    v.total_light_power = s.total_light_power;
    v.background_texture = s.background_texture;
    const Vec3 cam[4] = {s.camera.position, s.camera.corner, s.camera.horizontal, s.camera.vertical};
    for (int i = 0; i < 4; i++) {
        v.camera[3 * i + 0] = cam[i].x;
        v.camera[3 * i + 1] = cam[i].y;
        v.camera[3 * i + 2] = cam[i].z;
    }
}

template <typename F>
int load_common(F &&parse, lambda_scene **out, char *err, size_t err_len) {
    if (!out) {
        set_err(err, err_len, "lambda_scene_load: null output pointer");
        return LAMBDA_EINVAL;
    }
    *out = nullptr;
    lambda_scene *sc = new (std::nothrow) lambda_scene;
    if (!sc) {
        set_err(err, err_len, "lambda_scene_load: out of memory");
        return LAMBDA_ENOMEM;
    }
    try {
        parse(sc->s);
        flatten(*sc);
    } catch (const std::bad_alloc &) {
        delete sc;
        set_err(err, err_len, "lambda_scene_load: out of memory");
        return LAMBDA_ENOMEM;
    } catch (const std::exception &e) {
        delete sc;
        set_err(err, err_len, e.what());
        return LAMBDA_EPARSE;
    }
    *out = sc;
    return LAMBDA_OK;
}

// PCG32 (O'Neill's pcg32_srandom_r / pcg32_random_r), as math/random.h:9-31.
struct Pcg32 {
    uint64_t state, inc;
    Pcg32(uint64_t seed, uint64_t stream) : state(0), inc((stream << 1u) | 1u) {
        step();
        state += seed;
        step();
    }
    uint32_t step() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31));
    }
    float next() { return float(step()) / float(4294967296.0); }
};

float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// sRGB transfer curve on [0,1]-clamped linear values, then 8-bit with a
// uniform dither in [-0.5, 0.5) (core/utils.h:233-244, core/writer.h:23-41).
int quantize(float linear, Pcg32 &rng) {
    float c = clampf(linear, 0, 1);
    float s = c <= 0.0031308 ? 12.92f * c : 1.055f * std::pow(c, 1 / 2.4f) - 0.055f;
    float dither = rng.next() - 0.5f;
    return int(uint8_t(clampf(float(int(std::round(s * 255 + dither))), 0, 255)));
}

}  // namespace

extern "C" {

int lambda_scene_load(const char *path, lambda_scene **out, char *err, size_t err_len) {
    if (!path) {
        set_err(err, err_len, "lambda_scene_load: null path");
        return LAMBDA_EINVAL;
    }
    std::string p(path);
    return load_common([&](Scene &s) { parse_lrd_file(p, s); }, out, err, err_len);
}

int lambda_scene_parse_string(const char *text, const char *name, lambda_scene **out, char *err,
                              size_t err_len) {
    if (!text) {
        set_err(err, err_len, "lambda_scene_parse_string: null text");
        return LAMBDA_EINVAL;
    }
    std::string t(text), n(name ? name : "<string>");
    return load_common([&](Scene &s) { parse_lrd_text(t, n, s); }, out, err, err_len);
}

int lambda_scene_get_view(const lambda_scene *scene, lambda_scene_view *view) {
    if (!scene || !view) return LAMBDA_EINVAL;
    *view = scene->view;
    return LAMBDA_OK;
}

int lambda_scene_get_params(const lambda_scene *scene, lambda_render_params *p) {
    if (!scene || !p) return LAMBDA_EINVAL;
    std::memset(p, 0, sizeof(*p));
    p->width = scene->s.width;
    p->height = scene->s.height;
    p->depth = scene->s.depth;
    p->samples = scene->s.samples;
    p->sqrt_samples = scene->s.sqrt_samples;
    p->lambda_min = scene->s.lambda_min;
    p->lambda_max = scene->s.lambda_max;
    p->seed = 42;
    p->rng_mode = LAMBDA_RNG_REFERENCE;
    p->path_slots = 0;
    p->tail_paths = 0;
    p->cost_order = 0;
    p->graph_passes = 0;
    return LAMBDA_OK;
}

void lambda_scene_free(lambda_scene *scene) { delete scene; }

int lambda_write_ppm(const char *path, int32_t width, int32_t height, const float *rgb, char *err,
                     size_t err_len) {
    if (!path || !rgb || width <= 0 || height <= 0) {
        set_err(err, err_len, "lambda_write_ppm: invalid arguments");
        return LAMBDA_EINVAL;
    }
    std::string p(path);
    if (p.size() < 4 || p.substr(p.size() - 4) != ".ppm") {
        set_err(err, err_len, p + ": invalid file extension");
        return LAMBDA_EINVAL;
    }
    std::ofstream f(p);
    if (!f.is_open()) {
        std::string e = std::strerror(errno);
        if (!e.empty()) e[0] = char(std::tolower(e[0]));
        set_err(err, err_len, p + ": failed to open file (" + e + ")");
        return LAMBDA_EIO;
    }
    Pcg32 rng(0, 0);
    f << "P3\n" << width << " " << height << "\n255\n";
    const size_t n = size_t(width) * size_t(height);
    for (size_t i = 0; i < n; i++) {
        int r = quantize(rgb[i * 3 + 0], rng);
        int g = quantize(rgb[i * 3 + 1], rng);
        int b = quantize(rgb[i * 3 + 2], rng);
        f << r << " " << g << " " << b << "\n";
    }
    if (!f) {
        set_err(err, err_len, p + ": write failed");
        return LAMBDA_EIO;
    }
    return LAMBDA_OK;
}

}  // extern "C"
