// Host scene model: what the .lrd parser produces and the BVH builder
// reorders.  Records are plain structs; lambda_scene_view (lambda_hip.h) is
// filled from them by flattening.
//
// Geometry helpers (bounds, centroid, area, radius) follow scene/object.h:166-199
// and scene/instance.h:26-111 operation for operation, because their float
// results decide the BVH layout and the light powers.
#pragma once

#include <array>
#include <complex>
#include <string>
#include <vector>

#include "host_math.h"
#include "lambda_hip.h"

namespace lambda_host {

using Spectrum = std::array<float, kBins>;
using ComplexSpectrum = std::array<std::complex<double>, kBins>;

struct ScalarTexture {
    int type = LAMBDA_SCALAR_CONSTANT;
    float p[3] = {0, 0, 0};  // constant {value}; perlin/worley {min, max, frequency}
    float min() const { return type == LAMBDA_SCALAR_CONSTANT ? p[0] : p[0]; }
    float max() const { return type == LAMBDA_SCALAR_CONSTANT ? p[0] : p[1]; }
    float average() const { return type == LAMBDA_SCALAR_CONSTANT ? p[0] : (p[0] + p[1]) / 2; }
};

struct SpectrumTexture {
    int type = LAMBDA_SPECTRUM_CONSTANT;
    int a = 0, b = 0;     // constant {spectrum}; checker {spectrum1, spectrum2}; scalar {texture}
    float scale = 0;      // checker cell size in (u, v) units
};

struct Material {
    int type = LAMBDA_MATERIAL_LAMBERTIAN;
    int p[3] = {0, 0, 0};
    bool emissive() const { return type == LAMBDA_MATERIAL_EMISSIVE; }
};

struct Object {
    int type = LAMBDA_OBJECT_SPHERE;
    int material = 0;
    Vec3 a, h, v, n;     // sphere: a = centre; quad/tri: corner, horizontal, vertical, normal
    float radius = 0;
    Vec3 n0, n1, n2;     // triangle vertex normals
    float u0 = 0, u1 = 0, u2 = 0, v0 = 0, v1 = 0, v2 = 0;

    static Object sphere(int m, const Vec3 &c, float r);
    static Object quad(int m, const Vec3 &c, const Vec3 &h, const Vec3 &v);
    static Object tri(int m, const Vec3 &c, const Vec3 &h, const Vec3 &v, const Vec3 &n0 = Vec3(),
                      const Vec3 &n1 = Vec3(), const Vec3 &n2 = Vec3(), float u0 = 0, float u1 = 0,
                      float u2 = 0, float v0 = 0, float v1 = 0, float v2 = 0);

    Vec3 bmin() const;
    Vec3 bmax() const;
    Vec3 center() const;
    float bradius() const;
    float area() const;
};

struct Instance {
    int object = 0, count = 0, node = 0, material = -1;
    Mat4 M, Minv;

    Instance(int o, int c, int m, const Vec3 &translation, const Vec3 &rotation, const Vec3 &scale);

    int material_of(const std::vector<Object> &objects, int i) const {
        return material >= 0 ? material : objects[object + i].material;
    }
    Vec3 to_world(const Vec3 &p) const { return transform(M, p, 1); }
    Vec3 bmin(const std::vector<Object> &objects) const;
    Vec3 bmax(const std::vector<Object> &objects) const;
    void bounds(const std::vector<Object> &objects, Vec3 &lo, Vec3 &hi) const;
    Vec3 center(const std::vector<Object> &objects) const;
    Vec3 center(const std::vector<Object> &objects, int i) const { return to_world(objects[object + i].center()); }
    float radius(const std::vector<Object> &objects, int i) const;
    float area(const std::vector<Object> &objects, int i) const;
};

struct Node {
    Vec3 min, max;
    int index = -1, count = 0;
};

struct Camera {
    Vec3 position, corner, horizontal, vertical;
};

struct Scene {
    int width = 0, height = 0, depth = 0, samples = 0, sqrt_samples = 0;
    float lambda_min = 0, lambda_max = 0;
    Camera camera;

    std::vector<Spectrum> spectra;
    std::vector<ComplexSpectrum> complex_spectra;
    int background_texture = -1;
    std::vector<Object> objects;
    std::vector<Instance> instances;
    std::vector<Node> nodes;
    std::vector<int> light_instances, light_objects;
    std::vector<float> light_powers;
    float total_light_power = 0;
    std::vector<Material> materials;
    std::vector<int> material_properties;
    std::vector<ScalarTexture> scalar_textures;
    std::vector<SpectrumTexture> spectrum_textures;
};

// Average of a dense spectrum: float sum over the 471 bins, divided by 471
// (math/spectrum.h:203-210).
float spectrum_average(const Spectrum &s);
float spectrum_min(const Spectrum &s);
float spectrum_max(const Spectrum &s);

float spectrum_texture_min(const Scene &s, int t);
float spectrum_texture_max(const Scene &s, int t);
float spectrum_texture_average(const Scene &s, int t);

// Two-level binned-SAH BVH (scene/bvh.h:59-246).  Reorders s.instances and
// s.objects, fills s.nodes and each instance's BLAS root.
void build_bvh(Scene &s, const std::vector<int> &object_offsets);

// Scene bounding sphere, emissive light list and environment light
// (core/parser.h:335-378).
void build_lights(Scene &s);

// Parser entry points (core/parser.h:29).  Throw std::runtime_error with the
// reference's "<file>:<line>: <message>" texts.
void parse_lrd_file(const std::string &path, Scene &s);
void parse_lrd_text(const std::string &text, const std::string &name, Scene &s);

}  // namespace lambda_host
