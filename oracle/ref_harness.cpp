// Test infrastructure only: a driver around the reference's own headers
// (compiled from /root/reference/src by oracle/Makefile into oracle/_ref/).
// It emits golden vectors: the reference's parsed scene arrays, rendered
// float images, Scene::hit results, thin-film reflectances, PCG32 draws,
// randomUnitVector outputs and PPM files from its Writer.  Nothing in the product links or runs it.
//
// The reference keeps some of what is tested here (thinFilmReflectance,
// Renderer fields) private; the driver opens access with a macro after every
// standard header is in, which changes no layout and no arithmetic.
#include <atomic>
#include <cassert>
#include <cctype>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>
#include <omp.h>

#define private public
#include "core/constants.h"
#include "core/parser.h"
#include "core/utils.h"
#include "core/writer.h"
#include "render/renderer.h"
#include "scene/scene.h"
#undef private

namespace {

struct Loaded {
    Renderer renderer;
    std::vector<DenseSpectrum<Float>> spectra;
    std::vector<DenseSpectrum<Complex>> complexSpectra;
    Background background;
    std::vector<Object> objects;
    std::vector<Instance> instances;
    std::vector<BVHNode> nodes;
    std::vector<int> lightInstances, lightObjects;
    std::vector<Float> lightPowers;
    Float totalLightPower = 0;
    std::vector<Material> materials;
    std::vector<int> materialProperties;
    std::vector<ScalarTexture> scalarTextures;
    std::vector<SpectrumTexture> spectrumTextures;

    void load(const std::string &path) {
        PERMUTATION = const_cast<int *>(&h_PERMUTATION[0]);
        CIE_XYZ_1931 = const_cast<double *>(&h_CIE_XYZ_1931[0][0]);
        CIE_D65 = const_cast<double *>(&h_CIE_D65[0]);
        Parser::parseLRD(path, renderer, spectra, complexSpectra, background, objects, instances, nodes,
                         lightInstances, lightObjects, lightPowers, totalLightPower, materials,
                         materialProperties, scalarTextures, spectrumTextures);
    }

    Scene scene() {
        return Scene(spectra.data(), complexSpectra.data(), background, objects.data(), instances.data(),
                     nodes.data(), lightInstances.data(), lightObjects.data(), int(lightInstances.size()),
                     lightPowers.data(), totalLightPower, materials.data(), materialProperties.data(),
                     scalarTextures.data(), spectrumTextures.data());
    }
};

// Tagged binary records: 16-byte name, u64 element size, u64 count, data.
void put(std::ofstream &f, const char *name, const void *data, uint64_t elem, uint64_t count) {
    char tag[16] = {0};
    std::strncpy(tag, name, 15);
    f.write(tag, 16);
    f.write(reinterpret_cast<const char *>(&elem), 8);
    f.write(reinterpret_cast<const char *>(&count), 8);
    if (count) f.write(reinterpret_cast<const char *>(data), std::streamsize(elem * count));
}

std::vector<char> slurp(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + path);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

int cmd_scene(const std::string &lrd, const std::string &out) {
    Loaded L;
    L.load(lrd);
    std::ofstream f(out, std::ios::binary);
    const Renderer &r = L.renderer;
    int ints[6] = {r.width, r.height, r.depth, r.samples, r.sqrtSamples, 0};
    float floats[2] = {r.lambdaMin, r.lambdaMax};
    put(f, "render_i", ints, 4, 6);
    put(f, "render_f", floats, 4, 2);
    put(f, "camera", &r.camera, sizeof(Camera), 1);
    put(f, "spectra", L.spectra.data(), sizeof(DenseSpectrum<Float>), L.spectra.size());
    put(f, "cspectra", L.complexSpectra.data(), sizeof(DenseSpectrum<Complex>), L.complexSpectra.size());
    put(f, "background", &L.background, sizeof(Background), 1);
    put(f, "objects", L.objects.data(), sizeof(Object), L.objects.size());
    put(f, "instances", L.instances.data(), sizeof(Instance), L.instances.size());
    put(f, "nodes", L.nodes.data(), sizeof(BVHNode), L.nodes.size());
    put(f, "light_inst", L.lightInstances.data(), 4, L.lightInstances.size());
    put(f, "light_obj", L.lightObjects.data(), 4, L.lightObjects.size());
    put(f, "light_pow", L.lightPowers.data(), 4, L.lightPowers.size());
    put(f, "light_total", &L.totalLightPower, 4, 1);
    put(f, "materials", L.materials.data(), sizeof(Material), L.materials.size());
    put(f, "mat_props", L.materialProperties.data(), 4, L.materialProperties.size());
    put(f, "scalar_tex", L.scalarTextures.data(), sizeof(ScalarTexture), L.scalarTextures.size());
    put(f, "spectrum_tex", L.spectrumTextures.data(), sizeof(SpectrumTexture), L.spectrumTextures.size());
    return 0;
}

// Renders like src/main.cpp:105-156 (CPU flavour) and writes the float buffer.
int cmd_render(const std::string &lrd, const std::string &out, int threads) {
    Loaded L;
    auto t0 = std::chrono::steady_clock::now();
    L.load(lrd);
    auto t1 = std::chrono::steady_clock::now();
    if (threads > 0) omp_set_num_threads(threads);
    std::vector<Float> buffer(size_t(L.renderer.getTotalPixels()) * 3);
    L.renderer.setScene(L.spectra.data(), L.complexSpectra.data(), L.background, L.objects.data(),
                        L.instances.data(), L.nodes.data(), L.lightInstances.data(), L.lightObjects.data(),
                        int(L.lightInstances.size()), L.lightPowers.data(), L.totalLightPower,
                        L.materials.data(), L.materialProperties.data(), L.scalarTextures.data(),
                        L.spectrumTextures.data());
    L.renderer.setBuffer(buffer.data());
    int completed = 0;
    auto t2 = std::chrono::steady_clock::now();
    L.renderer.renderImage(&completed, 42ULL);
    auto t3 = std::chrono::steady_clock::now();
    std::ofstream f(out, std::ios::binary);
    f.write(reinterpret_cast<const char *>(buffer.data()), std::streamsize(buffer.size() * sizeof(Float)));
    double parse_s = std::chrono::duration<double>(t1 - t0).count();
    double render_s = std::chrono::duration<double>(t3 - t2).count();
    std::printf("{\"parse_s\": %.6f, \"render_s\": %.6f, \"threads\": %d, \"paths\": %lld}\n", parse_s, render_s,
                omp_get_max_threads(),
                (long long)L.renderer.getTotalPixels() * (long long)L.renderer.samples);
    return 0;
}

// rays: float[n][6]; occluders: int[n][2] (or absent for closest hit).
// Output per ray: int hit, instance, object; float t, px, py, pz, nx, ny, nz, u, v; int front.
int cmd_hits(const std::string &lrd, const std::string &rays_path, const std::string &occ_path,
             const std::string &out) {
    Loaded L;
    L.load(lrd);
    Scene s = L.scene();
    std::vector<char> rb = slurp(rays_path);
    size_t n = rb.size() / (6 * sizeof(float));
    const float *rays = reinterpret_cast<const float *>(rb.data());
    std::vector<char> ob;
    const int *occ = nullptr;
    if (occ_path != "-") {
        ob = slurp(occ_path);
        occ = reinterpret_cast<const int *>(ob.data());
    }
    std::ofstream f(out, std::ios::binary);
    SampledSpectrum lambdas(550);
    for (size_t i = 0; i < n; i++) {
        const float *q = rays + 6 * i;
        Ray r(Vector(q[0], q[1], q[2]), Vector(q[3], q[4], q[5]), lambdas);
        Intersection isect;
        bool h = occ ? s.hit(r, isect, occ[2 * i], occ[2 * i + 1]) : s.hit(r, isect);
        int32_t rec_i[3] = {h ? 1 : 0, isect.instance, isect.object};
        float rec_f[9] = {isect.t, isect.point[0], isect.point[1], isect.point[2], isect.normal[0],
                          isect.normal[1], isect.normal[2], isect.u, isect.v};
        int32_t front = isect.frontFacing ? 1 : 0;
        f.write(reinterpret_cast<const char *>(rec_i), 12);
        f.write(reinterpret_cast<const char *>(rec_f), 36);
        f.write(reinterpret_cast<const char *>(&front), 4);
    }
    return 0;
}

// queries: float[n][7] = {cos_theta, lambda, front, px, py, pz, pad}; out: float[n].
int cmd_thinfilm(const std::string &lrd, int material, const std::string &q_path, const std::string &out) {
    Loaded L;
    L.load(lrd);
    std::vector<char> qb = slurp(q_path);
    size_t n = qb.size() / (7 * sizeof(float));
    const float *q = reinterpret_cast<const float *>(qb.data());
    const Material &m = L.materials[size_t(material)];
    std::ofstream f(out, std::ios::binary);
    SampledSpectrum lambdas(550);
    for (size_t i = 0; i < n; i++) {
        const float *x = q + 7 * i;
        float c = x[0];
        float s = std::sqrt(std::max(0.0f, 1 - c * c));
        Intersection isect;
        isect.normal = Vector(0, 0, 1);
        isect.point = Vector(x[3], x[4], x[5]);
        isect.frontFacing = x[2] != 0;
        Ray r(Vector(0, 0, 0), Vector(s, 0, -c), lambdas);
        Float R = m.thinFilmReflectance(L.complexSpectra.data(), L.materialProperties.data(),
                                        L.scalarTextures.data(), isect, r, x[1]);
        f.write(reinterpret_cast<const char *>(&R), 4);
    }
    return 0;
}

int cmd_pcg(uint64_t seed, uint64_t stream, int n) {
    Random r(seed, stream);
    for (int i = 0; i < n; i++) std::printf("%u\n", r.step());
    return 0;
}

int cmd_unitvec(uint64_t seed, uint64_t stream, int n) {
    Random r(seed, stream);
    for (int i = 0; i < n; i++) {
        Vector v = randomUnitVector(r);
        std::printf("%a %a %a\n", double(v[0]), double(v[1]), double(v[2]));
    }
    return 0;
}

// Writes a float RGB buffer (raw float[H][W][3]) as a PPM with the
// reference's Writer (core/writer.h:17-41: tone map, Random(0, 0) dither).
int cmd_ppm(const std::string &in, int width, int height, const std::string &out) {
    std::vector<char> b = slurp(in);
    if (b.size() != size_t(width) * size_t(height) * 3 * sizeof(Float)) throw std::runtime_error("ppm: size mismatch");
    std::vector<Float> buffer(size_t(width) * size_t(height) * 3);
    std::memcpy(buffer.data(), b.data(), b.size());
    Writer::write(out, width, height, buffer.data());
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        std::string c = argc > 1 ? argv[1] : "";
        if (c == "scene" && argc == 4) return cmd_scene(argv[2], argv[3]);
        if (c == "render" && argc >= 4) return cmd_render(argv[2], argv[3], argc > 4 ? std::atoi(argv[4]) : 0);
        if (c == "hits" && argc == 6) return cmd_hits(argv[2], argv[3], argv[4], argv[5]);
        if (c == "thinfilm" && argc == 6) return cmd_thinfilm(argv[2], std::atoi(argv[3]), argv[4], argv[5]);
        if (c == "pcg" && argc == 5) return cmd_pcg(std::strtoull(argv[2], nullptr, 10), std::strtoull(argv[3], nullptr, 10), std::atoi(argv[4]));
        if (c == "unitvec" && argc == 5) return cmd_unitvec(std::strtoull(argv[2], nullptr, 10), std::strtoull(argv[3], nullptr, 10), std::atoi(argv[4]));
        if (c == "ppm" && argc == 6) return cmd_ppm(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), argv[5]);
        std::fprintf(stderr, "usage: ref_harness scene|render|hits|thinfilm|pcg|unitvec|ppm ...\n");
        return 2;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
}
