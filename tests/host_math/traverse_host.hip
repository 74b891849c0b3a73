// Test driver: runs the integrator's BVH traversal (dscene.h traverse, over
// the 4-wide BVH of scene_prep.h) on the CPU.  Built with -DRING=<n> to
// choose the LDS ring size, so that a tiny ring exercises the global spill
// path on every deep traversal.
//
// argv: <scene.lrd> closest <rays.f32 [n][6]> <out.i32 [n][4] = hit, instance, object, bits(t)>
//       <scene.lrd> occluded <rays.f32 [n][6]> <lights.i32 [n][2]> <out.i32 [n]>
// Prints the scene's stack bound on stdout.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "dscene.h"
#include "lambda_host.h"
#include "scene_prep.h"

using namespace lam;

#ifndef RING
#define RING 16
#endif

template <typename T>
static std::vector<T> read_file(const char *path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    size_t n = size_t(f.tellg());
    f.seekg(0);
    std::vector<T> v(n / sizeof(T));
    f.read(reinterpret_cast<char *>(v.data()), std::streamsize(n));
    return v;
}

template <typename T>
static void write_file(const char *path, const std::vector<T> &v) {
    std::ofstream o(path, std::ios::binary);
    o.write(reinterpret_cast<const char *>(v.data()), std::streamsize(v.size() * sizeof(T)));
}

int main(int argc, char **argv) {
    if (argc < 5) return 2;
    char err[512];
    lambda_scene *scene = nullptr;
    if (lambda_scene_load(argv[1], &scene, err, sizeof err)) {
        std::fprintf(stderr, "%s\n", err);
        return 1;
    }
    lambda_scene_view v;
    lambda_scene_get_view(scene, &v);

    WideBuilder wb(v);
    const int need = wb.run();
    std::vector<float4> obj = pack_objects(v), obj_s = pack_shading(v), inst;
    std::vector<int4> inst_i;
    pack_instances(v, inst, inst_i);
    std::vector<int> tlas_rank, obj_rank;
    reference_ranks(v, tlas_rank, obj_rank);

    DeviceScene S{};
    S.wide = wb.nodes.data();
    S.inst_root = wb.inst_root.data();
    S.obj = obj.data();
    S.obj_s = obj_s.data();
    S.inst = inst.data();
    S.inst_i = inst_i.data();
    S.tlas_rank = tlas_rank.data();
    S.obj_rank = obj_rank.data();
    S.num_objects = v.num_objects;
    S.num_instances = v.num_instances;
    std::printf("%d\n", need);

    std::vector<float> rays = read_file<float>(argv[3]);
    const size_t n = rays.size() / 6;
    int lcode[RING];
    float ldist[RING];
    std::vector<int2> spill(size_t(std::max(need, 1)));
    auto make = [&]() {
        TraversalStack<RING> s;
        s.lcode = (decltype(s.lcode))lcode;
        s.ldist = (decltype(s.ldist))ldist;
        s.stride = 1;
        s.spill = spill.data();
        s.spill_stride = 1;
        return s;
    };
    const std::string mode = argv[2];
    if (mode == "closest") {
        std::vector<int> out(4 * n);
        for (size_t i = 0; i < n; i++) {
            const float *r = &rays[6 * i];
            V3 o = v3(r[0], r[1], r[2]), d = v3(r[3], r[4], r[5]);
            float best = kMax;
            int hi = -1, ho = -1;
            TraversalStack<RING> stk = make();
            bool hit = traverse<false, RING>(S, o, d, best, hi, ho, -1, -2, stk);
            out[4 * i] = hit ? 1 : 0;
            out[4 * i + 1] = hi;
            out[4 * i + 2] = ho;
            std::memcpy(&out[4 * i + 3], &best, 4);
        }
        write_file(argv[4], out);
    } else if (mode == "occluded" && argc >= 6) {
        std::vector<int> lights = read_file<int>(argv[4]);
        std::vector<int> out(n);
        for (size_t i = 0; i < n; i++) {
            const float *r = &rays[6 * i];
            V3 o = v3(r[0], r[1], r[2]), d = v3(r[3], r[4], r[5]);
            float best = kMax;
            const int li = lights[2 * i], lo = lights[2 * i + 1];
            if (li != -1) {
                Hit lh;
                if (!hit_record(S, li, lo, o, d, lh)) {
                    out[i] = 1;
                    continue;
                }
                best = lh.t;
            }
            int hi = -1, ho = -1;
            TraversalStack<RING> stk = make();
            out[i] = traverse<true, RING>(S, o, d, best, hi, ho, li, lo, stk) ? 1 : 0;
        }
        write_file(argv[5], out);
    } else {
        return 2;
    }
    lambda_scene_free(scene);
    return 0;
}
