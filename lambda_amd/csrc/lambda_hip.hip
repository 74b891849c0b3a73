// MI355X (gfx950) wavefront spectral path tracer behind the lambda_hip.h ABI.
//
// Each pass over the in-flight paths is three kernels:
//   trace    closest-hit traversal for the paths about to shade, and, in the
//            same launch, the occlusion rays the previous shade pass queued
//            (their contributions land in the path radiance)   (scene/scene.h:51-124)
//   finalize samples that ended in the previous shade pass: radiance -> FP64
//            XYZ framebuffer, and a new camera sample for the slot
//                                                             (render/renderer.h:64-92)
//   shade    miss / emission / next-event estimation / BSDF sampling /
//            Russian roulette                                 (render/renderer.h:102-192)
// A path's next-event contribution is added before its next emission term
// and before its sample is retired, and every slot draws its random numbers
// in the reference's order, so reference-stream renders follow Lambda's
// estimator step for step.  Queues are compacted with wave-wide ballots: one
// atomic per wave per queue.  Path state lives in structure-of-arrays float4
// records so every queue pass reads and writes 16 bytes per lane.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "lambda_hip.h"
#include "guide_net.h"
#include "integrator.h"
#include "launch.h"
#include "lbvh.h"
#include "scene_prep.h"

namespace lam {
#include "spectral_tables.inc"
}

using namespace lam;

namespace {

// ------------------------------------------------------------ host: scene

void require(bool ok, const std::string &what) {
    if (!ok) throw Failure{LAMBDA_EINVAL, "invalid scene: " + what};
}

// Deepest stack a depth-first traversal of the tree at `root` needs, with
// the push-two/pop-one discipline of the kernels; also checks the indices.
int tree_stack_depth(const lambda_scene_view &v, int root, int leaf_limit, bool instance_level) {
    struct Item {
        int node, sp;
    };
    std::vector<Item> work{{root, 1}};
    int deepest = 1;
    long long visits = 0;
    while (!work.empty()) {
        Item it = work.back();
        work.pop_back();
        require(it.node >= 0 && it.node < v.num_nodes, "node index out of range");
        require(++visits <= 4LL * v.num_nodes + 4, "node graph is not a tree");
        int idx = v.node_ints[2 * it.node], cnt = v.node_ints[2 * it.node + 1];
        if (cnt > 0) {
            require(idx >= 0 && (long long)idx + cnt <= leaf_limit, "leaf range out of range");
            (void)instance_level;
            continue;
        }
        require(cnt == 0, "negative node count");
        // popping this node leaves sp-1 entries; pushing two children
        int sp = it.sp - 1 + 2;
        deepest = std::max(deepest, sp);
        work.push_back({idx, sp});      // right child is popped first
        work.push_back({it.node + 1, sp - 1});
    }
    return deepest;
}



void validate(const lambda_scene_view &v, int max_depth = kStack) {
    require(v.num_nodes > 0 && v.num_objects > 0 && v.num_instances > 0, "empty scene");
    require(v.num_lights >= 1, "light list must end with the environment entry");
    require(v.num_spectra > 0 && v.num_spectrum_textures > 0 && v.num_scalar_textures > 0 && v.num_materials > 0,
            "missing tables");
    require(v.spectra && v.spectrum_texture_type && v.spectrum_texture_index && v.spectrum_texture_scale &&
                v.scalar_texture_type && v.scalar_texture_params && v.material_type && v.material_params &&
                v.object_type && v.object_material && v.object_geometry && v.object_shading && v.instance_ints &&
                v.instance_matrices && v.node_bounds && v.node_ints && v.light_instances && v.light_objects &&
                v.light_powers,
            "null array");
    require(v.num_complex_spectra == 0 || v.complex_spectra, "null complex spectra");
    require(v.num_material_properties == 0 || v.material_properties, "null material properties");
    for (int i = 0; i < v.num_scalar_textures; i++) {
        int t = v.scalar_texture_type[i];
        require(t >= 0 && t <= 2, "scalar texture type");
    }
    for (int i = 0; i < v.num_spectrum_textures; i++) {
        int t = v.spectrum_texture_type[i];
        int a = v.spectrum_texture_index[2 * i], b = v.spectrum_texture_index[2 * i + 1];
        require(t >= 0 && t <= 2, "spectrum texture type");
        if (t == LAMBDA_SPECTRUM_SCALAR) require(a >= 0 && a < v.num_scalar_textures, "spectrum texture index");
        else require(a >= 0 && a < v.num_spectra, "spectrum texture spectrum");
        if (t == LAMBDA_SPECTRUM_CHECKER) require(b >= 0 && b < v.num_spectra, "checker spectrum");
    }
    for (int i = 0; i < v.num_materials; i++) {
        const int *p = v.material_params + 3 * i;
        int t = v.material_type[i];
        require(t >= 0 && t <= 4, "material type");
        if (t == LAMBDA_MATERIAL_LAMBERTIAN || t == LAMBDA_MATERIAL_MIRROR || t == LAMBDA_MATERIAL_EMISSIVE)
            require(p[0] >= 0 && p[0] < v.num_spectrum_textures, "material texture");
        if (t == LAMBDA_MATERIAL_DIELECTRIC)
            require(p[0] >= 0 && p[0] < v.num_spectra && p[1] >= 0 && p[1] < v.num_spectra, "dielectric spectra");
        if (t == LAMBDA_MATERIAL_THINFILM) {
            int layers = p[0];
            require(layers >= 0 && layers <= kMaxLayers, "thin film layer count (at most 8)");
            require(p[1] >= 0 && (long long)p[1] + layers + 2 <= v.num_material_properties, "thin film n range");
            require(p[2] >= 0 && (long long)p[2] + layers <= v.num_material_properties, "thin film d range");
            for (int j = 0; j < layers + 2; j++) {
                int c = v.material_properties[p[1] + j];
                require(c >= 0 && c < v.num_complex_spectra, "thin film index spectrum");
            }
            for (int j = 0; j < layers; j++) {
                int c = v.material_properties[p[2] + j];
                require(c >= 0 && c < v.num_scalar_textures, "thin film thickness texture");
            }
        }
    }
    for (int i = 0; i < v.num_objects; i++) {
        require(v.object_type[i] >= 0 && v.object_type[i] <= 2, "object type");
        require(v.object_material[i] >= 0 && v.object_material[i] < v.num_materials, "object material");
    }
    for (int i = 0; i < v.num_instances; i++) {
        const int *q = v.instance_ints + 4 * i;
        require(q[0] >= 0 && q[1] >= 1 && (long long)q[0] + q[1] <= v.num_objects, "instance object range");
        require(q[2] >= 0 && q[2] < v.num_nodes, "instance BLAS root");
        require(q[3] >= -1 && q[3] < v.num_materials, "instance material");
    }
    for (int i = 0; i < v.num_lights; i++) {
        int li = v.light_instances[i], lo = v.light_objects[i];
        if (li == -1) {
            require(lo == -1, "environment light object");
        } else {
            require(li >= 0 && li < v.num_instances, "light instance");
            const int *q = v.instance_ints + 4 * li;
            require(lo >= q[0] && lo < q[0] + q[1], "light object outside its instance");
        }
    }
    require(v.light_instances[v.num_lights - 1] == -1, "last light must be the environment");
    require(v.background_texture >= 0 && v.background_texture < v.num_spectrum_textures, "background texture");
    require(tree_stack_depth(v, 0, v.num_instances, true) <= max_depth, "instance BVH too deep");
    for (int i = 0; i < v.num_instances; i++)
        require(tree_stack_depth(v, v.instance_ints[4 * i + 2], v.num_objects, false) <= max_depth,
                "object BVH too deep");
}

}  // namespace

struct lambda_hip_scene {
    DevBuf<float4> wide, obj, obj_s, inst, stex;
    DevBuf<int> inst_root, obj_mat, props, perm, tlas_rank, obj_rank;
    int device = 0;         // HIP device holding the scene
    int stack_need = 0;     // deepest traversal stack of the scene
    float box[6] = {0, 0, 0, 0, 0, 0};  // scene bounds (instance-level root)
    DevBuf<float4> worley_points;
    DevBuf<int4> worley_box, worley_dim;
    DevBuf<int4> inst_i, mat, xtex, mat_tex;
    DevBuf<float> spectra, light_pow, cie, tex_avg;
    DevBuf<float2> cspec;
    DevBuf<int2> light_io;
    DeviceScene ds{};
    int num_materials = 0;
    std::vector<int> material_type;
    // the last render's buffers (a Renderer), reused by the next render of
    // the same size and slot count; renders of one scene serialize here
    std::shared_ptr<void> renderer;
    int r_width = 0, r_height = 0, r_slots = 0;
    std::mutex render_mutex;
};

struct lambda_guide {
    GuideNet net;
};

namespace {

// Deepest tree the GPU-built BVHs may have (binary radix trees of 30-bit
// codes); the traversal's spill area is sized from the scene's own need.
constexpr int kLbvhMaxDepth = 96;

void upload_scene(const lambda_scene_view &v, lambda_hip_scene &g, int max_depth = kStack) {
    validate(v, max_depth);
    HIP_OK(hipGetDevice(&g.device));
    const size_t O = size_t(v.num_objects), I = size_t(v.num_instances);
    WideBuilder wb(v);
    g.stack_need = wb.run();
    for (int k = 0; k < 6; k++) g.box[k] = v.node_bounds[k];
    std::vector<float4> ob4 = pack_objects(v), os = pack_shading(v);
    std::vector<int> om(O);
    for (size_t i = 0; i < O; i++) om[i] = v.object_material[i];
    std::vector<float4> im;
    std::vector<int4> ii;
    pack_instances(v, im, ii);
    std::vector<int4> mt(size_t(v.num_materials));
    g.material_type.resize(size_t(v.num_materials));
    for (int i = 0; i < v.num_materials; i++) {
        const int *p = v.material_params + 3 * i;
        int t = v.material_type[i];
        g.material_type[i] = t;
        // thin film: {type, layers, offset of n, offset of d}
        mt[i] = make_int4(t, p[0], p[1], p[2]);
    }
    g.num_materials = v.num_materials;
    std::vector<float4> st(size_t(v.num_scalar_textures));
    for (int i = 0; i < v.num_scalar_textures; i++) {
        int t = v.scalar_texture_type[i];
        float ft;
        std::memcpy(&ft, &t, 4);
        const float *p = v.scalar_texture_params + 3 * i;
        st[i] = make_float4(ft, p[0], p[1], p[2]);
    }
    std::vector<int4> xt(size_t(v.num_spectrum_textures));
    for (int i = 0; i < v.num_spectrum_textures; i++) {
        int sc;
        std::memcpy(&sc, v.spectrum_texture_scale + i, 4);
        xt[i] = make_int4(v.spectrum_texture_type[i], v.spectrum_texture_index[2 * i], v.spectrum_texture_index[2 * i + 1], sc);
    }
    // per material, the xtex record of its texture (the materials whose
    // first parameter is a spectrum texture), so that shading reads it with
    // the material record
    std::vector<int4> mtx(size_t(v.num_materials), make_int4(0, 0, 0, 0));
    for (int i = 0; i < v.num_materials; i++) {
        const int t = v.material_type[i];
        if (t == LAMBDA_MATERIAL_LAMBERTIAN || t == LAMBDA_MATERIAL_MIRROR || t == LAMBDA_MATERIAL_EMISSIVE)
            mtx[size_t(i)] = xt[size_t(v.material_params[3 * i])];
    }
    // mean of each spectrum texture over the 471 bins, summed in float in
    // bin order as the device formula (spectrum_texture_average) does
    std::vector<float> tavg(size_t(v.num_spectrum_textures));
    {
        auto spectrum_mean = [&](int k) {
            const float *p = v.spectra + size_t(k) * kBins;
            float sum = 0;
            for (int i = 0; i < kBins; i++) sum += p[i];
            return sum / kBins;
        };
        for (int i = 0; i < v.num_spectrum_textures; i++) {
            const int4 x = xt[size_t(i)];
            if (x.x == LAMBDA_SPECTRUM_CONSTANT) {
                tavg[size_t(i)] = spectrum_mean(x.y);
            } else if (x.x == LAMBDA_SPECTRUM_CHECKER) {
                tavg[size_t(i)] = float(0.5) * (spectrum_mean(x.y) + spectrum_mean(x.z));
            } else {
                const float4 q = st[size_t(x.y)];
                int qt;
                std::memcpy(&qt, &q.x, 4);
                tavg[size_t(i)] = qt == LAMBDA_SCALAR_CONSTANT ? q.y : (q.y + q.z) / 2;
            }
        }
    }
    std::vector<float2> cs(size_t(v.num_complex_spectra) * kBins);
    for (size_t i = 0; i < cs.size(); i++)
        cs[i] = make_float2(float(v.complex_spectra[2 * i]), float(v.complex_spectra[2 * i + 1]));
    std::vector<int2> lio(size_t(v.num_lights));
    for (int i = 0; i < v.num_lights; i++) lio[i] = make_int2(v.light_instances[i], v.light_objects[i]);
    std::vector<float> cie(kBins * 3);
    for (int i = 0; i < kBins; i++)
        for (int k = 0; k < 3; k++) cie[3 * i + k] = float(kCie1931Xyz[i][k]);

    g.wide.upload(wb.nodes);
    g.inst_root.upload(wb.inst_root);
    g.obj.upload(ob4);
    g.obj_s.upload(os);
    g.obj_mat.upload(om);
    g.inst.upload(im);
    g.inst_i.upload(ii);
    g.mat.upload(mt);
    g.stex.upload(st);
    g.xtex.upload(xt);
    g.mat_tex.upload(mtx);
    g.tex_avg.upload(tavg);
    g.spectra.upload(v.spectra, size_t(v.num_spectra) * kBins);
    if (v.num_complex_spectra) g.cspec.upload(cs);
    if (v.num_material_properties) g.props.upload(v.material_properties, size_t(v.num_material_properties));
    g.light_io.upload(lio);
    g.light_pow.upload(v.light_powers, size_t(v.num_lights));
    g.cie.upload(cie);
    g.perm.upload(kPerlinPermutation, 256);
    {
        std::vector<int> tr, orank;
        reference_ranks(v, tr, orank);
        g.tlas_rank.upload(tr);
        g.obj_rank.upload(orank);
    }
    {
        // Worley feature-point tables: every cell a lookup can reach, for
        // points on the geometry (inside the scene box) and for environment
        // directions (inside the unit cube), scaled by the frequency.
        const size_t kMaxCells = size_t(1) << 23;
        std::vector<float4> pts;
        std::vector<int4> box(size_t(v.num_scalar_textures), make_int4(0, 0, 0, 0));
        std::vector<int4> dim(size_t(v.num_scalar_textures), make_int4(0, 0, 0, 0));
        const float *root = v.node_bounds;
        for (int t = 0; t < v.num_scalar_textures; t++) {
            if (v.scalar_texture_type[t] != LAMBDA_SCALAR_WORLEY) continue;
            const float f = v.scalar_texture_params[3 * t + 2];
            long long lo[3], hi[3];
            bool ok = std::isfinite(f);
            for (int a = 0; a < 3 && ok; a++) {
                double c[4] = {double(root[a]) * f, double(root[3 + a]) * f, -1.0 * f, 1.0 * f};
                double mn = std::min(std::min(c[0], c[1]), std::min(c[2], c[3]));
                double mx = std::max(std::max(c[0], c[1]), std::max(c[2], c[3]));
                ok = std::isfinite(mn) && std::isfinite(mx) && std::fabs(mn) < 1e7 && std::fabs(mx) < 1e7;
                lo[a] = (long long)std::floor(mn) - 2;
                hi[a] = (long long)std::floor(mx) + 2;
            }
            if (!ok) continue;
            const size_t nx = size_t(hi[0] - lo[0] + 1), ny = size_t(hi[1] - lo[1] + 1), nz = size_t(hi[2] - lo[2] + 1);
            if (nx * ny * nz > kMaxCells || pts.size() + nx * ny * nz > 4 * kMaxCells) continue;
            box[size_t(t)] = make_int4(int(lo[0]), int(lo[1]), int(lo[2]), int(pts.size()));
            dim[size_t(t)] = make_int4(int(nx), int(ny), int(nz), 0);
            for (size_t z = 0; z < nz; z++)
                for (size_t y = 0; y < ny; y++)
                    for (size_t x = 0; x < nx; x++) {
                        V3 q = worley_feature(int(lo[0] + (long long)x), int(lo[1] + (long long)y), int(lo[2] + (long long)z));
                        pts.push_back(make_float4(q.x, q.y, q.z, 0));
                    }
        }
        if (pts.empty()) pts.push_back(make_float4(0, 0, 0, 0));
        g.worley_points.upload(pts);
        g.worley_box.upload(box);
        g.worley_dim.upload(dim);
    }

    DeviceScene &d = g.ds;
    d.wide = g.wide.p;
    d.inst_root = g.inst_root.p;
    d.obj = g.obj.p;
    d.obj_s = g.obj_s.p;
    d.obj_mat = g.obj_mat.p;
    d.inst = g.inst.p;
    d.inst_i = g.inst_i.p;
    d.mat = g.mat.p;
    d.stex = g.stex.p;
    d.xtex = g.xtex.p;
    d.mat_tex = g.mat_tex.p;
    d.tex_avg = g.tex_avg.p;
    d.spectra = g.spectra.p;
    d.cspectra = g.cspec.p;
    d.props = g.props.p;
    d.light_io = g.light_io.p;
    d.light_pow = g.light_pow.p;
    d.cie = g.cie.p;
    d.perm = g.perm.p;
    d.tlas_rank = g.tlas_rank.p;
    d.worley_points = g.worley_points.p;
    d.worley_box = g.worley_box.p;
    d.worley_dim = g.worley_dim.p;
    d.obj_rank = g.obj_rank.p;
    const float *c = v.camera;
    d.cam_pos = make_float4(c[0], c[1], c[2], 0);
    d.cam_corner = make_float4(c[3], c[4], c[5], 0);
    d.cam_h = make_float4(c[6], c[7], c[8], 0);
    d.cam_v = make_float4(c[9], c[10], c[11], 0);
    d.total_power = v.total_light_power;
    d.num_lights = v.num_lights;
    d.bg_tex = v.background_texture;
    d.num_nodes = v.num_nodes;
    d.num_objects = v.num_objects;
    d.num_instances = v.num_instances;
}

void check_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw Failure{LAMBDA_ENODEV, "no HIP device visible"};
}

// Pixels in tile x tile squares, tiles row-major: neighbouring lanes trace
// neighbouring pixels.
std::vector<int> tile_order(int w, int h, int tile) {
    std::vector<int> order;
    order.reserve(size_t(w) * h);
    for (int ty = 0; ty < h; ty += tile)
        for (int tx = 0; tx < w; tx += tile)
            for (int y = ty; y < std::min(ty + tile, h); y++)
                for (int x = tx; x < std::min(tx + tile, w); x++) order.push_back(y * w + x);
    return order;
}

struct RenderBuffers {
    DevBuf<float4> o, d, lam, beta, L, s_o, s_d, s_c;
    DevBuf<uint4> rng;
    DevBuf<int4> info, s_i;
    DevBuf<int2> hit;
    DevBuf<int> qa, qb, done, ctr, order;
    DevBuf<unsigned long long> rays;
    DevBuf<double> fb;
    DevBuf<float> rgb;
    DevBuf<int2> spill;
};

void validate_params(const lambda_render_params &p) {
    if (p.width <= 0 || p.height <= 0 || p.depth < 0 || p.samples < 1 || p.sqrt_samples < 1 ||
        p.sqrt_samples * p.sqrt_samples != p.samples)
        throw Failure{LAMBDA_EINVAL, "invalid render parameters (size, depth or samples)"};
    if (!(p.lambda_min >= 360 && p.lambda_max <= 830 && p.lambda_min < p.lambda_max))
        throw Failure{LAMBDA_EINVAL, "invalid wavelength range"};
    if (p.rng_mode != LAMBDA_RNG_REFERENCE && p.rng_mode != LAMBDA_RNG_THROUGHPUT)
        throw Failure{LAMBDA_EINVAL, "invalid rng_mode"};
    if (p.tile != 0 && p.tile != 4 && p.tile != 8 && p.tile != 16)
        throw Failure{LAMBDA_EINVAL, "invalid tile (4, 8 or 16; 0 = 8)"};
    if (p.occupancy != 0 && (p.occupancy < 2 || p.occupancy > 4))
        throw Failure{LAMBDA_EINVAL, "invalid occupancy (2, 3 or 4 waves per SIMD; 0 = 4)"};
    if (p.vote != 0 && p.vote != 1) throw Failure{LAMBDA_EINVAL, "invalid vote (0 or 1)"};
    if (p.vote && p.occupancy == 2) throw Failure{LAMBDA_EINVAL, "the voting schedule runs at 3 or 4 waves per SIMD"};
    if ((long long)p.width * p.height > (1LL << 28)) throw Failure{LAMBDA_EINVAL, "image too large"};
    // work-unit counters are 32-bit and may run past the last unit by the
    // number of slots (at most 2^28)
    if (p.rng_mode == LAMBDA_RNG_THROUGHPUT && (long long)p.width * p.height * p.samples >= (1LL << 31) - (1LL << 28))
        throw Failure{LAMBDA_EINVAL, "throughput mode supports fewer than 2^31 - 2^28 samples per render"};
    if ((long long)p.width * p.height * p.samples > (1LL << 40)) throw Failure{LAMBDA_EINVAL, "too many samples"};
}

int grid_for(long long n) { return int(std::max(1LL, (n + kBlock - 1) / kBlock)); }

// Global stack area for launches of up to `items` threads: entries beyond
// the LDS ring, [entry][thread]; none when the scene never needs them.
void alloc_spill(const lambda_hip_scene &g, long long items, DevBuf<int2> &spill) {
    if (g.stack_need <= kRing) return;
    const size_t threads = size_t(grid_for(items)) * kBlock;
    spill.alloc(size_t(g.stack_need - kRing) * threads);
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};


// Threads resident on the device at once for a kernel with `blocks` blocks
// of kBlock threads per compute unit.
long long resident_threads(int blocks) {
    int dev = 0, cus = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    return (long long)std::max(1, cus) * std::max(1, blocks) * kBlock;
}

long long resident_persist_threads(int occupancy = 0) { return resident_threads(persist_blocks_per_cu(occupancy)); }

// The persistent-kernel variant code of a render: waves per SIMD, plus
// kPersistVoteBit for the voting schedule.
int persist_variant(const lambda_render_params &p) { return p.occupancy | (p.vote ? kPersistVoteBit : 0); }

Frame make_frame(const lambda_render_params &p) {
    Frame F{};  // every optional pointer (cost, train) starts null
    F.width = p.width;
    F.height = p.height;
    F.depth = p.depth;
    F.spp = p.samples;
    F.sqrt_spp = p.sqrt_samples;
    F.mode = p.rng_mode;
    F.npix = p.width * p.height;
    F.lmin = p.lambda_min;
    F.lmax = p.lambda_max;
    F.lrange = p.lambda_max - p.lambda_min;
    F.seed = p.seed;
    F.total_work = (long long)F.npix * F.spp;
    F.cost = nullptr;
    F.guide_frac = kGuideFrac;
    F.rr_start = kRRStart;
    return F;
}


struct Renderer {
    lambda_hip_scene &g;
    RenderBuffers B;
    struct CostOrderBuffers {
        DevBuf<int> cost, order, tile_ids, offsets;
        DevBuf<long long> tile_cost;
        DevBuf<unsigned long long> t0;
        DevBuf<uint4> save;
    } C;                                  // two-launch cost order (allocated on first use)
    Paths P{};
    int ns = 0;
    int tile = 8;        // side of the tiles of the visiting order
    int occupancy = 0;   // persistent kernel variant (waves per SIMD, 0 = default)
    int *host_ctr = nullptr;
    hipStream_t st = nullptr;
    EventPair ev[4];

    Renderer(lambda_hip_scene &scene, const Frame &F, int slots) : g(scene), ns(slots) {
        B.o.alloc(ns); B.d.alloc(ns); B.lam.alloc(ns); B.beta.alloc(ns); B.L.alloc(ns);
        B.s_o.alloc(ns); B.s_d.alloc(ns); B.s_c.alloc(ns); B.s_i.alloc(ns);
        B.rng.alloc(ns); B.info.alloc(ns); B.hit.alloc(ns);
        B.qa.alloc(ns); B.qb.alloc(ns); B.done.alloc(ns);
        B.ctr.alloc(kCtrCount);
        B.rays.alloc(2);
        B.fb.alloc(size_t(F.npix) * 3);
        B.rgb.alloc(size_t(F.npix) * 3);
        B.order.upload(tile_order(F.width, F.height, tile));
        alloc_spill(g, 2LL * ns, B.spill);
        P.o = B.o.p; P.d = B.d.p; P.lam = B.lam.p; P.beta = B.beta.p; P.L = B.L.p;
        P.rng = B.rng.p; P.info = B.info.p; P.hit = B.hit.p;
        P.q_in = B.qa.p; P.q_out = B.qb.p; P.done = B.done.p;
        P.s_o = B.s_o.p; P.s_d = B.s_d.p; P.s_c = B.s_c.p; P.s_i = B.s_i.p;
        P.ctr = B.ctr.p; P.rays = B.rays.p; P.fb = B.fb.p; P.order = B.order.p;
        HIP_OK(hipHostMalloc((void **)&host_ctr, kCtrCount * sizeof(int), hipHostMallocDefault));
        HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    }
    ~Renderer() {
        if (host_ctr) (void)hipHostFree(host_ctr);
        if (host_prog) (void)hipHostFree(host_prog);
        if (st) (void)hipStreamDestroy(st);
        if (pst) (void)hipStreamDestroy(pst);
    }

    int *host_prog = nullptr;   // finished-pixel count read while a launch runs
    hipStream_t pst = nullptr;  // stream of those reads

    // Waits for the render stream; while it runs, copies the finished-pixel
    // counter (a copy on a stream of its own, every 20 ms) to *progress,
    // which only grows -- main.cpp's 100 ms progress poll sees a persistent
    // launch advance instead of jumping at its end.
    void wait_with_progress(const Frame &F, volatile int32_t *progress) {
        if (!progress) return;
        if (!pst) HIP_OK(hipStreamCreateWithFlags(&pst, hipStreamNonBlocking));
        if (!host_prog) HIP_OK(hipHostMalloc((void **)&host_prog, sizeof(int), hipHostMallocDefault));
        const long long units = F.mode == LAMBDA_RNG_REFERENCE ? F.npix : F.total_work;
        while (true) {
            const hipError_t q = hipStreamQuery(st);
            if (q == hipSuccess) return;
            if (q != hipErrorNotReady) HIP_OK(q);
            HIP_OK(hipMemcpyAsync(host_prog, P.ctr + kCtrPixels, sizeof(int), hipMemcpyDeviceToHost, pst));
            HIP_OK(hipStreamSynchronize(pst));
            long long done = host_prog[0];
            if (F.mode != LAMBDA_RNG_REFERENCE) done = done * F.npix / std::max<long long>(1, units);
            const int32_t v = int32_t(std::min<long long>(done, F.npix));
            if (v > *progress) *progress = v;
            std::this_thread::sleep_for(std::chrono::milliseconds(20));
        }
    }

    void sync_counters() {
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(host_ctr, P.ctr, kCtrCount * sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (host_ctr[kCtrError]) throw Failure{LAMBDA_EHIP, "device reported an inconsistent hit record"};
    }

    // Renders F into the FP64 framebuffer (added to it when clear is
    // false): one launch per bounce while at least `tail` paths are live,
    // then one persistent launch.  With a guide: fused, one launch of the
    // guided persistent kernel (the network inside it); otherwise every pass
    // runs the network on the paths at guided vertices (no persistent
    // launch).  With a table instead of a guide: the persistent launch is
    // the table-guided kernel.
    void run(const Frame &F, long long tail, volatile int32_t *progress, lambda_render_stats *stats,
             bool clear = true, const GuideNet *guide = nullptr, GuideQueue *Q = nullptr, bool fused = false,
             const GuideTableDev *table = nullptr) {
        const long long units = F.mode == LAMBDA_RNG_REFERENCE ? F.npix : F.total_work;
        if (stats) {
            for (auto &e : ev) {
                if (!e.a) HIP_OK(hipEventCreate(&e.a));
                if (!e.b) HIP_OK(hipEventCreate(&e.b));
            }
        }
        auto timed = [&](int k) {
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, ev[k].a, ev[k].b));
            stats->kernel_ms[k] += ms;
            stats->launches[k]++;
        };
        if (clear) launch_fill_d(grid_for((long long)F.npix * 3), kBlock, st, P.fb, size_t(F.npix) * 3);
        launch_fill_i(1, kBlock, st, P.ctr, kCtrCount);
        launch_fill_u64(1, kBlock, st, P.rays, 2);
        if (guide) tail = fused ? kAutoTailPaths : 0;
        launch_start(grid_for(ns), kBlock, st, g.ds, F, P, ns, 0);
        // The work counter starts after the slots' first units.
        launch_set_i(1, 1, st, P.ctr + kCtrWork, ns);
        HIP_OK(hipGetLastError());
        int n_closest = ns, n_shadow = 0, n_done = 0;
        // Every pass advances each live path by one segment; a slot runs at
        // most depth + 1 segments per sample it owns, plus one pass to retire.
        const long long per_slot = F.mode == LAMBDA_RNG_REFERENCE ? F.spp * ((units + ns - 1) / ns + 1)
                                                                  : (F.total_work + ns - 1) / ns + 1;
        const long long max_iters = (long long)(F.depth + 2) * per_slot + 16;
        long long iters = 0;
        while (n_closest > 0 || n_shadow > 0 || n_done > 0) {
            if (++iters > max_iters) throw Failure{LAMBDA_EHIP, "bounce loop did not drain"};
            if ((long long)n_closest + n_done < tail) {
                // settle pending occlusion rays and ended samples, then run
                // every live slot, and the units still to start, to the end
                if (stats) HIP_OK(hipEventRecord(ev[0].a, st));
                launch_trace(grid_for(n_shadow), kBlock, st, g.ds, P, 0, n_shadow, B.spill.p, nullptr);
                if (stats) HIP_OK(hipEventRecord(ev[0].b, st));
                launch_finalize(grid_for(n_done), kBlock, st, g.ds, F, P, n_done, nullptr);
                if (stats) HIP_OK(hipEventRecord(ev[3].a, st));
                if (guide)
                    launch_persist_guided(grid_for((long long)n_closest + n_done), kBlock, st, 
                        g.ds, F, P, n_closest, B.spill.p, guide->dev());
                else if (table)
                    launch_persist_table(grid_for((long long)n_closest + n_done), kBlock, st, g.ds, F, P, n_closest,
                                         B.spill.p, *table);
                else
                    launch_persist(grid_for((long long)n_closest + n_done), kBlock, st, g.ds, F, P, n_closest,
                                                                                          B.spill.p, occupancy);
                if (stats) HIP_OK(hipEventRecord(ev[3].b, st));
                HIP_OK(hipGetLastError());
                wait_with_progress(F, progress);
                sync_counters();
                if (stats) {
                    unsigned long long r[2];
                    HIP_OK(hipMemcpy(r, P.rays, sizeof r, hipMemcpyDeviceToHost));
                    stats->closest_rays += (long long)r[0];
                    stats->shadow_rays += n_shadow + (long long)r[1];
                    timed(0);
                    timed(3);
                }
                break;
            }
            if (stats) HIP_OK(hipEventRecord(ev[0].a, st));
            launch_trace(grid_for((long long)n_closest + n_shadow), kBlock, st, g.ds, P, n_closest, n_shadow, B.spill.p, nullptr);
            if (stats) {
                HIP_OK(hipEventRecord(ev[0].b, st));
                HIP_OK(hipEventRecord(ev[2].a, st));
            }
            launch_finalize(grid_for(n_done), kBlock, st, g.ds, F, P, n_done, nullptr);
            if (stats) {
                HIP_OK(hipEventRecord(ev[2].b, st));
                HIP_OK(hipEventRecord(ev[1].a, st));
            }
            if (guide) {
                launch_shade_guided_queue(grid_for(n_closest), kBlock, st, g.ds, F, P, n_closest, *Q);
                guide->forward_dev(Q->feat, P.ctr + kCtrGuided, n_closest, Q->out, st);
                launch_shade_guided(grid_for(n_closest), kBlock, st, g.ds, F, P, *Q);
            } else {
                launch_shade(grid_for(n_closest), kBlock, st, g.ds, F, P, n_closest, nullptr);
            }
            if (stats) HIP_OK(hipEventRecord(ev[1].b, st));
            sync_counters();
            if (stats) {
                stats->iterations++;
                stats->closest_rays += n_closest;
                stats->shadow_rays += n_shadow;
                for (int k = 0; k < 3; k++) timed(k);
            }
            n_closest = host_ctr[kCtrNext];
            n_shadow = host_ctr[kCtrShadow];
            n_done = host_ctr[kCtrDone];
            std::swap(P.q_in, P.q_out);
            if (progress) {
                long long done = host_ctr[kCtrPixels];
                if (F.mode != LAMBDA_RNG_REFERENCE) done = done * F.npix / std::max<long long>(1, units);
                *progress = int32_t(std::min<long long>(done, F.npix));
            }
        }
    }

    // Guided render whose network is used at the first F.guide_bounces
    // vertices only: the work units go in batches of one per slot; a batch
    // runs guided passes (k_trace, k_finalize, network, guided shading) while
    // its live paths are within the guided bounces, then one persistent
    // launch runs every remaining path of the batch to its end (no further
    // units: the work counter stays at the end).
    void run_guided_batches(const Frame &F, volatile int32_t *progress, lambda_render_stats *stats, bool clear,
                            const GuideNet *guide, GuideQueue *Q) {
        const long long units = F.mode == LAMBDA_RNG_REFERENCE ? F.npix : F.total_work;
        if (units > INT32_MAX / 2) throw Failure{LAMBDA_EINVAL, "too many work units for one guided render"};
        if (stats) {
            for (auto &e : ev) {
                if (!e.a) HIP_OK(hipEventCreate(&e.a));
                if (!e.b) HIP_OK(hipEventCreate(&e.b));
            }
        }
        auto timed = [&](int k) {
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, ev[k].a, ev[k].b));
            stats->kernel_ms[k] += ms;
            stats->launches[k]++;
        };
        if (clear) launch_fill_d(grid_for((long long)F.npix * 3), kBlock, st, P.fb, size_t(F.npix) * 3);
        launch_fill_i(1, kBlock, st, P.ctr, kCtrCount);
        launch_fill_u64(1, kBlock, st, P.rays, 2);
        // no unit is started after a batch's first ones
        launch_set_i(1, 1, st, P.ctr + kCtrWork, int(units));
        HIP_OK(hipGetLastError());
        for (long long base = 0; base < units; base += ns) {
            const int nb = int(std::min<long long>(ns, units - base));
            launch_start(grid_for(nb), kBlock, st, g.ds, F, P, nb, base);
            int n_closest = nb, n_shadow = 0, n_done = 0;
            for (int pass = 0; pass < F.guide_bounces && n_closest > 0; pass++) {
                if (stats) HIP_OK(hipEventRecord(ev[0].a, st));
                launch_trace(grid_for((long long)n_closest + n_shadow), kBlock, st, g.ds, P, n_closest, n_shadow,
                                                                                       B.spill.p, nullptr);
                if (stats) {
                    HIP_OK(hipEventRecord(ev[0].b, st));
                    HIP_OK(hipEventRecord(ev[2].a, st));
                }
                launch_finalize(grid_for(n_done), kBlock, st, g.ds, F, P, n_done, nullptr);
                if (stats) {
                    HIP_OK(hipEventRecord(ev[2].b, st));
                    HIP_OK(hipEventRecord(ev[1].a, st));
                }
                launch_shade_guided_queue(grid_for(n_closest), kBlock, st, g.ds, F, P, n_closest, *Q);
                guide->forward_dev(Q->feat, P.ctr + kCtrGuided, n_closest, Q->out, st);
                launch_shade_guided(grid_for(n_closest), kBlock, st, g.ds, F, P, *Q);
                if (stats) HIP_OK(hipEventRecord(ev[1].b, st));
                sync_counters();
                if (stats) {
                    stats->iterations++;
                    stats->closest_rays += n_closest;
                    stats->shadow_rays += n_shadow;
                    for (int k = 0; k < 3; k++) timed(k);
                }
                n_closest = host_ctr[kCtrNext];
                n_shadow = host_ctr[kCtrShadow];
                n_done = host_ctr[kCtrDone];
                std::swap(P.q_in, P.q_out);
            }
            // settle pending occlusion rays and ended samples, then run the
            // batch's live paths to their ends
            if (stats) HIP_OK(hipEventRecord(ev[0].a, st));
            launch_trace(grid_for(n_shadow), kBlock, st, g.ds, P, 0, n_shadow, B.spill.p, nullptr);
            if (stats) HIP_OK(hipEventRecord(ev[0].b, st));
            launch_finalize(grid_for(n_done), kBlock, st, g.ds, F, P, n_done, nullptr);
            if (stats) HIP_OK(hipEventRecord(ev[3].a, st));
            if (n_closest + n_done > 0)
                launch_persist(grid_for((long long)n_closest + n_done), kBlock, st, g.ds, F, P, n_closest,
                                                                                      B.spill.p, occupancy);
            if (stats) HIP_OK(hipEventRecord(ev[3].b, st));
            sync_counters();
            if (stats) {
                stats->shadow_rays += n_shadow;
                timed(0);
                timed(3);
            }
            if (progress) {
                const long long done = host_ctr[kCtrPixels];
                *progress = int32_t(std::min<long long>(done * F.npix / std::max<long long>(1, units), F.npix));
            }
        }
        if (stats) {
            unsigned long long r[2];
            HIP_OK(hipMemcpy(r, P.rays, sizeof r, hipMemcpyDeviceToHost));
            stats->closest_rays += (long long)r[0];
            stats->shadow_rays += (long long)r[1];
        }
    }

    // The pass schedule with the bounce loop captured in a hipGraph: a graph
    // of `passes` (even) passes -- each k_pass_begin, k_trace, k_finalize,
    // k_shade with queue lengths read on the device and grids sized for the
    // slot count -- is replayed until the queues are empty; the host reads
    // the counters once per replay.  The graph holds kernels only.
    void run_graph(const Frame &F, int passes, volatile int32_t *progress, lambda_render_stats *stats) {
        const long long units = F.mode == LAMBDA_RNG_REFERENCE ? F.npix : F.total_work;
        passes = std::max(2, passes + (passes & 1));
        DevBuf<int> counts;
        counts.alloc(3);
        launch_fill_d(grid_for((long long)F.npix * 3), kBlock, st, P.fb, size_t(F.npix) * 3);
        launch_fill_i(1, kBlock, st, P.ctr, kCtrCount);
        launch_fill_u64(1, kBlock, st, P.rays, 2);
        launch_start(grid_for(ns), kBlock, st, g.ds, F, P, ns, 0);
        launch_set_i(1, 1, st, P.ctr + kCtrWork, ns);
        launch_set_i(1, 1, st, P.ctr + kCtrNext, ns);  // every slot starts in the closest-hit queue
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(st));
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        struct GraphFree {
            hipGraph_t *g;
            hipGraphExec_t *e;
            ~GraphFree() {
                if (*e) (void)hipGraphExecDestroy(*e);
                if (*g) (void)hipGraphDestroy(*g);
            }
        } gf{&graph, &exec};
        HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        Paths Q = P;
        for (int k = 0; k < passes; k++) {
            launch_pass_begin(1, 1, st, Q.ctr, counts.p);
            launch_trace(grid_for(2LL * ns), kBlock, st, g.ds, Q, 0, 0, B.spill.p, counts.p);
            launch_finalize(grid_for(ns), kBlock, st, g.ds, F, Q, 0, counts.p);
            launch_shade(grid_for(ns), kBlock, st, g.ds, F, Q, 0, counts.p);
            std::swap(Q.q_in, Q.q_out);
        }
        HIP_OK(hipStreamEndCapture(st, &graph));
        HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        const long long per_slot = F.mode == LAMBDA_RNG_REFERENCE ? F.spp * ((units + ns - 1) / ns + 1)
                                                                  : (F.total_work + ns - 1) / ns + 1;
        const long long max_replays = ((long long)(F.depth + 2) * per_slot + 16) / passes + 2;
        for (long long r = 0;; r++) {
            if (r > max_replays) throw Failure{LAMBDA_EHIP, "bounce loop did not drain"};
            HIP_OK(hipGraphLaunch(exec, st));
            sync_counters();
            if (stats) stats->iterations += passes;
            if (progress) {
                long long done = host_ctr[kCtrPixels];
                if (F.mode != LAMBDA_RNG_REFERENCE) done = done * F.npix / std::max<long long>(1, units);
                *progress = int32_t(std::min<long long>(done, F.npix));
            }
            if (host_ctr[kCtrNext] == 0 && host_ctr[kCtrShadow] == 0 && host_ctr[kCtrDone] == 0) break;
        }
    }

    // Reference streams in two launches: the first `first` samples of every
    // pixel in tile order, timing each pixel on the device clock and saving
    // its stream; then the remaining samples with the tiles in decreasing
    // order of their time, so that the longest sample chains start first and
    // the launch's drain is short.  The image is the same as one launch's:
    // each pixel's samples still run in order on its own stream.
    void run_cost_ordered(const Frame &F, int first, long long tail, volatile int32_t *progress,
                          lambda_render_stats *stats) {
        const int tiles_x = (F.width + tile - 1) / tile, tiles_y = (F.height + tile - 1) / tile;
        const int ntiles = tiles_x * tiles_y;
        DevBuf<int> &cost = C.cost, &order = C.order, &tile_ids = C.tile_ids, &offsets = C.offsets;
        DevBuf<long long> &tile_cost = C.tile_cost;
        DevBuf<unsigned long long> &t0 = C.t0;
        DevBuf<uint4> &save = C.save;
        cost.ensure(size_t(F.npix));
        order.ensure(size_t(F.npix));
        tile_cost.ensure(size_t(ntiles));
        tile_ids.ensure(size_t(ntiles));
        offsets.ensure(size_t(ntiles));
        t0.ensure(size_t(ns));
        save.ensure(size_t(F.npix));
        launch_fill_i(grid_for(F.npix), kBlock, st, cost.p, F.npix);
        launch_fill_u64(grid_for(ntiles), kBlock, st, reinterpret_cast<unsigned long long *>(tile_cost.p),
                                                         size_t(ntiles));
        Frame F1 = F;
        F1.sample_end = first;
        F1.cost = cost.p;
        F1.unit_t0 = t0.p;
        F1.rng_save = save.p;
        run(F1, tail, nullptr, stats, true);
        launch_tile_cost(grid_for(F.npix), kBlock, st, cost.p, F.width, F.height, tile, tile_cost.p);
        std::vector<long long> tc(static_cast<size_t>(ntiles));
        HIP_OK(hipMemcpyAsync(tc.data(), tile_cost.p, tc.size() * sizeof(long long), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        std::vector<int> ids(static_cast<size_t>(ntiles)), offs(static_cast<size_t>(ntiles));
        for (int t = 0; t < ntiles; t++) ids[size_t(t)] = t;
        std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return tc[size_t(a)] > tc[size_t(b)]; });
        int pos = 0;
        for (int i = 0; i < ntiles; i++) {
            const int t = ids[size_t(i)];
            offs[size_t(i)] = pos;
            pos += std::min(tile, F.width - tile * (t % tiles_x)) * std::min(tile, F.height - tile * (t / tiles_x));
        }
        HIP_OK(hipMemcpyAsync(tile_ids.p, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(offsets.p, offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice, st));
        launch_tile_expand(grid_for((long long)ntiles * tile * tile), kBlock, st, tile_ids.p, offsets.p, ntiles, F.width,
                                                                           F.height, tile, order.p);
        HIP_OK(hipGetLastError());
        Frame F2 = F;
        F2.sample_begin = first;
        F2.rng_save = save.p;
        const int *tiles = P.order;
        P.order = order.p;
        try {
            run(F2, tail, progress, stats, false);
        } catch (...) {
            P.order = tiles;
            throw;
        }
        P.order = tiles;
    }
};

void render_with(Renderer &R, const Frame &F, const lambda_render_params &p, long long tail, float *out_rgb,
                 volatile int32_t *progress, lambda_render_stats *stats);

void render_impl(lambda_hip_scene &g, const lambda_render_params &p, float *out_rgb, volatile int32_t *progress,
                 lambda_render_stats *stats = nullptr) {
    auto wall0 = std::chrono::steady_clock::now();
    validate_params(p);
    const Frame F = make_frame(p);
    // Slots: paths in flight.  Reference mode: at most one per pixel (a slot
    // owns a pixel for all its samples).
    const long long units = F.mode == LAMBDA_RNG_REFERENCE ? F.npix : F.total_work;
    long long slots = p.path_slots > 0 ? p.path_slots : resident_persist_threads(persist_variant(p));
    slots = std::max<long long>(1, std::min<long long>(slots, units));
    if (slots > (1LL << 28)) throw Failure{LAMBDA_EINVAL, "too many path slots"};
    // Below this many live paths the rest of the render runs in one
    // persistent launch (per-thread chains instead of one pass per bounce).
    const long long tail = p.tail_paths < 0 ? 0 : p.tail_paths > 0 ? p.tail_paths : kAutoTailPaths;
    std::lock_guard<std::mutex> lock(g.render_mutex);
    if (!g.renderer || g.r_width != F.width || g.r_height != F.height || g.r_slots != int(slots)) {
        g.renderer.reset();  // the old buffers go first
        g.renderer = std::make_shared<Renderer>(g, F, int(slots));
        g.r_width = F.width;
        g.r_height = F.height;
        g.r_slots = int(slots);
    }
    Renderer &R = *static_cast<Renderer *>(g.renderer.get());
    try {
        render_with(R, F, p, tail, out_rgb, progress, stats);
    } catch (...) {
        g.renderer.reset();  // a failed render leaves no half-used buffers behind
        throw;
    }
    if (stats)
        stats->render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
}

// One render with the buffers of R.
void render_with(Renderer &R, const Frame &F, const lambda_render_params &p, long long tail, float *out_rgb,
                 volatile int32_t *progress, lambda_render_stats *stats) {
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->path_slots = R.ns;
    }
    const int tile = p.tile > 0 ? p.tile : 8;
    if (tile != R.tile) {
        R.B.order.upload(tile_order(F.width, F.height, tile));
        R.P.order = R.B.order.p;
        R.tile = tile;
    }
    R.occupancy = persist_variant(p);
    // reference streams, persistent schedule, more pixels than slots: the
    // two-launch cost order (cost_order samples first; 0 = automatic: 4, the
    // best measured on the 1080p instanced-mesh frame against 1, 2 and 8)
    const bool by_cost = F.mode == LAMBDA_RNG_REFERENCE && F.spp > 1 && p.graph_passes <= 0 &&
                         tail > R.ns && (p.cost_order > 0 || (p.cost_order == 0 && R.ns < F.npix));
    if (by_cost) R.run_cost_ordered(F, std::min(p.cost_order > 0 ? p.cost_order : 4, F.spp - 1), tail, progress, stats);
    else if (p.graph_passes > 0) R.run_graph(F, p.graph_passes, progress, stats);
    else R.run(F, tail, progress, stats);
    launch_resolve(grid_for(F.npix), kBlock, R.st, F, R.P.fb, R.B.rgb.p);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out_rgb, R.B.rgb.p, size_t(F.npix) * 3 * sizeof(float), hipMemcpyDeviceToHost, R.st));
    HIP_OK(hipStreamSynchronize(R.st));
    if (progress) *progress = F.npix;
}

}  // namespace

// Progressive rendering with throughput streams: each step adds samples
// with fresh PCG32 streams to one FP64 framebuffer; with a guide, Lambertian
// vertices sample the guiding mixture and each step can train the network on
// the radiance its own paths gathered.
// Guided vertices per sample when a session does not set them: the first
// eight bounces (on the guiding room with roulette on, 2, 4, 6, 10, 12 or 16
// gave less variance reduction per unit time; with roulette off, as guided
// sessions run, 12 and every bounce were within the run-to-run spread).
constexpr int kDefaultGuidedBounces = 8;
// Tabulated guide limits: cells along the longest side, entries in all
// (2^24 entries of 40 floats = 2.7 GB).
constexpr int kMaxTableCells = 1024;
constexpr long long kMaxTableEntries = 1LL << 24;

struct lambda_hip_progressive {
    lambda_hip_scene *scene = nullptr;
    lambda_render_params params{};
    GuideNet *net = nullptr;
    std::unique_ptr<Renderer> R;
    DevBuf<int> q_slot;
    DevBuf<float> q_feat, q_out, train;
    DevBuf<float4> grec;
    GuideQueue Q{};
    long long total = 0;
    int steps = 0;
    int train_cap = 1 << 18;
    int batch = 1 << 14;
    int epochs = 1;
    int guide_bounces = kDefaultGuidedBounces;
    float guide_frac = kGuideFrac;  // probability of sampling the guide
    // bounce index from which Russian roulette applies: the reference's for
    // plain sessions; guided sessions run without it (set at creation)
    int rr_start = kRRStart;
    bool fused = false;          // guided: network inside the persistent kernel
    double seen_per_sample = 0;  // training candidates per sample per pixel, last step
    DevBuf<double> scratch_fb;   // framebuffer of steps that are not kept
    // tabulated guide (cells along the box's longest side; 0 = off)
    int table_cells = 0;
    long long table_entries = 0;
    GuideTableDev table{};
    DevBuf<float> table_rows, table_feat;
    bool table_stale = true;     // the weights changed since the table was built
};

namespace {

// Frame of one step; guided vertices are recorded for training only in
// steps that train.
Frame progressive_frame(const lambda_hip_progressive &pr, int samples, int root, bool train) {
    Frame F = make_frame(pr.params);
    F.mode = LAMBDA_RNG_THROUGHPUT;
    F.spp = samples;
    F.sqrt_spp = root;
    F.total_work = (long long)F.npix * samples;
    F.seed = pr.params.seed + 0x9e3779b97f4a7c15ull * uint64_t(pr.steps + 1);
    F.box_lo = make_float3(pr.scene->box[0], pr.scene->box[1], pr.scene->box[2]);
    F.box_hi = make_float3(pr.scene->box[3], pr.scene->box[4], pr.scene->box[5]);
    if (pr.net) {
        F.guide_bounces = pr.guide_bounces;
        F.guide_frac = pr.guide_frac;
    }
    F.rr_start = pr.rr_start;
    if (pr.net && train) {
        F.train = pr.train.p;
        F.train_cap = pr.train_cap;
        F.grec_stride = pr.params.depth + 1;
        // expected candidates from the last step (first step: 4 per path)
        const double expect = (pr.seen_per_sample > 0 ? pr.seen_per_sample : 4.0 * F.npix) * samples;
        F.train_keep = float(std::min(1.0, 1.1 * pr.train_cap / std::max(1.0, expect)));
    }
    return F;
}

template <typename F>
int guarded(char *err, size_t len, F &&fn) {
    try {
        fn();
        return LAMBDA_OK;
    } catch (const Failure &f) {
        set_err(err, len, f.msg);
        return f.code;
    } catch (const std::bad_alloc &) {
        set_err(err, len, "host allocation failed");
        return LAMBDA_ENOMEM;
    } catch (const std::exception &e) {
        set_err(err, len, e.what());
        return LAMBDA_EINVAL;
    } catch (...) {
        set_err(err, len, "unknown failure");
        return LAMBDA_EINVAL;
    }
}

}  // namespace


// ------------------------------------------------------------ live serving
//
// A scheduler thread owns every HIP call of a server.  Submitted jobs wait
// in a FIFO; a group of them is admitted while the in-flight jobs' buffers
// (36 bytes per pixel: the FP64 XYZ sums and the RGB result) fit the HBM
// budget, and runs as one k_serve launch on one of the server's launch
// slots (stream, path state, buffers that only grow).  Launches on
// different slots overlap: a launch sized to fewer pixels than the resident
// threads, or one draining its last pixels, leaves compute units to the
// next.  Buffers are never freed while launches run (hipFree may wait for
// the device); they are reused by the slot's next launch.

struct lambda_server {
    struct Job {
        lambda_job job;
        int64_t ticket = 0;
        double submit_ms = 0;
        int32_t state = 0;          // LAMBDA_JOB_*
        lambda_job_result result{};
        std::string error;
    };
    struct Slot {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        RenderBuffers B;            // path state of the launch's threads
        DevBuf<ServeCtl> ctl;
        DevBuf<double> fb;
        DevBuf<float> rgb;
        DevBuf<int> done;
        DevBuf<unsigned long long> finish;
        DevBuf<ServeJob> djobs;
        std::vector<ServeJob> hjobs;
        float *staging = nullptr;   // pinned: RGB of the launch's jobs
        size_t staging_n = 0;
        unsigned long long *hfinish = nullptr;  // pinned: finish ticks, then t0, error flag, rays
        size_t hfinish_n = 0;
        std::vector<int64_t> tickets;
        long long bytes = 0;        // budget held by the launch
        int launch_index = -1;
        bool busy = false;
    };
    std::vector<DeviceScene> table;
    DevBuf<DeviceScene> dscenes;
    int stack_need = 0;
    long long threads = 0;
    int64_t budget = 0;
    double ticks_per_ms = 100000.0;
    unsigned long long cal_tick = 0;
    double cal_ms = 0;
    std::chrono::steady_clock::time_point t0;
    std::vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int64_t> pending;
    std::unordered_map<int64_t, Job> jobs;
    int64_t next_ticket = 1;
    long long inflight_bytes = 0;
    int launches = 0;
    long long paths = 0, closest = 0, shadow = 0;
    double last_finish_ms = 0;
    std::string fatal;
    bool stop = false;
    std::thread worker;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    double tick_ms(unsigned long long tick) const { return cal_ms + (double(tick) - double(cal_tick)) / ticks_per_ms; }
    void run();
    void launch(Slot &sl, std::vector<int64_t> batch);
    void finalize(Slot &sl);
    ~lambda_server() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv.notify_all();
        if (worker.joinable()) worker.join();
        for (auto &sl : slots) {
            if (sl->staging) (void)hipHostFree(sl->staging);
            if (sl->hfinish) (void)hipHostFree(sl->hfinish);
            if (sl->ev) (void)hipEventDestroy(sl->ev);
            if (sl->st) (void)hipStreamDestroy(sl->st);
        }
    }
};

namespace {

constexpr long long kServeBytesPerPixel = 3 * 8 + 3 * 4;

template <typename T>
void grow(DevBuf<T> &b, size_t n) {
    if (b.n < n) b.alloc(n);
}

template <typename T>
void grow_pinned(T *&p, size_t &cap, size_t n) {
    if (cap >= n) return;
    if (p) HIP_OK(hipHostFree(p));
    p = nullptr;
    cap = 0;
    HIP_OK(hipHostMalloc((void **)&p, n * sizeof(T), hipHostMallocDefault));
    cap = n;
}

}  // namespace

// One launch for the admitted jobs `batch` (the caller has taken their
// budget): the jobs with the longest per-pixel chains first.
void lambda_server::launch(Slot &sl, std::vector<int64_t> batch) {
    std::vector<const Job *> js;
    {
        std::lock_guard<std::mutex> g(mu);
        for (int64_t t : batch) js.push_back(&jobs.at(t));
    }
    std::vector<size_t> order(js.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const lambda_render_params &pa = js[a]->job.params, &pb = js[b]->job.params;
        return (long long)pa.samples * (pa.depth + 1) > (long long)pb.samples * (pb.depth + 1);
    });
    sl.tickets.clear();
    long long px_total = 0;
    for (size_t k : order) {
        sl.tickets.push_back(js[k]->ticket);
        px_total += (long long)js[k]->job.params.width * js[k]->job.params.height;
    }
    const size_t nj = sl.tickets.size();
    grow(sl.fb, size_t(px_total) * 3);
    grow(sl.rgb, size_t(px_total) * 3);
    grow(sl.done, nj);
    grow(sl.finish, nj);
    grow(sl.djobs, nj);
    grow_pinned(sl.staging, sl.staging_n, size_t(px_total) * 3);
    grow_pinned(sl.hfinish, sl.hfinish_n, nj + 4);
    sl.hjobs.assign(nj, ServeJob{});
    long long units = 0;
    for (size_t k = 0; k < nj; k++) {
        lambda_render_params p = js[order[k]]->job.params;
        p.rng_mode = LAMBDA_RNG_REFERENCE;
        ServeJob &J = sl.hjobs[k];
        J.F = make_frame(p);
        J.fb = sl.fb.p + units * 3;
        J.first_unit = units;
        J.arrival = 0;
        J.scene = js[order[k]]->job.scene;
        J.units = J.F.npix;
        units += J.units;
    }
    Paths P{};
    RenderBuffers &B = sl.B;
    P.o = B.o.p; P.d = B.d.p; P.lam = B.lam.p; P.beta = B.beta.p; P.L = B.L.p;
    P.rng = B.rng.p; P.info = B.info.p; P.hit = B.hit.p; P.ctr = B.ctr.p; P.rays = B.rays.p;
    const hipStream_t st = sl.st;
    HIP_OK(hipMemcpyAsync(sl.djobs.p, sl.hjobs.data(), nj * sizeof(ServeJob), hipMemcpyHostToDevice, st));
    launch_fill_d(grid_for(px_total * 3), kBlock, st, sl.fb.p, size_t(px_total) * 3);
    launch_fill_i(grid_for((long long)nj), kBlock, st, sl.done.p, int(nj));
    launch_fill_i(1, kBlock, st, P.ctr, kCtrCount);
    launch_fill_u64(1, kBlock, st, P.rays, 2);
    launch_serve_clock(1, 1, st, sl.ctl.p);
    const long long grid_threads = std::min<long long>(threads, (units + kBlock - 1) / kBlock * kBlock);
    launch_serve(int(grid_threads / kBlock), kBlock, st, dscenes.p, sl.djobs.p, int(nj), units, P, sl.ctl.p,
                 sl.done.p, sl.finish.p, B.spill.p);
    long long off = 0;
    for (size_t k = 0; k < nj; k++) {
        launch_resolve(grid_for(sl.hjobs[k].F.npix), kBlock, st, sl.hjobs[k].F, sl.hjobs[k].fb, sl.rgb.p + off * 3);
        off += sl.hjobs[k].F.npix;
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(sl.staging, sl.rgb.p, size_t(px_total) * 3 * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(sl.hfinish, sl.finish.p, nj * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(sl.hfinish + nj, &sl.ctl.p->t0, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(sl.hfinish + nj + 1, P.ctr + kCtrError, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(sl.hfinish + nj + 2, P.rays, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_OK(hipEventRecord(sl.ev, st));
    std::lock_guard<std::mutex> g(mu);
    sl.launch_index = launches++;
    for (int64_t t : sl.tickets) jobs.at(t).state = LAMBDA_JOB_RUNNING;
}

// Results of a finished launch: images to the callers' buffers, latencies
// from the device finish ticks.
void lambda_server::finalize(Slot &sl) {
    const size_t nj = sl.tickets.size();
    const bool bad = (sl.hfinish[nj + 1] & 0xffffffffull) != 0;
    long long off = 0;
    std::lock_guard<std::mutex> g(mu);
    for (size_t k = 0; k < nj; k++) {
        Job &jb = jobs.at(sl.tickets[k]);
        const ServeJob &J = sl.hjobs[k];
        if (bad) {
            jb.state = LAMBDA_JOB_FAILED;
            jb.error = "device reported an inconsistent hit record";
        } else {
            std::memcpy(jb.job.out_rgb, sl.staging + off * 3, size_t(J.F.npix) * 3 * sizeof(float));
            jb.result.finish_ms = tick_ms(sl.hfinish[k]);
            jb.result.latency_ms = jb.result.finish_ms - jb.submit_ms;
            jb.result.batch = sl.launch_index;
            jb.state = LAMBDA_JOB_DONE;
            last_finish_ms = std::max(last_finish_ms, jb.result.finish_ms);
        }
        off += J.F.npix;
        paths += (long long)J.F.npix * J.F.spp;
    }
    closest += (long long)sl.hfinish[nj + 2];
    shadow += (long long)sl.hfinish[nj + 3];
    inflight_bytes -= sl.bytes;
    sl.bytes = 0;
    sl.busy = false;
    cv.notify_all();
}

void lambda_server::run() {
    try {
        while (true) {
            bool progressed = false;
            for (auto &slp : slots) {
                Slot &sl = *slp;
                if (!sl.busy) continue;
                const hipError_t q = hipEventQuery(sl.ev);
                if (q == hipErrorNotReady) continue;
                HIP_OK(q);
                finalize(sl);
                progressed = true;
            }
            Slot *free_slot = nullptr;
            std::vector<int64_t> batch;
            {
                std::unique_lock<std::mutex> g(mu);
                bool any_busy = false;
                for (auto &slp : slots) {
                    if (slp->busy) any_busy = true;
                    else if (!free_slot) free_slot = slp.get();
                }
                if (stop && pending.empty() && !any_busy) return;
                if (free_slot) {
                    long long bytes = 0;
                    while (!pending.empty()) {
                        const Job &jb = jobs.at(pending.front());
                        const long long jbytes =
                            (long long)jb.job.params.width * jb.job.params.height * kServeBytesPerPixel;
                        const bool alone = !any_busy && batch.empty();
                        if (budget > 0 && inflight_bytes + bytes + jbytes > budget && !alone) break;
                        bytes += jbytes;
                        batch.push_back(pending.front());
                        pending.pop_front();
                    }
                    if (!batch.empty()) {
                        // the slot is taken now, so that a drain cannot see
                        // the group neither queued nor running
                        inflight_bytes += bytes;
                        free_slot->bytes = bytes;
                        free_slot->busy = true;
                    }
                }
                if (batch.empty() && !progressed) {
                    // nothing to start: wait for a submission, or poll the launches
                    cv.wait_for(g, std::chrono::microseconds(any_busy ? 200 : 20000));
                    continue;
                }
            }
            if (!batch.empty()) launch(*free_slot, std::move(batch));
        }
    } catch (const Failure &f) {
        std::lock_guard<std::mutex> g(mu);
        fatal = f.msg;
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> g(mu);
        fatal = e.what();
    }
    std::lock_guard<std::mutex> g(mu);
    for (auto &kv : jobs)
        if (kv.second.state != LAMBDA_JOB_DONE) {
            kv.second.state = LAMBDA_JOB_FAILED;
            kv.second.error = fatal;
        }
    for (auto &slp : slots) slp->busy = false;
    pending.clear();
    inflight_bytes = 0;
    cv.notify_all();
}

namespace {

void server_setup(lambda_server &srv, lambda_hip_scene *const *scenes, int32_t num_scenes, int32_t max_launches) {
    srv.t0 = std::chrono::steady_clock::now();
    srv.table.resize(size_t(num_scenes));
    for (int k = 0; k < num_scenes; k++) {
        if (!scenes[k]) throw Failure{LAMBDA_EINVAL, "null scene"};
        srv.table[size_t(k)] = scenes[k]->ds;
        srv.stack_need = std::max(srv.stack_need, scenes[k]->stack_need);
    }
    srv.dscenes.upload(srv.table);
    srv.threads = resident_persist_threads();
    int dev = 0, rate_khz = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
    srv.ticks_per_ms = double(std::max(rate_khz, 1));
    for (int k = 0; k < max_launches; k++) {
        auto sl = std::make_unique<lambda_server::Slot>();
        HIP_OK(hipStreamCreateWithFlags(&sl->st, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&sl->ev, hipEventDisableTiming));
        const size_t n = size_t(srv.threads);
        RenderBuffers &B = sl->B;
        B.o.alloc(n); B.d.alloc(n); B.lam.alloc(n); B.beta.alloc(n); B.L.alloc(n); B.rng.alloc(n);
        B.info.alloc(n); B.hit.alloc(n);
        B.ctr.alloc(kCtrCount);
        B.rays.alloc(2);
        if (srv.stack_need > kRing) B.spill.alloc(size_t(srv.stack_need - kRing) * n);
        sl->ctl.alloc(1);
        srv.slots.push_back(std::move(sl));
    }
    // device clock -> server milliseconds: one clock read bracketed by host times
    lambda_server::Slot &s0 = *srv.slots[0];
    unsigned long long tick = 0;
    const double before = srv.now_ms();
    launch_serve_clock(1, 1, s0.st, s0.ctl.p);
    HIP_OK(hipMemcpyAsync(&tick, &s0.ctl.p->t0, sizeof tick, hipMemcpyDeviceToHost, s0.st));
    HIP_OK(hipStreamSynchronize(s0.st));
    srv.cal_ms = 0.5 * (before + srv.now_ms());
    srv.cal_tick = tick;
}

}  // namespace

extern "C" {

int lambda_hip_scene_create(const lambda_scene_view *scene, lambda_hip_scene **out, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || !out) throw Failure{LAMBDA_EINVAL, "null argument"};
        *out = nullptr;
        check_device();
        auto *g = new lambda_hip_scene;
        try {
            upload_scene(*scene, *g);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

void lambda_hip_scene_destroy(lambda_hip_scene *scene) { delete scene; }

int lambda_hip_scene_create_lbvh(const lambda_scene_view *scene, lambda_hip_scene **out, int32_t *object_order,
                                 int32_t *instance_order, double *build_ms, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || !out) throw Failure{LAMBDA_EINVAL, "null argument"};
        *out = nullptr;
        check_device();
        const lambda_scene_view &v = *scene;
        if (v.num_objects <= 0 || v.num_instances <= 0 || !v.object_type || !v.object_geometry || !v.instance_ints ||
            !v.instance_matrices || !v.object_material || !v.object_shading || v.num_lights < 1 || !v.light_instances ||
            !v.light_objects)
            throw Failure{LAMBDA_EINVAL, "invalid scene: empty scene or null array"};
        const auto t0 = std::chrono::steady_clock::now();
        LbvhScene L;
        build_lbvh(v, L);
        HIP_OK(hipDeviceSynchronize());
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        const size_t O = size_t(v.num_objects), I = size_t(v.num_instances);
        std::vector<int> inv_obj(O), inv_inst(I);
        for (size_t k = 0; k < O; k++) inv_obj[size_t(L.object_order[k])] = int(k);
        for (size_t k = 0; k < I; k++) inv_inst[size_t(L.instance_order[k])] = int(k);
        std::vector<int> otype(O), omat(O), iints(I * 4), lins(size_t(v.num_lights)), lobj(size_t(v.num_lights));
        std::vector<float> ogeom(O * 12), oshade(O * 15), imats(I * 32);
        for (size_t k = 0; k < O; k++) {
            const size_t src = size_t(L.object_order[k]);
            otype[k] = v.object_type[src];
            omat[k] = v.object_material[src];
            std::memcpy(&ogeom[k * 12], v.object_geometry + src * 12, 12 * sizeof(float));
            std::memcpy(&oshade[k * 15], v.object_shading + src * 15, 15 * sizeof(float));
        }
        for (size_t k = 0; k < I; k++) {
            const size_t src = size_t(L.instance_order[k]);
            iints[k * 4 + 0] = v.instance_ints[src * 4 + 0];
            iints[k * 4 + 1] = v.instance_ints[src * 4 + 1];
            iints[k * 4 + 2] = L.blas_root[src];
            iints[k * 4 + 3] = v.instance_ints[src * 4 + 3];
            std::memcpy(&imats[k * 32], v.instance_matrices + src * 32, 32 * sizeof(float));
        }
        for (int k = 0; k < v.num_lights; k++) {
            const int li = v.light_instances[k], lo = v.light_objects[k];
            if ((li != -1 && (li < 0 || size_t(li) >= I)) || (lo != -1 && (lo < 0 || size_t(lo) >= O)))
                throw Failure{LAMBDA_EINVAL, "invalid scene: light index out of range"};
            lins[size_t(k)] = li == -1 ? -1 : inv_inst[size_t(li)];
            lobj[size_t(k)] = lo == -1 ? -1 : inv_obj[size_t(lo)];
        }
        lambda_scene_view w = v;
        w.object_type = otype.data();
        w.object_material = omat.data();
        w.object_geometry = ogeom.data();
        w.object_shading = oshade.data();
        w.instance_ints = iints.data();
        w.instance_matrices = imats.data();
        w.num_nodes = int32_t(L.node_ints.size() / 2);
        w.node_bounds = L.node_bounds.data();
        w.node_ints = L.node_ints.data();
        w.light_instances = lins.data();
        w.light_objects = lobj.data();
        auto g = std::make_unique<lambda_hip_scene>();
        upload_scene(w, *g, kLbvhMaxDepth);
        if (object_order) std::memcpy(object_order, L.object_order.data(), O * sizeof(int32_t));
        if (instance_order) std::memcpy(instance_order, L.instance_order.data(), I * sizeof(int32_t));
        if (build_ms) *build_ms = ms;
        *out = g.release();
    });
}

int lambda_hip_scene_render(lambda_hip_scene *scene, const lambda_render_params *params, float *out_rgb,
                            volatile int32_t *progress, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || !params || !out_rgb) throw Failure{LAMBDA_EINVAL, "null argument"};
        render_impl(*scene, *params, out_rgb, progress);
    });
}

int lambda_hip_scene_render_stats(lambda_hip_scene *scene, const lambda_render_params *params, float *out_rgb,
                                  lambda_render_stats *stats, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || !params || !out_rgb || !stats) throw Failure{LAMBDA_EINVAL, "null argument"};
        render_impl(*scene, *params, out_rgb, nullptr, stats);
    });
}

int lambda_hip_render(const lambda_scene_view *scene, const lambda_render_params *params, float *out_rgb,
                      volatile int32_t *progress, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || !params || !out_rgb) throw Failure{LAMBDA_EINVAL, "null argument"};
        validate_params(*params);
        check_device();
        lambda_hip_scene g;
        upload_scene(*scene, g);
        render_impl(g, *params, out_rgb, progress);
    });
}

int lambda_hip_closest_hit(lambda_hip_scene *scene, const float *rays, int32_t n, lambda_hit *hits, char *err,
                           size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || (n > 0 && (!rays || !hits)) || n < 0) throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        if (n == 0) return;
        DevBuf<float> r;
        r.upload(rays, size_t(n) * 6);
        DevBuf<lambda_hit> h;
        h.alloc(size_t(n));
        DevBuf<int2> spill;
        alloc_spill(*scene, n, spill);
        launch_query_closest(grid_for(n), kBlock, nullptr, scene->ds, r.p, n, h.p, spill.p);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(hits, h.p, size_t(n) * sizeof(lambda_hit), hipMemcpyDeviceToHost));
    });
}

int lambda_hip_closest_hit_device(lambda_hip_scene *scene, const float *rays, int32_t n, lambda_hit *hits,
                                  void *stream, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || n < 0 || (n > 0 && (!rays || !hits))) throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        int dev = 0;
        HIP_OK(hipGetDevice(&dev));
        if (dev != scene->device) throw Failure{LAMBDA_EINVAL, "the scene is on another device than the current one"};
        if (scene->stack_need > kNoAllocMaxStack)
            throw Failure{LAMBDA_EINVAL, "scene traversal needs a " + std::to_string(scene->stack_need) +
                                             "-entry stack; the no-allocation query holds at most " +
                                             std::to_string(kNoAllocMaxStack)};
        if (n == 0) return;
        for (const void *p : {static_cast<const void *>(rays), static_cast<const void *>(hits)}) {
            hipPointerAttribute_t a;
            if (hipPointerGetAttributes(&a, p) != hipSuccess || a.type != hipMemoryTypeDevice || a.device != dev) {
                (void)hipGetLastError();
                throw Failure{LAMBDA_EINVAL, "rays and hits must be device memory of the current device"};
            }
        }
        int2 *spill = noalloc_scratch();
        launch_query_closest_fixed(kNoAllocThreads / kBlock, kBlock, static_cast<hipStream_t>(stream), scene->ds,
                                   rays, n, hits, spill);
        HIP_OK(hipGetLastError());
    });
}

int lambda_hip_occluded(lambda_hip_scene *scene, const float *rays, const int32_t *light_instance,
                        const int32_t *light_object, int32_t n, int32_t *occluded, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || n < 0 || (n > 0 && (!rays || !light_instance || !light_object || !occluded)))
            throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        if (n == 0) return;
        const lambda_hip_scene &g = *scene;
        for (int i = 0; i < n; i++) {
            int li = light_instance[i], lo = light_object[i];
            if (li == -1) {
                if (lo != -1) throw Failure{LAMBDA_EINVAL, "environment query needs object -1"};
            } else if (li < 0 || li >= g.ds.num_instances || lo < 0 || lo >= g.ds.num_objects) {
                throw Failure{LAMBDA_EINVAL, "light index out of range"};
            }
        }
        DevBuf<float> r;
        r.upload(rays, size_t(n) * 6);
        DevBuf<int> li, lo, out;
        li.upload(light_instance, size_t(n));
        lo.upload(light_object, size_t(n));
        out.alloc(size_t(n));
        DevBuf<int2> spill;
        alloc_spill(*scene, n, spill);
        launch_query_occluded(grid_for(n), kBlock, nullptr, scene->ds, r.p, li.p, lo.p, n, out.p, spill.p);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(occluded, out.p, size_t(n) * sizeof(int), hipMemcpyDeviceToHost));
    });
}

int lambda_hip_thinfilm_reflectance(lambda_hip_scene *scene, int32_t material, const float *queries,
                                    const float *points, int32_t n, float *out, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || n < 0 || (n > 0 && (!queries || !points || !out))) throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        if (material < 0 || material >= scene->num_materials ||
            scene->material_type[size_t(material)] != LAMBDA_MATERIAL_THINFILM)
            throw Failure{LAMBDA_EINVAL, "material is not a thin film"};
        if (n == 0) return;
        DevBuf<float> q, p, o;
        q.upload(queries, size_t(n) * 4);
        p.upload(points, size_t(n) * 3);
        o.alloc(size_t(n));
        launch_query_thinfilm(grid_for(n), kBlock, nullptr, scene->ds, material, q.p, p.p, n, o.p);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(out, o.p, size_t(n) * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int lambda_hip_guide_create(uint64_t seed, lambda_guide **out, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!out) throw Failure{LAMBDA_EINVAL, "null argument"};
        check_device();
        auto g = std::make_unique<lambda_guide>();
        g->net.init(seed);
        *out = g.release();
    });
}

void lambda_hip_guide_destroy(lambda_guide *guide) { delete guide; }

int lambda_hip_guide_get_params(lambda_guide *guide, float *params, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!guide || !params) throw Failure{LAMBDA_EINVAL, "null argument"};
        HIP_OK(hipMemcpy(params, guide->net.params.p, kGuideParams * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int lambda_hip_guide_set_params(lambda_guide *guide, const float *params, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!guide || !params) throw Failure{LAMBDA_EINVAL, "null argument"};
        guide->net.set(params, nullptr);
        guide->net.reset_adam(nullptr);
    });
}

int lambda_hip_guide_forward(lambda_guide *guide, const float *features, int32_t n, float *out, char *err,
                             size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!guide || n < 0 || (n > 0 && (!features || !out))) throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        if (n == 0) return;
        DevBuf<float> f, o;
        f.upload(features, size_t(n) * kGuideScalars);
        o.alloc(size_t(n) * kLobeOut);
        // the row count is read on the device, as in the guided passes
        DevBuf<int> cnt;
        cnt.upload(&n, 1);
        guide->net.forward_dev(f.p, cnt.p, n, o.p, nullptr);
        HIP_OK(hipMemcpy(out, o.p, size_t(n) * kLobeOut * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int lambda_hip_guide_gradient(lambda_guide *guide, const float *samples, int32_t n, float *grads, double *loss,
                              char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!guide || n < 0 || !grads || (n > 0 && !samples)) throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        GuideNet &net = guide->net;
        DevBuf<float> sm;
        if (n > 0) sm.upload(samples, size_t(n) * kGuideSample);
        HIP_OK(hipMemset(net.grads.p, 0, kGuideParams * sizeof(float)));
        HIP_OK(hipMemset(net.loss.p, 0, sizeof(float)));
        net.gradient(sm.p, n, nullptr);
        HIP_OK(hipMemcpy(grads, net.grads.p, kGuideParams * sizeof(float), hipMemcpyDeviceToHost));
        if (loss) {
            float l = 0;
            HIP_OK(hipMemcpy(&l, net.loss.p, sizeof(float), hipMemcpyDeviceToHost));
            *loss = l;
        }
        HIP_OK(hipMemset(net.grads.p, 0, kGuideParams * sizeof(float)));
    });
}

int lambda_hip_guide_train(lambda_guide *guide, const float *samples, int32_t n, int32_t batch, float lr,
                           double *loss, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!guide || n < 0 || batch <= 0 || (n > 0 && !samples) || !(lr > 0))
            throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        GuideNet &net = guide->net;
        DevBuf<float> sm;
        if (n > 0) sm.upload(samples, size_t(n) * kGuideSample);
        HIP_OK(hipMemset(net.loss.p, 0, sizeof(float)));
        net.lr = lr;
        net.train(sm.p, n, batch, nullptr);
        HIP_OK(hipDeviceSynchronize());
        if (loss) {
            float l = 0;
            HIP_OK(hipMemcpy(&l, net.loss.p, sizeof(float), hipMemcpyDeviceToHost));
            *loss = l;
        }
    });
}

int lambda_hip_guide_adam_step(lambda_guide *guide, const float *grads, float scale, float lr, char *err,
                               size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!guide || !grads || !(lr > 0) || !std::isfinite(scale)) throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        GuideNet &net = guide->net;
        HIP_OK(hipMemcpy(net.grads.p, grads, kGuideParams * sizeof(float), hipMemcpyHostToDevice));
        net.lr = lr;
        net.adam(scale, nullptr);
        HIP_OK(hipDeviceSynchronize());
    });
}

int lambda_hip_guide_get_state(lambda_guide *guide, float *m, float *v, int32_t *step, uint16_t *w_nk,
                               uint16_t *w_kn, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!guide) throw Failure{LAMBDA_EINVAL, "null argument"};
        GuideNet &net = guide->net;
        HIP_OK(hipDeviceSynchronize());
        if (m) HIP_OK(hipMemcpy(m, net.m.p, kGuideParams * sizeof(float), hipMemcpyDeviceToHost));
        if (v) HIP_OK(hipMemcpy(v, net.v.p, kGuideParams * sizeof(float), hipMemcpyDeviceToHost));
        if (step) *step = net.step;
        if (w_nk) HIP_OK(hipMemcpy(w_nk, net.w_nk.p, kGuideWeightsBf * sizeof(uint16_t), hipMemcpyDeviceToHost));
        if (w_kn) HIP_OK(hipMemcpy(w_kn, net.w_kn.p, kGuideWeightsBf * sizeof(uint16_t), hipMemcpyDeviceToHost));
    });
}

int lambda_hip_scene_info(const lambda_hip_scene *scene, lambda_scene_info *info, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || !info) throw Failure{LAMBDA_EINVAL, "null argument"};
        std::memset(info, 0, sizeof(*info));
        info->stack_need = scene->stack_need;
        info->stack_lds = kRing;
        info->wide_nodes = int32_t(scene->wide.n);
        info->num_objects = scene->ds.num_objects;
        info->num_instances = scene->ds.num_instances;
        info->num_nodes = scene->ds.num_nodes;
        for (int k = 0; k < 6; k++) info->bounds[k] = scene->box[k];
    });
}

int lambda_hip_server_create(lambda_hip_scene *const *scenes, int32_t num_scenes, int64_t hbm_budget,
                             int32_t max_launches, lambda_server **out, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scenes || num_scenes < 1 || !out || hbm_budget < 0 || max_launches < 0 || max_launches > 8)
            throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        *out = nullptr;
        check_device();
        auto srv = std::make_unique<lambda_server>();
        srv->budget = hbm_budget;
        server_setup(*srv, scenes, num_scenes, max_launches > 0 ? max_launches : 2);
        lambda_server *p = srv.get();
        srv->worker = std::thread([p] { p->run(); });
        *out = srv.release();
    });
}

int lambda_hip_server_submit(lambda_server *srv, const lambda_job *job, int64_t *ticket, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!srv || !job || !ticket) throw Failure{LAMBDA_EINVAL, "null argument"};
        if (job->scene < 0 || job->scene >= int(srv->table.size())) throw Failure{LAMBDA_EINVAL, "job scene out of range"};
        if (!job->out_rgb) throw Failure{LAMBDA_EINVAL, "job without output buffer"};
        lambda_render_params p = job->params;
        p.rng_mode = LAMBDA_RNG_REFERENCE;
        validate_params(p);
        std::lock_guard<std::mutex> g(srv->mu);
        if (!srv->fatal.empty()) throw Failure{LAMBDA_EHIP, "server stopped: " + srv->fatal};
        if (srv->pending.size() >= 1u << 20) throw Failure{LAMBDA_EBUSY, "job queue full"};
        lambda_server::Job jb;
        jb.job = *job;
        jb.job.params = p;
        jb.ticket = srv->next_ticket++;
        jb.submit_ms = srv->now_ms();
        jb.state = LAMBDA_JOB_QUEUED;
        *ticket = jb.ticket;
        srv->jobs.emplace(jb.ticket, jb);
        srv->pending.push_back(jb.ticket);
        srv->cv.notify_all();
    });
}

int lambda_hip_server_poll(lambda_server *srv, int64_t ticket, int32_t *state, lambda_job_result *result, char *err,
                           size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!srv || !state) throw Failure{LAMBDA_EINVAL, "null argument"};
        std::lock_guard<std::mutex> g(srv->mu);
        auto it = srv->jobs.find(ticket);
        if (it == srv->jobs.end()) throw Failure{LAMBDA_EINVAL, "unknown ticket"};
        *state = it->second.state;
        if (it->second.state == LAMBDA_JOB_FAILED) throw Failure{LAMBDA_EHIP, it->second.error};
        if (result && it->second.state == LAMBDA_JOB_DONE) *result = it->second.result;
    });
}

int lambda_hip_server_drain(lambda_server *srv, lambda_serve_stats *stats, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!srv) throw Failure{LAMBDA_EINVAL, "null argument"};
        std::unique_lock<std::mutex> g(srv->mu);
        srv->cv.wait(g, [&] {
            if (!srv->fatal.empty()) return true;
            if (!srv->pending.empty()) return false;
            for (auto &sl : srv->slots)
                if (sl->busy) return false;
            return true;
        });
        if (!srv->fatal.empty()) throw Failure{LAMBDA_EHIP, "server stopped: " + srv->fatal};
        if (stats) {
            stats->batches = srv->launches;
            stats->paths = srv->paths;
            stats->closest_rays = srv->closest;
            stats->shadow_rays = srv->shadow;
            double first = -1;
            for (auto &kv : srv->jobs) first = first < 0 ? kv.second.submit_ms : std::min(first, kv.second.submit_ms);
            stats->wall_ms = first < 0 ? 0 : srv->last_finish_ms - first;
        }
    });
}

void lambda_hip_server_destroy(lambda_server *srv) { delete srv; }

int lambda_hip_serve(lambda_hip_scene *const *scenes, int32_t num_scenes, const lambda_job *jobs, int32_t num_jobs,
                     int64_t hbm_budget, lambda_job_result *results, lambda_serve_stats *stats, char *err,
                     size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scenes || num_scenes < 1 || num_jobs < 0 || (num_jobs > 0 && (!jobs || !results)))
            throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        check_device();
        const auto wall0 = std::chrono::steady_clock::now();
        std::vector<DeviceScene> table(static_cast<size_t>(num_scenes));
        int need = 0;
        for (int k = 0; k < num_scenes; k++) {
            if (!scenes[k]) throw Failure{LAMBDA_EINVAL, "null scene"};
            table[size_t(k)] = scenes[k]->ds;
            need = std::max(need, scenes[k]->stack_need);
        }
        // jobs in arrival order
        std::vector<int> order(static_cast<size_t>(num_jobs));
        for (int i = 0; i < num_jobs; i++) {
            const lambda_job &jb = jobs[i];
            if (jb.scene < 0 || jb.scene >= num_scenes) throw Failure{LAMBDA_EINVAL, "job scene out of range"};
            if (!jb.out_rgb) throw Failure{LAMBDA_EINVAL, "job without output buffer"};
            if (!(jb.arrival_ms >= 0 && jb.arrival_ms <= 600000)) throw Failure{LAMBDA_EINVAL, "job arrival out of range"};
            lambda_render_params p = jb.params;
            p.rng_mode = LAMBDA_RNG_REFERENCE;
            validate_params(p);
            order[size_t(i)] = i;
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return jobs[a].arrival_ms < jobs[b].arrival_ms; });
        int dev = 0, rate_khz = 0;
        HIP_OK(hipGetDevice(&dev));
        HIP_OK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
        const double ticks_per_ms = double(std::max(rate_khz, 1));
        DevBuf<DeviceScene> dscenes;
        dscenes.upload(table);
        const int threads = int(resident_persist_threads());
        RenderBuffers B;
        B.o.alloc(size_t(threads)); B.d.alloc(size_t(threads)); B.lam.alloc(size_t(threads));
        B.beta.alloc(size_t(threads)); B.L.alloc(size_t(threads)); B.rng.alloc(size_t(threads));
        B.info.alloc(size_t(threads)); B.hit.alloc(size_t(threads));
        B.ctr.alloc(kCtrCount);
        B.rays.alloc(2);
        if (need > kRing) B.spill.alloc(size_t(need - kRing) * size_t(threads));
        Paths P{};
        P.o = B.o.p; P.d = B.d.p; P.lam = B.lam.p; P.beta = B.beta.p; P.L = B.L.p;
        P.rng = B.rng.p; P.info = B.info.p; P.hit = B.hit.p; P.ctr = B.ctr.p; P.rays = B.rays.p;
        DevBuf<ServeCtl> ctl;
        ctl.alloc(1);
        hipStream_t st = nullptr;
        HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamFree {
            hipStream_t s;
            ~StreamFree() { if (s) (void)hipStreamDestroy(s); }
        } sfree{st};
        launch_fill_i(1, kBlock, st, P.ctr, kCtrCount);
        launch_fill_u64(1, kBlock, st, P.rays, 2);
        int batches = 0;
        double first_start_ms = -1;
        size_t next = 0;
        while (next < order.size()) {
            // admission: jobs in arrival order while their buffers fit the budget
            std::vector<int> batch;
            long long bytes = 0;
            while (next < order.size()) {
                const lambda_job &jb = jobs[order[next]];
                const long long px = (long long)jb.params.width * jb.params.height;
                const long long jbytes = px * (3 * 8 + 3 * 4);
                if (!batch.empty() && hbm_budget > 0 && bytes + jbytes > hbm_budget) break;
                bytes += jbytes;
                batch.push_back(order[next++]);
            }
            const double now_ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
            if (first_start_ms < 0) first_start_ms = now_ms;
            const double batch_start = now_ms - first_start_ms;  // relative to the first batch
            std::vector<ServeJob> sj(batch.size());
            long long units = 0, px_total = 0;
            for (size_t k = 0; k < batch.size(); k++) px_total += (long long)jobs[batch[k]].params.width * jobs[batch[k]].params.height;
            DevBuf<double> fb;
            DevBuf<float> rgb;
            fb.alloc(size_t(px_total) * 3);
            rgb.alloc(size_t(px_total) * 3);
            DevBuf<int> done;
            DevBuf<unsigned long long> finish;
            done.alloc(batch.size());
            finish.alloc(batch.size());
            launch_fill_d(grid_for(px_total * 3), kBlock, st, fb.p, size_t(px_total) * 3);
            launch_fill_i(grid_for((long long)batch.size()), kBlock, st, done.p, int(batch.size()));
            long long off = 0;
            for (size_t k = 0; k < batch.size(); k++) {
                const lambda_job &jb = jobs[batch[k]];
                lambda_render_params p = jb.params;
                p.rng_mode = LAMBDA_RNG_REFERENCE;
                ServeJob &J = sj[k];
                J.F = make_frame(p);
                J.fb = fb.p + off * 3;
                J.first_unit = units;
                J.arrival = (unsigned long long)(std::max(0.0, jb.arrival_ms - batch_start) * ticks_per_ms);
                J.scene = jb.scene;
                J.units = J.F.npix;
                units += J.units;
                off += J.F.npix;
            }
            if (units >= (1LL << 31) - (1LL << 28)) throw Failure{LAMBDA_EINVAL, "batch too large"};
            DevBuf<ServeJob> djobs;
            djobs.upload(sj);
            launch_serve_clock(1, 1, st, ctl.p);
            launch_serve(threads / kBlock, kBlock, st, dscenes.p, djobs.p, int(sj.size()), units, P, ctl.p, done.p,
                                                        finish.p, B.spill.p);
            HIP_OK(hipGetLastError());
            off = 0;
            for (size_t k = 0; k < batch.size(); k++) {
                const ServeJob &J = sj[k];
                launch_resolve(grid_for(J.F.npix), kBlock, st, J.F, J.fb, rgb.p + off * 3);
                off += J.F.npix;
            }
            std::vector<float> host_rgb(size_t(px_total) * 3);
            std::vector<unsigned long long> fin(batch.size());
            ServeCtl hc;
            HIP_OK(hipMemcpyAsync(host_rgb.data(), rgb.p, host_rgb.size() * sizeof(float), hipMemcpyDeviceToHost, st));
            HIP_OK(hipMemcpyAsync(fin.data(), finish.p, fin.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
            HIP_OK(hipMemcpyAsync(&hc, ctl.p, sizeof hc, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            int herr = 0;
            HIP_OK(hipMemcpy(&herr, P.ctr + kCtrError, sizeof(int), hipMemcpyDeviceToHost));
            if (herr) throw Failure{LAMBDA_EHIP, "device reported an inconsistent hit record"};
            off = 0;
            for (size_t k = 0; k < batch.size(); k++) {
                const lambda_job &jb = jobs[batch[k]];
                const ServeJob &J = sj[k];
                std::memcpy(jb.out_rgb, host_rgb.data() + off * 3, size_t(J.F.npix) * 3 * sizeof(float));
                off += J.F.npix;
                lambda_job_result &r = results[batch[k]];
                const double fin_ms = double(fin[k] - hc.t0) / ticks_per_ms;  // after this batch's start
                r.finish_ms = batch_start + fin_ms;
                r.latency_ms = r.finish_ms - jb.arrival_ms;
                r.batch = batches;
            }
            batches++;
        }
        if (stats) {
            unsigned long long r[2] = {0, 0};
            HIP_OK(hipMemcpy(r, P.rays, sizeof r, hipMemcpyDeviceToHost));
            stats->batches = batches;
            stats->paths = 0;
            for (int i = 0; i < num_jobs; i++)
                stats->paths += (long long)jobs[i].params.width * jobs[i].params.height * jobs[i].params.samples;
            stats->closest_rays = (long long)r[0];
            stats->shadow_rays = (long long)r[1];
            stats->wall_ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count() -
                std::max(0.0, first_start_ms);
        }
    });
}

namespace {

// Path slots and per-slot guide buffers of a session, for the kernel that
// renders its steps (path_slots > 0: the caller's count).
void size_session(lambda_hip_progressive &pr, int path_slots) {
    const bool guided = pr.net != nullptr;
    long long slots = path_slots > 0                  ? path_slots
                      : guided && pr.table_cells > 0  ? resident_threads(persist_table_blocks_per_cu())
                      : pr.fused                      ? resident_threads(persist_guided_blocks_per_cu())
                      : guided                        ? (1LL << 20)
                                                      : resident_persist_threads();
    slots = std::max<long long>(1, std::min<long long>(slots, 1LL << 26));
    pr.R.reset();
    pr.R = std::make_unique<Renderer>(*pr.scene, make_frame(pr.params), int(slots));
    if (guided) {
        const size_t ns = size_t(slots);
        pr.q_slot.alloc(ns);
        pr.q_feat.alloc(ns * kGuideScalars);
        pr.q_out.alloc(ns * kLobeOut);
        pr.grec.alloc(ns * size_t(pr.params.depth + 1) * 3);
        pr.Q = GuideQueue{pr.q_slot.p, pr.q_feat.p, pr.q_out.p};
        pr.R->P.grec = pr.grec.p;
    }
}

}  // namespace

int lambda_hip_progressive_create(lambda_hip_scene *scene, const lambda_render_params *params, lambda_guide *guide,
                                  lambda_hip_progressive **out, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!scene || !params || !out) throw Failure{LAMBDA_EINVAL, "null argument"};
        lambda_render_params p = *params;
        p.rng_mode = LAMBDA_RNG_THROUGHPUT;
        p.samples = 1;
        p.sqrt_samples = 1;
        validate_params(p);
        auto pr = std::make_unique<lambda_hip_progressive>();
        pr->scene = scene;
        pr->params = p;
        pr->net = guide ? &guide->net : nullptr;
        // Throughput-based roulette works against a guide: after a guided
        // bounce toward bright directions the throughput is small because
        // the sampling density there is high, so roulette would end most of
        // exactly those paths and weight the survivors up to 20x.  Guided
        // sessions therefore run every path to the depth (or to a miss or an
        // emitter) unless the caller sets a start.
        if (guide) pr->rr_start = p.depth + 1;
        // guided: the network inside the persistent kernel unless the
        // caller asks for one launch per bounce (tail_paths -1)
        pr->fused = guide && params->tail_paths != -1;
        size_session(*pr, params->path_slots);
        if (guide) pr->train.alloc(size_t(pr->train_cap) * kGuideSample);
        *out = pr.release();
    });
}

int lambda_hip_progressive_step(lambda_hip_progressive *pr, int32_t samples, int32_t flags,
                                lambda_progressive_stats *stats, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!pr || samples < 1) throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        int root = int(std::lround(std::sqrt(double(samples))));
        if (root * root != samples) throw Failure{LAMBDA_EINVAL, "samples per step must be a perfect square"};
        Renderer &R = *pr->R;
        const bool train = (flags & LAMBDA_STEP_TRAIN) != 0;
        const Frame F = progressive_frame(*pr, samples, root, train);
        if ((long long)F.npix * samples >= (1LL << 31) - (1LL << 28))
            throw Failure{LAMBDA_EINVAL, "too many samples in one step"};
        EventPair ev;
        HIP_OK(hipEventCreate(&ev.a));
        HIP_OK(hipEventCreate(&ev.b));
        const bool keep = (flags & LAMBDA_STEP_DISCARD) == 0;
        double *fb = R.P.fb;
        if (!keep) {
            if (!pr->scratch_fb.p) pr->scratch_fb.alloc(size_t(F.npix) * 3);
            R.P.fb = pr->scratch_fb.p;
        }
        const bool use_table = pr->net && pr->table_cells > 0;
        if (use_table && pr->table_stale) {
            const float *b = pr->scene->box;
            pr->net->tabulate(pr->table, make_float3(b[0], b[1], b[2]), make_float3(b[3], b[4], b[5]),
                              pr->table_feat.p, pr->table_rows.p, pr->table_entries, R.st);
            pr->table_stale = false;
        }
        HIP_OK(hipEventRecord(ev.a, R.st));
        const bool clear = !keep || pr->total == 0;
        lambda_render_stats rs{};
        if (use_table)
            R.run(F, kAutoTailPaths, nullptr, &rs, clear, nullptr, nullptr, false, &pr->table);
        else if (pr->net && pr->fused)
            R.run(F, kAutoTailPaths, nullptr, &rs, clear, pr->net, &pr->Q, true);
        else if (pr->net && pr->guide_bounces <= pr->params.depth)
            R.run_guided_batches(F, nullptr, &rs, clear, pr->net, &pr->Q);
        else
            R.run(F, pr->net ? 0 : kAutoTailPaths, nullptr, &rs, clear, pr->net, pr->net ? &pr->Q : nullptr);
        HIP_OK(hipEventRecord(ev.b, R.st));
        R.P.fb = fb;
        HIP_OK(hipEventSynchronize(ev.b));
        float render_ms = 0;
        HIP_OK(hipEventElapsedTime(&render_ms, ev.a, ev.b));
        const long long collected = pr->net && train ? std::min<long long>(R.host_ctr[kCtrTrain], pr->train_cap) : 0;
        if (pr->net && train) pr->seen_per_sample = double(R.host_ctr[kCtrTrainSeen]) / samples;
        float train_ms = 0;
        double loss = 0;
        if (pr->net && train && collected >= pr->batch) {
            GuideNet &net = *pr->net;
            HIP_OK(hipMemsetAsync(net.loss.p, 0, sizeof(float), R.st));
            HIP_OK(hipEventRecord(ev.a, R.st));
            net.train(pr->train.p, int(collected), pr->batch, R.st, kWeightCap, pr->epochs);
            pr->table_stale = true;
            HIP_OK(hipEventRecord(ev.b, R.st));
            HIP_OK(hipEventSynchronize(ev.b));
            HIP_OK(hipEventElapsedTime(&train_ms, ev.a, ev.b));
            float lsum = 0;
            HIP_OK(hipMemcpy(&lsum, net.loss.p, sizeof(float), hipMemcpyDeviceToHost));
            loss = lsum;
            loss /= double(collected / pr->batch * pr->batch) * pr->epochs;
        }
        if (keep) pr->total += samples;
        pr->steps++;
        if (stats) {
            stats->samples = pr->total;
            stats->train_samples = collected;
            stats->render_ms = render_ms;
            stats->train_ms = train_ms;
            stats->loss = loss;
            stats->closest_rays = rs.closest_rays;
            stats->shadow_rays = rs.shadow_rays;
        }
    });
}

int lambda_hip_progressive_image(lambda_hip_progressive *pr, float *out_rgb, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!pr || !out_rgb) throw Failure{LAMBDA_EINVAL, "null argument"};
        if (pr->total == 0) throw Failure{LAMBDA_EINVAL, "no samples rendered yet"};
        Renderer &R = *pr->R;
        Frame F = make_frame(pr->params);
        F.spp = int(std::min<long long>(pr->total, 1LL << 30));
        launch_resolve(grid_for(F.npix), kBlock, R.st, F, R.P.fb, R.B.rgb.p);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(out_rgb, R.B.rgb.p, size_t(F.npix) * 3 * sizeof(float), hipMemcpyDeviceToHost, R.st));
        HIP_OK(hipStreamSynchronize(R.st));
    });
}

int lambda_hip_progressive_set_training(lambda_hip_progressive *pr, int32_t epochs, int32_t batch, float lr,
                                        int32_t guided_bounces, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!pr || epochs < 1 || batch < kGuideRowsPerBlock || batch > pr->train_cap || !(lr > 0))
            throw Failure{LAMBDA_EINVAL, "invalid training settings"};
        pr->epochs = epochs;
        pr->batch = batch;
        if (guided_bounces < 0) throw Failure{LAMBDA_EINVAL, "invalid training settings"};
        pr->guide_bounces = guided_bounces > 0 ? guided_bounces : kDefaultGuidedBounces;
        if (pr->net) pr->net->lr = lr;
    });
}

int lambda_hip_progressive_set_guide_table(lambda_hip_progressive *pr, int32_t cells, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!pr) throw Failure{LAMBDA_EINVAL, "null argument"};
        if (!pr->net) throw Failure{LAMBDA_EINVAL, "the session has no guide"};
        if (cells < 0 || cells > kMaxTableCells) throw Failure{LAMBDA_EINVAL, "invalid table size"};
        if (cells > 0 && !pr->fused) throw Failure{LAMBDA_EINVAL, "a table needs the persistent schedule"};
        GuideTableDev T{};
        long long entries = 0;
        if (cells > 0) {
            const float *b = pr->scene->box;
            float ext[3], longest = 0;
            for (int a = 0; a < 3; a++) {
                ext[a] = std::max(0.0f, b[3 + a] - b[a]);
                longest = std::max(longest, ext[a]);
            }
            const float cell = longest > 0 ? longest / float(cells) : 1.0f;
            entries = kTableNormals;
            for (int a = 0; a < 3; a++) {
                T.lo[a] = b[a];
                T.n[a] = std::max(1, int(std::ceil(ext[a] / cell)));
                entries *= T.n[a];
            }
            if (entries > kMaxTableEntries) throw Failure{LAMBDA_EINVAL, "table too large"};
            T.inv_cell = 1 / cell;
            T.lambda = 0.5f * (pr->params.lambda_min + pr->params.lambda_max);
        }
        const bool resize = (cells > 0) != (pr->table_cells > 0);
        if (resize && pr->steps > 0) throw Failure{LAMBDA_EINVAL, "switch to or from a table before the first step"};
        pr->table_rows.release();
        pr->table_feat.release();
        pr->table_cells = cells;
        pr->table_entries = entries;
        if (cells > 0) {
            pr->table_rows.alloc(size_t(entries) * kLobeOut);
            pr->table_feat.alloc(size_t(entries) * kGuideScalars);
            T.rows = pr->table_rows.p;
        }
        pr->table = T;
        pr->table_stale = true;
        // the table kernel holds more paths in flight than the network's
        if (resize) size_session(*pr, pr->params.path_slots);
    });
}

int lambda_hip_progressive_set_guide_mix(lambda_hip_progressive *pr, float frac, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!pr) throw Failure{LAMBDA_EINVAL, "null argument"};
        if (!pr->net) throw Failure{LAMBDA_EINVAL, "the session has no guide"};
        if (!(frac > 0 && frac <= kGuideFracMax)) throw Failure{LAMBDA_EINVAL, "invalid guide fraction"};
        pr->guide_frac = frac;
    });
}

int lambda_hip_progressive_set_roulette(lambda_hip_progressive *pr, int32_t start_bounce, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!pr) throw Failure{LAMBDA_EINVAL, "null argument"};
        if (start_bounce < 1) throw Failure{LAMBDA_EINVAL, "invalid roulette start"};
        pr->rr_start = start_bounce;
    });
}

void lambda_hip_progressive_destroy(lambda_hip_progressive *pr) { delete pr; }

int lambda_hip_guide_benchmark(lambda_guide *guide, int32_t rows, int32_t iters, double *forward_ms,
                               double *gradient_ms, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        if (!guide || rows < kGuideRowsPerBlock || iters < 1) throw Failure{LAMBDA_EINVAL, "invalid arguments"};
        GuideNet &net = guide->net;
        // device-resident samples: features on a lattice, unit directions, c = 1
        std::vector<float> host(size_t(rows) * kGuideSample);
        Pcg rng = pcg_seed(12345, 3);
        for (int r = 0; r < rows; r++) {
            float *o = &host[size_t(r) * kGuideSample];
            for (int k = 0; k < kGuideScalars; k++) o[k] = rng.next();
            V3 w = normalize(v3(rng.next() - 0.5f, rng.next() - 0.5f, rng.next() - 0.5f + 1e-3f));
            o[6] = w.x; o[7] = w.y; o[8] = w.z; o[9] = 1.0f;
        }
        DevBuf<float> sm, out;
        sm.upload(host);
        out.alloc(size_t(rows) * kLobeOut);
        EventPair ev;
        HIP_OK(hipEventCreate(&ev.a));
        HIP_OK(hipEventCreate(&ev.b));
        net.forward(sm.p, kGuideSample, rows, out.p, kLobeOut, nullptr);
        HIP_OK(hipEventRecord(ev.a, nullptr));
        for (int i = 0; i < iters; i++) net.forward(sm.p, kGuideSample, rows, out.p, kLobeOut, nullptr);
        HIP_OK(hipEventRecord(ev.b, nullptr));
        HIP_OK(hipEventSynchronize(ev.b));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, ev.a, ev.b));
        if (forward_ms) *forward_ms = ms / iters;
        net.gradient(sm.p, rows, nullptr);
        HIP_OK(hipEventRecord(ev.a, nullptr));
        for (int i = 0; i < iters; i++) net.gradient(sm.p, rows, nullptr);
        HIP_OK(hipEventRecord(ev.b, nullptr));
        HIP_OK(hipEventSynchronize(ev.b));
        HIP_OK(hipEventElapsedTime(&ms, ev.a, ev.b));
        if (gradient_ms) *gradient_ms = ms / iters;
        HIP_OK(hipMemset(net.grads.p, 0, kGuideParams * sizeof(float)));
    });
}

int lambda_hip_device_info(char *name, size_t name_len, int32_t *compute_units, char *err, size_t err_len) {
    return guarded(err, err_len, [&] {
        check_device();
        int dev = 0;
        HIP_OK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIP_OK(hipGetDeviceProperties(&prop, dev));
        std::string label = prop.name[0] ? std::string(prop.name) + " (" + prop.gcnArchName + ")" : prop.gcnArchName;
        if (name && name_len) set_err(name, name_len, label);
        if (compute_units) *compute_units = prop.multiProcessorCount;
    });
}

}  // extern "C"
