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
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "lambda_hip.h"
#include "dscene.h"
#include "guide.h"
#include "scene_prep.h"

namespace lam {
#include "spectral_tables.inc"
}

using namespace lam;

namespace {

constexpr int kBlock = 256;

// ------------------------------------------------------------------ errors

struct Failure {
    int code;
    std::string msg;
};

void set_err(char *err, size_t len, const std::string &msg) {
    if (!err || len == 0) return;
    size_t n = std::min(msg.size(), len - 1);
    std::memcpy(err, msg.data(), n);
    err[n] = 0;
}

#define HIP_OK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw Failure{e_ == hipErrorOutOfMemory ? LAMBDA_ENOMEM : LAMBDA_EHIP,            \
                          std::string(#call) + ": " + hipGetErrorString(e_)};                 \
    } while (0)

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return;
        HIP_OK(hipMalloc(&p, count * sizeof(T)));
    }
    // alloc unless it already holds exactly count elements
    void ensure(size_t count) {
        if (!p || n != count) alloc(count);
    }
    void upload(const T *src, size_t count) {
        alloc(count);
        if (count) HIP_OK(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
    }
    void upload(const std::vector<T> &v) { upload(v.data(), v.size()); }
};

#include "guide_mlp.h"

// --------------------------------------------------------- wave utilities

// Wave-aggregated slot reservation: one atomicAdd per wave; every lane of
// the wave must call it (pred false for lanes with nothing to push).
__device__ __forceinline__ int wave_push(int *counter, bool pred) {
    const unsigned long long mask = __ballot(pred);
    if (mask == 0) return -1;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
    return pred ? base + rank : -1;
}

// --------------------------------------------------------------- state

enum : int {
    kFlagSpecular = 1,   // previous vertex was specular (or the camera)
    kFlagFinished = 2,   // the sample's path has ended
    kFlagGuidePending = 4,  // a guided vertex awaits the radiance base of its continuation
};

enum : int {
    kCtrNext = 0,        // next closest-hit queue length
    kCtrShadow = 1,      // occlusion queue length
    kCtrWork = 2,        // next work item (throughput mode)
    kCtrPixels = 3,      // finished pixels (reference mode) / samples (throughput)
    kCtrError = 4,       // error bits
    kCtrDone = 5,        // samples ended by the last shade pass
    kCtrGuided = 6,      // paths queued for the guiding network this pass
    kCtrTrain = 7,       // training samples written since the last training
    kCtrTrainSeen = 8,   // training candidates met (kept or not)
    kCtrCount = 16
};

struct Paths {
    float4 *o;        // origin, previous scatter pdf
    float4 *d;        // direction
    float4 *lam;      // hero wavelengths
    float4 *beta;     // throughput
    float4 *L;        // radiance
    uint4 *rng;       // PCG32 state and increment (lo, hi, lo, hi)
    int4 *info;       // pixel, sample index, bounce, flags
    int2 *hit;        // instance, object (-1 on a miss)
    int *q_in, *q_out;
    int *done;                // slots whose sample ended in the last shade pass
    float4 *s_o, *s_d, *s_c;  // shadow: origin+tmax, direction, contribution
    int4 *s_i;                // shadow: slot, light instance, light object
    int *ctr;
    unsigned long long *rays; // rays traced by the persistent kernel: closest, occlusion
    double *fb;               // XYZ sums, [pixel][3]
    const int *order;         // pixel visiting order (8x8 tiles)
    float4 *grec;             // guided vertices, [slot][bounce][3] (guided training steps only)
};

struct Frame {
    int width, height, depth, spp, sqrt_spp, mode, npix;
    float lmin, lmax, lrange;
    uint64_t seed;
    long long total_work;     // npix * spp (throughput mode)
    // Reference mode: this launch runs samples [sample_begin, sample_end) of
    // each pixel (sample_end 0 = spp); a pixel that stops before spp saves
    // its PCG32 state in rng_save, where a later launch resumes it.
    int sample_begin, sample_end;
    uint4 *rng_save;
    int *cost;                // when set (persistent kernel): per-pixel device-clock ticks
    unsigned long long *unit_t0;  // with cost: start time of each slot's unit
    float *train;             // when set: training samples [train_cap][10]
    int train_cap;
    float train_keep;         // probability of keeping a training candidate
    int grec_stride;          // guided vertices per slot (depth + 1)
    int guide_bounces;        // vertices with bounce index below this are guided
    float3 box_lo, box_hi;    // scene box for the guiding features
};

// New camera sample `sample` of `pixel` into slot s (render/renderer.h:71-82,
// render/camera.h:13-16).  Consumes three draws.
__device__ __forceinline__ void camera_sample(const DeviceScene &S, const Frame &F, const Paths &P, int s, int pixel,
                                              int sample, Pcg &rng) {
    const int px = pixel % F.width, py = pixel / F.width;
    const int i = sample / F.sqrt_spp, j = sample % F.sqrt_spp;
    float u = (float(px) + (float(i) + rng.next()) / float(F.sqrt_spp)) / float(F.width);
    float v = (float(py) + (float(j) + rng.next()) / float(F.sqrt_spp)) / float(F.height);
    float lambda = F.lmin + rng.next() * (F.lmax - F.lmin);
    float4 lam;
    lam.x = F.lmin + fmodf(lambda - F.lmin + float(0) * F.lrange / kHero, F.lrange);
    lam.y = F.lmin + fmodf(lambda - F.lmin + float(1) * F.lrange / kHero, F.lrange);
    lam.z = F.lmin + fmodf(lambda - F.lmin + float(2) * F.lrange / kHero, F.lrange);
    lam.w = F.lmin + fmodf(lambda - F.lmin + float(3) * F.lrange / kHero, F.lrange);
    V3 pos = v3(S.cam_pos), corner = v3(S.cam_corner), ch = v3(S.cam_h), cv = v3(S.cam_v);
    V3 dir = normalize(corner + ch * u + cv * v - pos);
    P.o[s] = make_float4(pos.x, pos.y, pos.z, 1.0f);
    P.d[s] = make_float4(dir.x, dir.y, dir.z, 0.0f);
    P.lam[s] = lam;
    P.beta[s] = make_float4(1, 1, 1, 1);
    P.L[s] = make_float4(0, 0, 0, 0);
    P.info[s] = make_int4(pixel, sample, 0, kFlagSpecular);
}

__device__ __forceinline__ void store_rng(const Paths &P, int s, const Pcg &r) {
    P.rng[s] = make_uint4(uint32_t(r.state), uint32_t(r.state >> 32), uint32_t(r.inc), uint32_t(r.inc >> 32));
}
__device__ __forceinline__ Pcg load_rng(const Paths &P, int s) {
    uint4 v = P.rng[s];
    Pcg r;
    r.state = uint64_t(v.x) | (uint64_t(v.y) << 32);
    r.inc = uint64_t(v.z) | (uint64_t(v.w) << 32);
    return r;
}

// ---------------------------------------------------------------- kernels

// Traversal stacks: kRing entries per lane in LDS, deeper entries in a
// global area of the launch, [entry][thread] (allocated by the host when the
// scene's worst-case depth exceeds kRing).
constexpr int kRing = 16;
// Waves per SIMD the persistent kernel is compiled for (its registers are
// capped so that this many fit; measured best on gfx950).
constexpr int kPersistWaves = 3;
// Automatic schedule: live paths below which a render switches to the
// persistent schedule (above any slot count: persistent from the first pass
// on).  Slots default to the number of persistent threads resident at once.
constexpr long long kAutoTailPaths = 1LL << 40;
using Stack = TraversalStack<kRing>;

__device__ __forceinline__ Stack make_stack(int *s_code, float *s_dist, int2 *spill) {
    Stack s;
    s.lcode = s_code + threadIdx.x;
    s.ldist = s_dist + threadIdx.x;
    s.stride = kBlock;
    const size_t threads = size_t(gridDim.x) * blockDim.x;
    s.spill = spill ? spill + (size_t(blockIdx.x) * blockDim.x + threadIdx.x) : nullptr;
    s.spill_stride = threads;
    return s;
}

template <bool kAny>
__device__ __forceinline__ bool trace_ray(const DeviceScene &S, V3 o, V3 d, float &best, int &hit_inst, int &hit_obj,
                                         int skip_inst, int skip_obj, Stack &stk) {
    return lam::traverse<kAny, kRing>(S, o, d, best, hit_inst, hit_obj, skip_inst, skip_obj, stk);
}

// Reference mode: the pixel of rank `rank` in the visiting order starts
// sample F.sample_begin in slot s, with PCG32 stream (seed, pixel index)
// (render/renderer.h:47-49) from its start or from the state a previous
// launch saved; the slot keeps the pixel for the launch's samples.
__device__ __forceinline__ void start_pixel(const DeviceScene &S, const Frame &F, const Paths &P, int s, int rank) {
    const int pixel = P.order[rank];
    Pcg rng;
    if (F.sample_begin > 0) {
        const uint4 v = F.rng_save[pixel];
        rng.state = uint64_t(v.x) | (uint64_t(v.y) << 32);
        rng.inc = uint64_t(v.z) | (uint64_t(v.w) << 32);
    } else {
        rng = pcg_seed(F.seed, uint64_t(pixel));
    }
    camera_sample(S, F, P, s, pixel, F.sample_begin, rng);
    store_rng(P, s, rng);
}

// Throughput mode: work item w = sample * npix + rank starts in slot s with
// its own PCG32 stream (seed, w).
__device__ __forceinline__ void start_work(const DeviceScene &S, const Frame &F, const Paths &P, int s, long long w) {
    int rank = int(w % F.npix);
    int sample = int(w / F.npix);
    int pixel = P.order[rank];
    Pcg rng = pcg_seed(F.seed, uint64_t(w));
    camera_sample(S, F, P, s, pixel, sample, rng);
    store_rng(P, s, rng);
}

// Work units: pixels (reference mode) or (pixel, sample) items.
__device__ __forceinline__ long long work_units(const Frame &F) {
    return F.mode == LAMBDA_RNG_REFERENCE ? F.npix : F.total_work;
}

__device__ __forceinline__ void start_unit(const DeviceScene &S, const Frame &F, const Paths &P, int s, long long w) {
    if (F.mode == LAMBDA_RNG_REFERENCE) start_pixel(S, F, P, s, int(w));
    else start_work(S, F, P, s, w);
}

// First work unit of every slot: slot s takes unit base + s.
__global__ void __launch_bounds__(kBlock) k_start(DeviceScene S, Frame F, Paths P, int nslots, long long base) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    start_unit(S, F, P, s, base + s);
    P.q_in[s] = s;
}

// Threads [0, n_closest) find the closest hit of queued path q_in[t];
// threads [n_closest, n_closest + n_shadow) test queued occlusion ray
// t - n_closest and, if the light is visible, add its contribution to the
// path's radiance.
// Queue lengths of a pass taken on the device (graph-captured passes):
// {closest, shadow, done}; a kernel given counts reads its lengths there.
__global__ void k_pass_begin(const int *ctr, int *counts) {
    counts[0] = ctr[kCtrNext];
    counts[1] = ctr[kCtrShadow];
    counts[2] = ctr[kCtrDone];
}

__global__ void __launch_bounds__(kBlock) k_trace(DeviceScene S, Paths P, int n_closest, int n_shadow,
                                                  int2 *spill, const int *counts) {
    __shared__ int s_code[kRing * kBlock];
    __shared__ float s_dist[kRing * kBlock];
    if (counts) {
        n_closest = counts[0];
        n_shadow = counts[1];
    }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) P.ctr[kCtrNext] = 0;
    Stack stk = make_stack(s_code, s_dist, spill);
    if (t < n_closest) {
        const int s = P.q_in[t];
        float4 o = P.o[s], d = P.d[s];
        float best = kMax;
        int hi = -1, ho = -1;
        trace_ray<false>(S, v3(o), v3(d), best, hi, ho, -1, -2, stk);
        P.hit[s] = make_int2(hi, ho);
    } else if (t < n_closest + n_shadow) {
        const int q = t - n_closest;
        float4 so = P.s_o[q], sd = P.s_d[q];
        int4 si = P.s_i[q];
        float best = so.w;
        int hi = -1, ho = -1;
        bool blocked = trace_ray<true>(S, v3(so), v3(sd), best, hi, ho, si.y, si.z, stk);
        if (!blocked) {
            float4 c = P.s_c[q];
            float4 L = P.L[si.x];
            L.x += c.x;
            L.y += c.y;
            L.z += c.z;
            L.w += c.w;
            P.L[si.x] = L;
        }
    }
}

__device__ __forceinline__ S4 lambertian_albedo(const DeviceScene &S, int4 mt, const Hit &h, const S4 &lam) {
    return spectrum_texture4(S, mt.y, h, lam);
}

// Occlusion ray queued by a shading step: origin and light distance,
// direction, contribution, and {slot, light instance, light object}.
struct ShadowRay {
    float4 o, d, c;
    int4 i;
};

// Guided vertex record: {features 0-3}, {features 4-5, w.xy}, {w.z,
// cos / pdf, hero throughput after the vertex, radiance base}; radiance base
// -2 = no record, -1 = awaiting the next vertex.
constexpr float kRecNone = -2.0f, kRecPending = -1.0f;

// One iteration body of Renderer::trace (render/renderer.h:102-192) for the
// path in slot s after its closest-hit query (P.hit[s]).  Updates the slot's
// state; returns true when the sample's path has ended.  *push_shadow tells
// whether next-event estimation produced an occlusion ray.
//
// kGuided with gout set: a Lambertian vertex samples the next direction from
// the one-sample mixture 0.5 cos/pi + 0.5 q (q: the guiding distribution in
// the compact form of guide_compact, features gfeat); the mixture density is the
// scatter pdf of both MIS weights (next-event estimation and emitter hits),
// so the estimator stays unbiased.  The vertex is recorded for training in
// P.grec.  Without gout the vertex is shaded as in the plain tracer.
template <bool kGuided = false>
__device__ __forceinline__ bool shade_slot(const DeviceScene &S, const Frame &F, const Paths &P, int s,
                                           bool &push_shadow, ShadowRay &sh, const float *gout = nullptr,
                                           const float *gfeat = nullptr) {
    push_shadow = false;
    {
        float4 o4 = P.o[s], d4 = P.d[s];
        int4 info = P.info[s];
        if (info.w & kFlagGuidePending) {
            // the radiance gathered so far is the base of the previous guided
            // vertex's continuation (never set in unguided renders)
            if (info.z >= 1 && info.z - 1 < F.grec_stride)
                P.grec[(size_t(s) * F.grec_stride + info.z - 1) * 3 + 2].w = P.L[s].x;
            info.w &= ~kFlagGuidePending;
        }
        int2 hit = P.hit[s];
        S4 lam = s4(P.lam[s]);
        S4 beta = s4(P.beta[s]);
        S4 L = s4(P.L[s]);
        V3 o = v3(o4), d = v3(d4);
        float prev_pdf = o4.w;
        bool specular = (info.w & kFlagSpecular) != 0;
        const int bounce = info.z;
        const float total = S.total_power;
        bool finished = false;
        Pcg rng = load_rng(P, s);

        if (hit.x < 0) {
            // Miss: environment with MIS against its NEE (render/renderer.h:105-113).
            float w = 1;
            if (!specular && total > 0)
                w = power_heuristic(prev_pdf, 1 / (4 * kPi) * S.light_pow[S.num_lights - 1] / total);
            S4 bg = background(S, d, lam);
            S4 c = beta * bg * w;
#pragma unroll
            for (int k = 0; k < kHero; k++) L.v[k] += c.v[k];
            finished = true;
        } else {
            Hit h;
            if (!hit_record(S, hit.x, hit.y, o, d, h)) {
                atomicOr(&P.ctr[kCtrError], 1);
                finished = true;
            } else {
                const int mat = material_of(S, hit.x, hit.y);
                const int4 mt = S.mat[mat];
                if (mt.x == LAMBDA_MATERIAL_EMISSIVE) {
                    // Emitter hit (render/renderer.h:119-131).
                    float w = 1;
                    if (!specular && total > 0) {
                        float lp = instance_area(S, hit.x, hit.y) * spectrum_texture_average(S, mt.y);
                        w = power_heuristic(prev_pdf, instance_pdf(S, hit.x, hit.y, o, d) * lp / total);
                    }
                    S4 e = spectrum_texture4(S, mt.y, h, lam);
                    S4 c = beta * e * w;
#pragma unroll
                    for (int k = 0; k < kHero; k++) L.v[k] += c.v[k];
                    finished = true;
                } else {
                    const bool lambertian = mt.x == LAMBDA_MATERIAL_LAMBERTIAN;
                    S4 albedo = s4(0);
                    if (lambertian) albedo = lambertian_albedo(S, mt, h, lam);
                    const bool guided = kGuided && lambertian && gout != nullptr;
                    if (lambertian && total > 0) {
                        // Next-event estimation (render/renderer.h:133-175).
                        float target = total * rng.next();
                        float cum = 0;
                        int2 lio = S.light_io[0];
                        float lp = S.light_pow[0];
                        for (int j = 0; j < S.num_lights; j++) {
                            cum += S.light_pow[j];
                            if (cum >= target) {
                                lio = S.light_io[j];
                                lp = S.light_pow[j];
                                break;
                            }
                        }
                        V3 ldir = lio.x != -1 ? instance_sample(S, lio.x, lio.y, h.p, rng) : random_unit_vector(rng);
                        float lprob = lio.x != -1 ? instance_pdf(S, lio.x, lio.y, h.p, ldir) : 1 / (4 * kPi);
                        float ct = dot(h.n, ldir);
                        if (lprob > 0 && ct > 0) {
                            lprob *= lp / total;
                            float tmax = kMax;
                            S4 em;
                            bool visible_light = true;
                            if (lio.x != -1) {
                                // The occlusion query first intersects the
                                // light itself; a miss counts as occluded
                                // (scene/scene.h:61-70).
                                Hit lh;
                                if (hit_record(S, lio.x, lio.y, h.p, ldir, lh)) {
                                    tmax = lh.t;
                                    em = spectrum_texture4(S, S.mat[material_of(S, lio.x, lio.y)].y, lh, lam);
                                } else {
                                    visible_light = false;
                                }
                            } else {
                                em = background(S, ldir, lam);
                            }
                            if (visible_light) {
                                S4 att = s4(0);
                                if (dot(h.n, ldir) > 0) att = albedo / kPi;
                                float sp = ct > 0 ? ct / kPi : 0;
                                if constexpr (kGuided) {
                                    if (guided) sp = 0.5f * sp + 0.5f * guide_pdf_c(gout, ldir);
                                }
                                float w = power_heuristic(lprob, sp);
                                S4 c = beta * att * ct * em * w / lprob;
                                push_shadow = true;
                                sh.o = make_float4(h.p.x, h.p.y, h.p.z, tmax);
                                sh.d = make_float4(ldir.x, ldir.y, ldir.z, 0);
                                sh.c = f4(c);
                                sh.i = make_int4(s, lio.x, lio.y, 0);
                            }
                        }
                    }
                    // BSDF sampling (scene/material.h:117-261).
                    S4 att;
                    V3 nd;
                    if (lambertian) {
                        bool from_guide = false;
                        if constexpr (kGuided) {
                            if (guided && rng.next() >= 0.5f) {
                                float u0 = rng.next(), u1 = rng.next(), u2 = rng.next();
                                nd = guide_sample_c(gout, u0, u1, u2);
                                from_guide = true;
                            }
                        }
                        if (!from_guide) nd = random_in_hemisphere(h.n, rng);
                        att = dot(h.n, nd) <= 0 ? s4(0) : albedo / kPi;
                    } else if (mt.x == LAMBDA_MATERIAL_MIRROR) {
                        nd = reflected(d, h.n);
                        att = spectrum_texture4(S, mt.y, h, lam);
                    } else if (mt.x == LAMBDA_MATERIAL_DIELECTRIC) {
                        float factor = powf(1 + dot(d, h.n), 5.0f);
                        const float *s0 = S.spectra + size_t(mt.y) * kBins;
                        const float *s1 = S.spectra + size_t(mt.z) * kBins;
                        float n0h = dense(s0, lam.v[0]);
                        float n1h = dense(s1, lam.v[0]);
                        float ratio = h.front ? n0h / n1h : n1h / n0h;
                        float r0 = (n0h - n1h) / (n0h + n1h);
                        r0 *= r0;
                        float rh = r0 + (1 - r0) * factor;
                        bool refl = rng.next() < rh;
                        nd = refl ? reflected(d, h.n) : refracted(d, h.n, ratio);
                        att.v[0] = 1;
#pragma unroll
                        for (int j = 1; j < kHero; j++) {
                            float n0 = dense(s0, lam.v[j]);
                            float n1 = dense(s1, lam.v[j]);
                            float rr = (n0 - n1) / (n0 + n1);
                            rr *= rr;
                            float rj = rr + (1 - rr) * factor;
                            att.v[j] = refl ? rj / rh
                                            : (fabsf(n0 - n0h) < kEps && fabsf(n1 - n1h) < kEps) ? (1 - rj) / (1 - rh) : 0;
                        }
                    } else {
                        // Thin film: hero wavelength picks reflect/refract.
                        const int layers = mt.y;
                        float thick[kMaxLayers];
                        for (int j = 0; j < layers && j < kMaxLayers; j++)
                            thick[j] = scalar_texture(S, S.props[mt.w + j], h.p);
                        const float2 *c_out = S.cspectra + size_t(S.props[mt.z]) * kBins;
                        const float2 *c_in = S.cspectra + size_t(S.props[mt.z + layers + 1]) * kBins;
                        float n0h = dense_c(c_out, lam.v[0]).x;
                        float n1h = dense_c(c_in, lam.v[0]).x;
                        float ratio = h.front ? n0h / n1h : n1h / n0h;
                        float rh = thinfilm_reflectance(S, mt, h.front, h.n, d, lam.v[0], thick);
                        bool refl = rng.next() < rh;
                        nd = refl ? reflected(d, h.n) : refracted(d, h.n, ratio);
                        att.v[0] = 1;
                        for (int j = 1; j < kHero; j++) {
                            float n0 = dense_c(c_out, lam.v[j]).x;
                            float n1 = dense_c(c_in, lam.v[j]).x;
                            float rj = thinfilm_reflectance(S, mt, h.front, h.n, d, lam.v[j], thick);
                            att.v[j] = refl ? rj / rh
                                            : (fabsf(n0 - n0h) < kEps && fabsf(n1 - n1h) < kEps) ? (1 - rj) / (1 - rh) : 0;
                        }
                    }
                    o = h.p;
                    d = nd;
                    float ct = dot(h.n, d);
                    prev_pdf = lambertian ? (ct > 0 ? ct / kPi : 0) : 0;
                    if constexpr (kGuided) {
                        if (guided) {
                            prev_pdf = 0.5f * prev_pdf + 0.5f * guide_pdf_c(gout, d);
                            // below the surface the contribution is zero
                            if (!(ct > 0) || !(prev_pdf > 0)) finished = true;
                        }
                    }
                    specular = !lambertian;
                    if (specular) {
                        beta = beta * att;
                    } else {
                        S4 f = att * ct / prev_pdf;
                        beta = beta * f;
                    }
                    if (bounce >= kRRStart) {
                        float qq = 1 - clampf(s4max(beta), float(0.05), float(0.95));
                        if (rng.next() < qq) finished = true;
                        else {
#pragma unroll
                            for (int k = 0; k < kHero; k++) beta.v[k] /= 1 - qq;
                        }
                    }
                    if (bounce >= F.depth) finished = true;
                    if constexpr (kGuided) {
                        if (guided && !finished && bounce < F.grec_stride) {
                            float4 *r = P.grec + (size_t(s) * F.grec_stride + bounce) * 3;
                            r[0] = make_float4(gfeat[0], gfeat[1], gfeat[2], gfeat[3]);
                            r[1] = make_float4(gfeat[4], gfeat[5], d.x, d.y);
                            r[2] = make_float4(d.z, ct / prev_pdf, beta.v[0], kRecPending);
                            info.w |= kFlagGuidePending;
                        }
                    }
                }
            }
        }
        P.o[s] = make_float4(o.x, o.y, o.z, prev_pdf);
        P.d[s] = make_float4(d.x, d.y, d.z, 0);
        P.beta[s] = f4(beta);
        P.L[s] = f4(L);
        store_rng(P, s, rng);
        P.info[s] = make_int4(info.x, info.y, bounce + 1,
                              (specular ? kFlagSpecular : 0) | (finished ? kFlagFinished : 0) |
                                  (info.w & kFlagGuidePending));
        return finished;
    }
}

// Shading of every queued path.
__global__ void __launch_bounds__(kBlock) k_shade(DeviceScene S, Frame F, Paths P, int count, const int *counts) {
    if (counts) count = counts[0];
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = q < count;
    bool push_shadow = false;
    bool path_done = false;
    int s = 0;
    ShadowRay sh;
    if (active) {
        s = P.q_in[q];
        path_done = shade_slot(S, F, P, s, push_shadow, sh);
    }
    int pos = wave_push(&P.ctr[kCtrShadow], push_shadow);
    if (push_shadow) {
        P.s_o[pos] = sh.o;
        P.s_d[pos] = sh.d;
        P.s_c[pos] = sh.c;
        P.s_i[pos] = sh.i;
    }
    // Continuing paths trace again next pass; ended ones wait for their
    // occlusion ray and are then retired by finalize.
    const bool cont = active && !path_done;
    const int pn = wave_push(&P.ctr[kCtrNext], cont);
    if (cont) P.q_out[pn] = s;
    const bool ended = active && path_done;
    const int pd = wave_push(&P.ctr[kCtrDone], ended);
    if (ended) P.done[pd] = s;
}

// XYZ of one sample's radiance: the first half of spectrumToRGB
// (core/utils.h:212-228), in float, normalised by 106.856895.
__device__ __forceinline__ void sample_xyz(const DeviceScene &S, const Frame &F, float4 Lr, float4 lamr, float &x,
                                           float &y, float &z) {
    const float L[4] = {Lr.x, Lr.y, Lr.z, Lr.w};
    const float lam[4] = {lamr.x, lamr.y, lamr.z, lamr.w};
    x = 0;
    y = 0;
    z = 0;
    float weight = F.lrange / kHero;
#pragma unroll
    for (int i = 0; i < kHero; i++) {
        float l = lam[i];
        int lo = int(l);
        float c[3];
        if (lo == 830) {
            c[0] = S.cie[(lo - 360) * 3 + 0];
            c[1] = S.cie[(lo - 360) * 3 + 1];
            c[2] = S.cie[(lo - 360) * 3 + 2];
        } else {
            int hi = lo + 1;
#pragma unroll
            for (int k = 0; k < 3; k++)
                c[k] = (float(hi) - l) * S.cie[(lo - 360) * 3 + k] + (l - float(lo)) * S.cie[(hi - 360) * 3 + k];
        }
        x += c[0] * L[i] * weight;
        y += c[1] * L[i] * weight;
        z += c[2] * L[i] * weight;
    }
    x /= float(106.856895);
    y /= float(106.856895);
    z /= float(106.856895);
}

// Retires the samples that ended in the previous shade pass (their
// occlusion rays have been traced since): radiance to the framebuffer, then
// the slot's next sample, which joins the closest-hit queue.
// Paths queued for the guiding network in a guided pass.
struct GuideQueue {
    int *slot;       // [n] slot of each queued path
    float *feat;     // [n][6] features
    float *out;      // [n][40] network outputs
};

// Guided pass, step 1: paths on a Lambertian surface within the guided
// bounces queue their features for the network, all others are shaded at
// once (shading closes the record of a path's previous guided vertex).
__global__ void __launch_bounds__(kBlock) k_shade_guided_queue(DeviceScene S, Frame F, Paths P, int count,
                                                               GuideQueue Q) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = q < count;
    bool guided = false, push_shadow = false, path_done = false;
    int s = 0;
    ShadowRay sh;
    GuideFeatures gf;
    if (active) {
        s = P.q_in[q];
        const int4 info = P.info[s];
        if (info.z < F.grec_stride) P.grec[(size_t(s) * F.grec_stride + info.z) * 3 + 2].w = kRecNone;
        const int2 hit = P.hit[s];
        if (hit.x >= 0 && info.z < F.guide_bounces &&
            S.mat[material_of(S, hit.x, hit.y)].x == LAMBDA_MATERIAL_LAMBERTIAN) {
            const float4 o4 = P.o[s], d4 = P.d[s];
            Hit h;
            if (hit_record(S, hit.x, hit.y, v3(o4), v3(d4), h)) {
                gf = guide_features(h.p, v3(F.box_lo.x, F.box_lo.y, F.box_lo.z), v3(F.box_hi.x, F.box_hi.y, F.box_hi.z),
                                    h.n, P.lam[s].x);
                guided = true;
            }
        }
        if (!guided) path_done = shade_slot<false>(S, F, P, s, push_shadow, sh);
    }
    const int gp = wave_push(&P.ctr[kCtrGuided], guided);
    if (guided) {
        Q.slot[gp] = s;
#pragma unroll
        for (int k = 0; k < kGuideScalars; k++) Q.feat[size_t(gp) * kGuideScalars + k] = gf.f[k];
    }
    const bool plain = active && !guided;
    int pos = wave_push(&P.ctr[kCtrShadow], push_shadow);
    if (push_shadow) {
        P.s_o[pos] = sh.o;
        P.s_d[pos] = sh.d;
        P.s_c[pos] = sh.c;
        P.s_i[pos] = sh.i;
    }
    const bool cont = plain && !path_done;
    const int pn = wave_push(&P.ctr[kCtrNext], cont);
    if (cont) P.q_out[pn] = s;
    const bool ended = plain && path_done;
    const int pd = wave_push(&P.ctr[kCtrDone], ended);
    if (ended) P.done[pd] = s;
}

// Guided pass, step 2: the queued Lambertian paths shade with the network's
// mixture (P.ctr[kCtrGuided] of them).
__global__ void __launch_bounds__(kBlock) k_shade_guided(DeviceScene S, Frame F, Paths P, GuideQueue Q) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = q < P.ctr[kCtrGuided];
    bool push_shadow = false, path_done = false;
    int s = 0;
    ShadowRay sh;
    if (active) {
        s = Q.slot[q];
        float *lobes = Q.out + size_t(q) * kLobeOut;
        guide_compact(lobes, lobes);
        path_done = shade_slot<true>(S, F, P, s, push_shadow, sh, lobes, Q.feat + size_t(q) * kGuideScalars);
    }
    int pos = wave_push(&P.ctr[kCtrShadow], push_shadow);
    if (push_shadow) {
        P.s_o[pos] = sh.o;
        P.s_d[pos] = sh.d;
        P.s_c[pos] = sh.c;
        P.s_i[pos] = sh.i;
    }
    const bool cont = active && !path_done;
    const int pn = wave_push(&P.ctr[kCtrNext], cont);
    if (cont) P.q_out[pn] = s;
    const bool ended = active && path_done;
    const int pd = wave_push(&P.ctr[kCtrDone], ended);
    if (ended) P.done[pd] = s;
}

// 32-bit integer hash of three words (lowbias32 finaliser by C. Wellons).
__device__ __forceinline__ uint32_t hash3(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t x = a * 0x9e3779b1u ^ (b + 0x7f4a7c15u) * 0x85ebca6bu ^ (c + 0x165667b1u) * 0xc2b2ae35u;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Adds the ended sample of slot s to its pixel's XYZ sum.  In reference
// mode the pixel's next sample then starts in the same slot, continuing the
// pixel's PCG32 stream (render/renderer.h:69-91); returns true when it did.
// A pixel that reaches the launch's last sample before spp saves its stream.
__device__ __forceinline__ bool retire_sample(const DeviceScene &S, const Frame &F, const Paths &P, int s) {
    int4 info = P.info[s];
    if (F.train) {
        // training samples from the sample's guided vertices: the radiance
        // that reached vertex b along its sampled direction is the radiance
        // gathered after it over the throughput after it (hero wavelength)
        const float Lend = P.L[s].x;
        const int nb = min(min(info.z, F.grec_stride), F.guide_bounces);
        for (int b = 0; b < nb; b++) {
            float4 *r = P.grec + (size_t(s) * F.grec_stride + b) * 3;
            const float4 r2 = r[2];
            if (!(r2.w >= 0)) continue;
            r[2].w = kRecNone;
            const float c = (Lend - r2.w) / r2.z * r2.y;
            if (!(c > 0) || !isfinite(c)) continue;
            // a uniform subsample of the candidates (by a hash of slot,
            // sample and vertex), so that the buffer does not fill with the
            // paths that end first
            atomicAdd(&P.ctr[kCtrTrainSeen], 1);
            const uint32_t hsh = hash3(uint32_t(s), uint32_t(info.y) * 64u + uint32_t(b), uint32_t(F.seed));
            if (float(hsh >> 8) * (1.0f / 16777216.0f) >= F.train_keep) continue;
            const int k = atomicAdd(&P.ctr[kCtrTrain], 1);
            if (k >= F.train_cap) continue;
            const float4 r0 = r[0], r1 = r[1];
            float *o = F.train + size_t(k) * kGuideSample;
            o[0] = r0.x; o[1] = r0.y; o[2] = r0.z; o[3] = r0.w; o[4] = r1.x; o[5] = r1.y;
            o[6] = r1.z; o[7] = r1.w; o[8] = r2.x; o[9] = c;
        }
    }
    float x, y, z;
    sample_xyz(S, F, P.L[s], P.lam[s], x, y, z);
    double *px = P.fb + size_t(info.x) * 3;
    unsafeAtomicAdd(px + 0, double(x));
    unsafeAtomicAdd(px + 1, double(y));
    unsafeAtomicAdd(px + 2, double(z));
    if (F.mode == LAMBDA_RNG_REFERENCE && info.y + 1 < (F.sample_end ? F.sample_end : F.spp)) {
        Pcg rng = load_rng(P, s);
        camera_sample(S, F, P, s, info.x, info.y + 1, rng);
        store_rng(P, s, rng);
        return true;
    }
    if (F.mode == LAMBDA_RNG_REFERENCE && info.y + 1 < F.spp) {
        F.rng_save[info.x] = P.rng[s];  // paused: a later launch resumes the stream here
        return false;
    }
    atomicAdd(&P.ctr[kCtrPixels], 1);
    return false;
}

// Retires the samples that ended in the previous shade pass (their
// occlusion rays have been traced since): radiance to the framebuffer, then
// the slot's next sample, which joins the closest-hit queue.
__global__ void __launch_bounds__(kBlock) k_finalize(DeviceScene S, Frame F, Paths P, int n_done,
                                                     const int *counts) {
    if (counts) n_done = counts[2];
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0) {
        P.ctr[kCtrShadow] = 0;
        P.ctr[kCtrDone] = 0;
        P.ctr[kCtrGuided] = 0;
    }
    bool keep = false;
    bool need_work = false;
    int s = 0;
    if (q < n_done) {
        s = P.done[q];
        keep = retire_sample(S, F, P, s);
        need_work = !keep;
    }
    // a slot whose unit is finished takes the next one
    int w = wave_push(&P.ctr[kCtrWork], need_work);
    if (need_work && w < work_units(F)) {
        start_unit(S, F, P, s, w);
        keep = true;
    }
    int pos = wave_push(&P.ctr[kCtrNext], keep);
    if (keep) P.q_out[pos] = s;
}

// Persistent per-thread schedule: each thread takes one live slot and runs
// it -- closest hit, shading, the occlusion ray at once, retirement, the
// slot's next sample, then further work units -- without waiting for other
// paths.  Threads [0, n_in) take q_in (paths that need their closest hit);
// the next P.ctr[kCtrNext] threads take q_out (samples just started by
// k_finalize).  No occlusion ray or ended sample may be pending when it
// starts.  Threads exit when no work unit is left.
__global__ void __launch_bounds__(kBlock, kPersistWaves) k_persist(DeviceScene S, Frame F, Paths P, int n_in,
                                                                 int2 *spill) {
    __shared__ int s_code[kRing * kBlock];
    __shared__ float s_dist[kRing * kBlock];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_new = P.ctr[kCtrNext];
    if (t >= n_in + n_new) return;
    const int s = t < n_in ? P.q_in[t] : P.q_out[t - n_in];
    Stack stk = make_stack(s_code, s_dist, spill);
    unsigned n_closest = 0, n_shadow = 0;
    if (F.cost) F.unit_t0[s] = wall_clock64();
    while (true) {
        n_closest++;
        {
            float4 o = P.o[s], d = P.d[s];
            float best = kMax;
            int hi = -1, ho = -1;
            stk.sp = stk.lo = 0;
            trace_ray<false>(S, v3(o), v3(d), best, hi, ho, -1, -2, stk);
            P.hit[s] = make_int2(hi, ho);
        }
        bool push = false;
        ShadowRay sh;
        const bool ended = shade_slot(S, F, P, s, push, sh);
        if (push) {
            n_shadow++;
            float best = sh.o.w;
            int hi = -1, ho = -1;
            stk.sp = stk.lo = 0;
            if (!trace_ray<true>(S, v3(sh.o), v3(sh.d), best, hi, ho, sh.i.y, sh.i.z, stk)) {
                float4 L = P.L[s];
                L.x += sh.c.x;
                L.y += sh.c.y;
                L.z += sh.c.z;
                L.w += sh.c.w;
                P.L[s] = L;
            }
        }
        if (!ended) continue;
        if (retire_sample(S, F, P, s)) continue;
        if (F.cost) {
            // device-clock time of the unit (its pixel's samples in this launch)
            const unsigned long long now = wall_clock64();
            const unsigned long long dt = now - F.unit_t0[s];
            atomicAdd(F.cost + P.info[s].x, int(dt < (1ull << 30) ? dt : (1ull << 30)));
            F.unit_t0[s] = now;
        }
        const int w = atomicAdd(&P.ctr[kCtrWork], 1);
        if (w >= work_units(F)) break;
        start_unit(S, F, P, s, w);
    }
    atomicAdd(&P.rays[0], (unsigned long long)n_closest);
    atomicAdd(&P.rays[1], (unsigned long long)n_shadow);
}

// Waves per SIMD of the guided persistent kernel (its LDS, stacks plus the
// network's staging, allows two blocks per CU).  Staging in the traversal
// stacks' LDS instead, at 3 waves (168 VGPRs), spilled 158 registers and
// ran 4 % slower on the guiding room.
constexpr int kPersistGuidedWaves = 2;

// The persistent schedule with the guiding network evaluated inside the
// kernel.  Every thread runs its paths as in k_persist; after each closest
// hit, the lanes at a guided vertex (Lambertian, bounce index below
// F.guide_bounces) put their features in the wave's LDS rows, and if any
// lane has one, the whole wave runs the network on the matrix cores for the
// guided lanes only (packed into the first rows, empty 16-row tiles
// skipped) and each guided lane shades with its own row of outputs.  The
// loop ends for the wave together, when no lane has a path.
__global__ void __launch_bounds__(kBlock, kPersistGuidedWaves) k_persist_guided(DeviceScene S, Frame F, Paths P,
                                                                                int n_in, int2 *spill, GuideDev G) {
    __shared__ int s_code[kRing * kBlock];
    __shared__ float s_dist[kRing * kBlock];
    // per wave: features [64][6] in, network outputs [64][40] out
    __shared__ __attribute__((aligned(16))) float s_io[kBlock / 64][64 * kLobeOut];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    float *io = s_io[threadIdx.x >> 6];
    const int n_new = P.ctr[kCtrNext];
    bool alive = t < n_in + n_new;
    const int s = alive ? (t < n_in ? P.q_in[t] : P.q_out[t - n_in]) : 0;
    Stack stk = make_stack(s_code, s_dist, spill);
    unsigned n_closest = 0, n_shadow = 0;
    while (__any(alive)) {
        bool guided = false;
        GuideFeatures gf{};
        if (alive) {
            n_closest++;
            float4 o = P.o[s], d = P.d[s];
            float best = kMax;
            int hi = -1, ho = -1;
            stk.sp = stk.lo = 0;
            trace_ray<false>(S, v3(o), v3(d), best, hi, ho, -1, -2, stk);
            P.hit[s] = make_int2(hi, ho);
            const int bounce = P.info[s].z;
            if (bounce < F.grec_stride && bounce < F.guide_bounces)
                P.grec[(size_t(s) * F.grec_stride + bounce) * 3 + 2].w = kRecNone;
            if (hi >= 0 && bounce < F.guide_bounces && S.mat[material_of(S, hi, ho)].x == LAMBDA_MATERIAL_LAMBERTIAN) {
                Hit h;
                if (hit_record(S, hi, ho, v3(o), v3(d), h)) {
                    gf = guide_features(h.p, v3(F.box_lo.x, F.box_lo.y, F.box_lo.z),
                                        v3(F.box_hi.x, F.box_hi.y, F.box_hi.z), h.n, P.lam[s].x);
                    guided = true;
                }
            }
        }
        // the wave's guided lanes take rows 0 .. n-1 in lane order
        const unsigned long long gmask = __ballot(guided);
        const int grow = __popcll(gmask & ((1ull << lane) - 1ull));
        if (gmask) {
            // every lane of the wave is here (the condition is uniform)
            if (guided) {
#pragma unroll
                for (int k = 0; k < kGuideScalars; k++) io[grow * kGuideScalars + k] = gf.f[k];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // reads the features in its first layer, writes [n][40] at the end
            mlp_rows64_call(G, io, kGuideScalars, __popcll(gmask), io, kLobeOut);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // each guided lane turns its own row into the compact mixture
            if (guided) guide_compact(io + grow * kLobeOut, io + grow * kLobeOut);
        }
        if (alive) {
            bool push = false;
            ShadowRay sh;
            const bool ended = shade_slot<true>(S, F, P, s, push, sh, guided ? io + grow * kLobeOut : nullptr, gf.f);
            if (push) {
                n_shadow++;
                float best = sh.o.w;
                int hi = -1, ho = -1;
                stk.sp = stk.lo = 0;
                if (!trace_ray<true>(S, v3(sh.o), v3(sh.d), best, hi, ho, sh.i.y, sh.i.z, stk)) {
                    float4 L = P.L[s];
                    L.x += sh.c.x;
                    L.y += sh.c.y;
                    L.z += sh.c.z;
                    L.w += sh.c.w;
                    P.L[s] = L;
                }
            }
            if (ended && !retire_sample(S, F, P, s)) {
                const int w = atomicAdd(&P.ctr[kCtrWork], 1);
                if (w < work_units(F)) start_unit(S, F, P, s, w);
                else alive = false;
            }
        }
    }
    atomicAdd(&P.rays[0], (unsigned long long)n_closest);
    atomicAdd(&P.rays[1], (unsigned long long)n_shadow);
}

// ------------------------------------------------------ batched serving
//
// One persistent launch renders a batch of jobs.  A job is a scene and a
// frame rendered with reference streams; its units are its pixels, numbered
// consecutively across the batch in arrival order.  Threads claim units
// from one counter; a thread that claims a unit of a job that has not
// arrived yet sleeps until the job's arrival time on the device wall clock.
// Each unit runs all samples of its pixel, so every job's image equals the
// image of a single render of the same frame bit for bit.  The thread that
// finishes a job's last pixel records the finishing time.
struct ServeJob {
    Frame F;                        // reference streams; fb below
    double *fb;                     // XYZ sums of the job, [pixel][3]
    long long first_unit;           // first unit of the job in the batch
    unsigned long long arrival;     // wall-clock ticks after the batch start
    int scene;                      // index into the scene table
    int units;                      // pixels
};

struct ServeCtl {
    unsigned long long t0;          // wall clock at the batch start
    int next_unit;                  // unit counter
    int pad;
};

__global__ void k_serve_clock(ServeCtl *ctl) {
    ctl->t0 = wall_clock64();
    ctl->next_unit = 0;
}

__global__ void __launch_bounds__(kBlock, kPersistWaves) k_serve(const DeviceScene *scenes, const ServeJob *jobs,
                                                                  int njobs, long long total_units, Paths P,
                                                                  ServeCtl *ctl, int *done, unsigned long long *finish,
                                                                  int2 *spill) {
    __shared__ int s_code[kRing * kBlock];
    __shared__ float s_dist[kRing * kBlock];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    Stack stk = make_stack(s_code, s_dist, spill);
    const unsigned long long t0 = ctl->t0;
    unsigned n_closest = 0, n_shadow = 0;
    int j = 0;
    while (true) {
        const int w = atomicAdd(&ctl->next_unit, 1);
        if (w >= total_units) break;
        // claims only grow, so the job index only moves forward
        while (j + 1 < njobs && w >= jobs[j + 1].first_unit) j++;
        const ServeJob &J = jobs[j];
        const unsigned long long due = t0 + J.arrival;
        while (wall_clock64() < due) __builtin_amdgcn_s_sleep(127);
        const DeviceScene &S = scenes[J.scene];
        const Frame &F = J.F;
        Paths Pj = P;
        Pj.fb = J.fb;
        const int pixel = int(w - J.first_unit);
        {
            Pcg rng = pcg_seed(F.seed, uint64_t(pixel));
            camera_sample(S, F, Pj, s, pixel, 0, rng);
            store_rng(Pj, s, rng);
        }
        while (true) {
            n_closest++;
            {
                float4 o = Pj.o[s], d = Pj.d[s];
                float best = kMax;
                int hi = -1, ho = -1;
                stk.sp = stk.lo = 0;
                trace_ray<false>(S, v3(o), v3(d), best, hi, ho, -1, -2, stk);
                Pj.hit[s] = make_int2(hi, ho);
            }
            bool push = false;
            ShadowRay sh;
            const bool ended = shade_slot(S, F, Pj, s, push, sh);
            if (push) {
                n_shadow++;
                float best = sh.o.w;
                int hi = -1, ho = -1;
                stk.sp = stk.lo = 0;
                if (!trace_ray<true>(S, v3(sh.o), v3(sh.d), best, hi, ho, sh.i.y, sh.i.z, stk)) {
                    float4 L = Pj.L[s];
                    L.x += sh.c.x;
                    L.y += sh.c.y;
                    L.z += sh.c.z;
                    L.w += sh.c.w;
                    Pj.L[s] = L;
                }
            }
            if (!ended) continue;
            if (retire_sample(S, F, Pj, s)) continue;
            break;
        }
        if (atomicAdd(&done[j], 1) == J.units - 1) finish[j] = wall_clock64();
    }
    atomicAdd(&P.rays[0], (unsigned long long)n_closest);
    atomicAdd(&P.rays[1], (unsigned long long)n_shadow);
}

// Mean XYZ over samples to linear sRGB (core/utils.h:230), in double.
__global__ void k_resolve(Frame F, const double *fb, float *rgb) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= F.npix) return;
    double x = fb[3 * size_t(p) + 0] / F.spp;
    double y = fb[3 * size_t(p) + 1] / F.spp;
    double z = fb[3 * size_t(p) + 2] / F.spp;
    rgb[3 * size_t(p) + 0] = float(x * double(float(3.2406255)) - y * double(float(1.5372080)) - z * double(float(0.4986286)));
    rgb[3 * size_t(p) + 1] = float(-x * double(float(0.9689307)) + y * double(float(1.8757561)) + z * double(float(0.0415175)));
    rgb[3 * size_t(p) + 2] = float(x * double(float(0.0557101)) - y * double(float(0.2040211)) + z * double(float(1.0569959)));
}

__global__ void k_set_i(int *p, int v) { *p = v; }

__global__ void k_fill_d(double *p, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_fill_u64(unsigned long long *p, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
__global__ void k_fill_i(int *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ------------------------------------------------------------ query kernels

__global__ void __launch_bounds__(kBlock) k_query_closest(DeviceScene S, const float *rays, int n, lambda_hit *out,
                                                          int2 *spill) {
    __shared__ int s_code[kRing * kBlock];
    __shared__ float s_dist[kRing * kBlock];
    Stack stk = make_stack(s_code, s_dist, spill);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 o = v3(rays[6 * i + 0], rays[6 * i + 1], rays[6 * i + 2]);
    V3 d = v3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
    float best = kMax;
    int hi = -1, ho = -1;
    bool ok = trace_ray<false>(S, o, d, best, hi, ho, -1, -2, stk);
    lambda_hit r;
    r.hit = 0;
    r.instance = -1;
    r.object = -1;
    r.t = kMax;
    r.point[0] = r.point[1] = r.point[2] = 0;
    r.normal[0] = r.normal[1] = r.normal[2] = 0;
    r.u = r.v = 0;
    r.front_facing = 0;
    if (ok) {
        Hit h;
        if (hit_record(S, hi, ho, o, d, h)) {
            r.hit = 1;
            r.instance = hi;
            r.object = ho;
            r.t = h.t;
            r.point[0] = h.p.x; r.point[1] = h.p.y; r.point[2] = h.p.z;
            r.normal[0] = h.n.x; r.normal[1] = h.n.y; r.normal[2] = h.n.z;
            r.u = h.u;
            r.v = h.v;
            r.front_facing = h.front ? 1 : 0;
        }
    }
    out[i] = r;
}

__global__ void __launch_bounds__(kBlock) k_query_occluded(DeviceScene S, const float *rays, const int *li,
                                                           const int *lo, int n, int *out, int2 *spill) {
    __shared__ int s_code[kRing * kBlock];
    __shared__ float s_dist[kRing * kBlock];
    Stack stk = make_stack(s_code, s_dist, spill);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 o = v3(rays[6 * i + 0], rays[6 * i + 1], rays[6 * i + 2]);
    V3 d = v3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
    float best = kMax;
    int inst = li[i], obj = lo[i];
    if (inst != -1) {
        Hit lh;
        if (!hit_record(S, inst, obj, o, d, lh)) {
            out[i] = 1;
            return;
        }
        best = lh.t;
    }
    int hi = -1, ho = -1;
    out[i] = trace_ray<true>(S, o, d, best, hi, ho, inst, obj, stk) ? 1 : 0;
}

__global__ void k_query_thinfilm(DeviceScene S, int mat, const float *q, const float *pts, int n, float *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float c = q[4 * i + 0];
    float lambda = q[4 * i + 1];
    bool front = q[4 * i + 2] != 0;
    float s = sqrtf(fmaxf(0.0f, 1 - c * c));
    V3 p = v3(pts[3 * i + 0], pts[3 * i + 1], pts[3 * i + 2]);
    int4 mt = S.mat[mat];
    float thick[kMaxLayers];
    for (int j = 0; j < mt.y && j < kMaxLayers; j++) thick[j] = scalar_texture(S, S.props[mt.w + j], p);
    out[i] = thinfilm_reflectance(S, mt, front, v3(0, 0, 1), v3(s, 0, -c), lambda, thick);
}

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



void validate(const lambda_scene_view &v) {
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
    require(tree_stack_depth(v, 0, v.num_instances, true) <= kStack, "instance BVH deeper than 20");
    for (int i = 0; i < v.num_instances; i++)
        require(tree_stack_depth(v, v.instance_ints[4 * i + 2], v.num_objects, false) <= kStack,
                "object BVH deeper than 20");
}

}  // namespace

struct lambda_hip_scene {
    DevBuf<float4> wide, obj, obj_s, inst, stex;
    DevBuf<int> inst_root, obj_mat, props, perm, tlas_rank, obj_rank;
    int stack_need = 0;     // deepest traversal stack of the scene
    float box[6] = {0, 0, 0, 0, 0, 0};  // scene bounds (instance-level root)
    DevBuf<float4> worley_points;
    DevBuf<int4> worley_box, worley_dim;
    DevBuf<int4> inst_i, mat, xtex;
    DevBuf<float> spectra, light_pow, cie;
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

void upload_scene(const lambda_scene_view &v, lambda_hip_scene &g) {
    validate(v);
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

// Pixels in 8x8 tiles, tiles row-major: neighbouring lanes trace
// neighbouring pixels.
std::vector<int> tile_order(int w, int h) {
    std::vector<int> order;
    order.reserve(size_t(w) * h);
    for (int ty = 0; ty < h; ty += 8)
        for (int tx = 0; tx < w; tx += 8)
            for (int y = ty; y < std::min(ty + 8, h); y++)
                for (int x = tx; x < std::min(tx + 8, w); x++) order.push_back(y * w + x);
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


// Threads of k_persist resident on the device at once.
template <typename Kernel>
long long resident_threads(Kernel kernel) {
    int dev = 0, cus = 0, blocks = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kBlock, 0));
    return (long long)std::max(1, cus) * std::max(1, blocks) * kBlock;
}

long long resident_persist_threads() { return resident_threads(k_persist); }

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
    return F;
}

// Device buffers, stream and counters of one render.
// Visiting order from measured pixel costs: 8x8 tiles (the unit of the
// default order, which keeps neighbouring paths in flight together) are
// summed on the device, sorted on the host (longest first) and expanded back
// to pixels, row order within a tile.  The order changes when a pixel runs,
// never what it computes.
__global__ void k_tile_cost(const int *cost, int w, int h, long long *tile_cost) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= w * h) return;
    const int tx = (p % w) >> 3, ty = (p / w) >> 3;
    atomicAdd(reinterpret_cast<unsigned long long *>(tile_cost + size_t(ty) * ((w + 7) >> 3) + tx),
              (unsigned long long)cost[p]);
}

// Thread (i, k): pixel k of the tile at position i of the sorted list.
__global__ void k_tile_expand(const int *tiles, const int *offsets, int ntiles, int w, int h, int *order) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = int(e >> 6), k = int(e & 63);
    if (i >= ntiles) return;
    const int t = tiles[i], tiles_x = (w + 7) >> 3;
    const int tx = t % tiles_x, ty = t / tiles_x;
    const int tw = min(8, w - 8 * tx), th = min(8, h - 8 * ty);
    if (k >= tw * th) return;
    order[offsets[i] + k] = (8 * ty + k / tw) * w + 8 * tx + k % tw;
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
        B.order.upload(tile_order(F.width, F.height));
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
    // launch).
    void run(const Frame &F, long long tail, volatile int32_t *progress, lambda_render_stats *stats,
             bool clear = true, const GuideNet *guide = nullptr, GuideQueue *Q = nullptr, bool fused = false) {
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
        if (clear) k_fill_d<<<grid_for((long long)F.npix * 3), kBlock, 0, st>>>(P.fb, size_t(F.npix) * 3);
        k_fill_i<<<1, kBlock, 0, st>>>(P.ctr, kCtrCount);
        k_fill_u64<<<1, kBlock, 0, st>>>(P.rays, 2);
        if (guide) tail = fused ? kAutoTailPaths : 0;
        k_start<<<grid_for(ns), kBlock, 0, st>>>(g.ds, F, P, ns, 0);
        // The work counter starts after the slots' first units.
        k_set_i<<<1, 1, 0, st>>>(P.ctr + kCtrWork, ns);
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
                k_trace<<<grid_for(n_shadow), kBlock, 0, st>>>(g.ds, P, 0, n_shadow, B.spill.p, nullptr);
                if (stats) HIP_OK(hipEventRecord(ev[0].b, st));
                k_finalize<<<grid_for(n_done), kBlock, 0, st>>>(g.ds, F, P, n_done, nullptr);
                if (stats) HIP_OK(hipEventRecord(ev[3].a, st));
                if (guide)
                    k_persist_guided<<<grid_for((long long)n_closest + n_done), kBlock, 0, st>>>(
                        g.ds, F, P, n_closest, B.spill.p, guide->dev());
                else
                    k_persist<<<grid_for((long long)n_closest + n_done), kBlock, 0, st>>>(g.ds, F, P, n_closest,
                                                                                          B.spill.p);
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
            k_trace<<<grid_for((long long)n_closest + n_shadow), kBlock, 0, st>>>(g.ds, P, n_closest, n_shadow, B.spill.p, nullptr);
            if (stats) {
                HIP_OK(hipEventRecord(ev[0].b, st));
                HIP_OK(hipEventRecord(ev[2].a, st));
            }
            k_finalize<<<grid_for(n_done), kBlock, 0, st>>>(g.ds, F, P, n_done, nullptr);
            if (stats) {
                HIP_OK(hipEventRecord(ev[2].b, st));
                HIP_OK(hipEventRecord(ev[1].a, st));
            }
            if (guide) {
                k_shade_guided_queue<<<grid_for(n_closest), kBlock, 0, st>>>(g.ds, F, P, n_closest, *Q);
                guide->forward_dev(Q->feat, P.ctr + kCtrGuided, n_closest, Q->out, st);
                k_shade_guided<<<grid_for(n_closest), kBlock, 0, st>>>(g.ds, F, P, *Q);
            } else {
                k_shade<<<grid_for(n_closest), kBlock, 0, st>>>(g.ds, F, P, n_closest, nullptr);
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
        if (clear) k_fill_d<<<grid_for((long long)F.npix * 3), kBlock, 0, st>>>(P.fb, size_t(F.npix) * 3);
        k_fill_i<<<1, kBlock, 0, st>>>(P.ctr, kCtrCount);
        k_fill_u64<<<1, kBlock, 0, st>>>(P.rays, 2);
        // no unit is started after a batch's first ones
        k_set_i<<<1, 1, 0, st>>>(P.ctr + kCtrWork, int(units));
        HIP_OK(hipGetLastError());
        for (long long base = 0; base < units; base += ns) {
            const int nb = int(std::min<long long>(ns, units - base));
            k_start<<<grid_for(nb), kBlock, 0, st>>>(g.ds, F, P, nb, base);
            int n_closest = nb, n_shadow = 0, n_done = 0;
            for (int pass = 0; pass < F.guide_bounces && n_closest > 0; pass++) {
                if (stats) HIP_OK(hipEventRecord(ev[0].a, st));
                k_trace<<<grid_for((long long)n_closest + n_shadow), kBlock, 0, st>>>(g.ds, P, n_closest, n_shadow,
                                                                                       B.spill.p, nullptr);
                if (stats) {
                    HIP_OK(hipEventRecord(ev[0].b, st));
                    HIP_OK(hipEventRecord(ev[2].a, st));
                }
                k_finalize<<<grid_for(n_done), kBlock, 0, st>>>(g.ds, F, P, n_done, nullptr);
                if (stats) {
                    HIP_OK(hipEventRecord(ev[2].b, st));
                    HIP_OK(hipEventRecord(ev[1].a, st));
                }
                k_shade_guided_queue<<<grid_for(n_closest), kBlock, 0, st>>>(g.ds, F, P, n_closest, *Q);
                guide->forward_dev(Q->feat, P.ctr + kCtrGuided, n_closest, Q->out, st);
                k_shade_guided<<<grid_for(n_closest), kBlock, 0, st>>>(g.ds, F, P, *Q);
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
            k_trace<<<grid_for(n_shadow), kBlock, 0, st>>>(g.ds, P, 0, n_shadow, B.spill.p, nullptr);
            if (stats) HIP_OK(hipEventRecord(ev[0].b, st));
            k_finalize<<<grid_for(n_done), kBlock, 0, st>>>(g.ds, F, P, n_done, nullptr);
            if (stats) HIP_OK(hipEventRecord(ev[3].a, st));
            if (n_closest + n_done > 0)
                k_persist<<<grid_for((long long)n_closest + n_done), kBlock, 0, st>>>(g.ds, F, P, n_closest,
                                                                                      B.spill.p);
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
        k_fill_d<<<grid_for((long long)F.npix * 3), kBlock, 0, st>>>(P.fb, size_t(F.npix) * 3);
        k_fill_i<<<1, kBlock, 0, st>>>(P.ctr, kCtrCount);
        k_fill_u64<<<1, kBlock, 0, st>>>(P.rays, 2);
        k_start<<<grid_for(ns), kBlock, 0, st>>>(g.ds, F, P, ns, 0);
        k_set_i<<<1, 1, 0, st>>>(P.ctr + kCtrWork, ns);
        k_set_i<<<1, 1, 0, st>>>(P.ctr + kCtrNext, ns);  // every slot starts in the closest-hit queue
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
            k_pass_begin<<<1, 1, 0, st>>>(Q.ctr, counts.p);
            k_trace<<<grid_for(2LL * ns), kBlock, 0, st>>>(g.ds, Q, 0, 0, B.spill.p, counts.p);
            k_finalize<<<grid_for(ns), kBlock, 0, st>>>(g.ds, F, Q, 0, counts.p);
            k_shade<<<grid_for(ns), kBlock, 0, st>>>(g.ds, F, Q, 0, counts.p);
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
        const int tiles_x = (F.width + 7) / 8, tiles_y = (F.height + 7) / 8, ntiles = tiles_x * tiles_y;
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
        k_fill_i<<<grid_for(F.npix), kBlock, 0, st>>>(cost.p, F.npix);
        k_fill_u64<<<grid_for(ntiles), kBlock, 0, st>>>(reinterpret_cast<unsigned long long *>(tile_cost.p),
                                                         size_t(ntiles));
        Frame F1 = F;
        F1.sample_end = first;
        F1.cost = cost.p;
        F1.unit_t0 = t0.p;
        F1.rng_save = save.p;
        run(F1, tail, nullptr, stats, true);
        k_tile_cost<<<grid_for(F.npix), kBlock, 0, st>>>(cost.p, F.width, F.height, tile_cost.p);
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
            pos += std::min(8, F.width - 8 * (t % tiles_x)) * std::min(8, F.height - 8 * (t / tiles_x));
        }
        HIP_OK(hipMemcpyAsync(tile_ids.p, ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(offsets.p, offs.data(), offs.size() * sizeof(int), hipMemcpyHostToDevice, st));
        k_tile_expand<<<grid_for((long long)ntiles * 64), kBlock, 0, st>>>(tile_ids.p, offsets.p, ntiles, F.width,
                                                                           F.height, order.p);
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
    long long slots = p.path_slots > 0 ? p.path_slots : resident_persist_threads();
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
    // reference streams, persistent schedule, more pixels than slots: the
    // two-launch cost order (cost_order samples first; 0 = automatic: 4, the
    // best measured on the 1080p instanced-mesh frame against 1, 2 and 8)
    const bool by_cost = F.mode == LAMBDA_RNG_REFERENCE && F.spp > 1 && p.graph_passes <= 0 &&
                         tail > R.ns && (p.cost_order > 0 || (p.cost_order == 0 && R.ns < F.npix));
    if (by_cost) R.run_cost_ordered(F, std::min(p.cost_order > 0 ? p.cost_order : 4, F.spp - 1), tail, progress, stats);
    else if (p.graph_passes > 0) R.run_graph(F, p.graph_passes, progress, stats);
    else R.run(F, tail, progress, stats);
    k_resolve<<<grid_for(F.npix), kBlock, 0, R.st>>>(F, R.P.fb, R.B.rgb.p);
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
// eight bounces (on the guiding room, 2, 4, 6, 10, 12, 16 or every bounce
// gave less variance reduction per unit time).
constexpr int kDefaultGuidedBounces = 8;

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
    bool fused = false;          // guided: network inside the persistent kernel
    double seen_per_sample = 0;  // training candidates per sample per pixel, last step
    DevBuf<double> scratch_fb;   // framebuffer of steps that are not kept
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
    if (pr.net) F.guide_bounces = pr.guide_bounces;
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
        k_query_closest<<<grid_for(n), kBlock>>>(scene->ds, r.p, n, h.p, spill.p);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpy(hits, h.p, size_t(n) * sizeof(lambda_hit), hipMemcpyDeviceToHost));
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
        k_query_occluded<<<grid_for(n), kBlock>>>(scene->ds, r.p, li.p, lo.p, n, out.p, spill.p);
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
        k_query_thinfilm<<<grid_for(n), kBlock>>>(scene->ds, material, q.p, p.p, n, o.p);
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
        k_fill_i<<<1, kBlock, 0, st>>>(P.ctr, kCtrCount);
        k_fill_u64<<<1, kBlock, 0, st>>>(P.rays, 2);
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
            k_fill_d<<<grid_for(px_total * 3), kBlock, 0, st>>>(fb.p, size_t(px_total) * 3);
            k_fill_i<<<grid_for((long long)batch.size()), kBlock, 0, st>>>(done.p, int(batch.size()));
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
            k_serve_clock<<<1, 1, 0, st>>>(ctl.p);
            k_serve<<<threads / kBlock, kBlock, 0, st>>>(dscenes.p, djobs.p, int(sj.size()), units, P, ctl.p, done.p,
                                                        finish.p, B.spill.p);
            HIP_OK(hipGetLastError());
            off = 0;
            for (size_t k = 0; k < batch.size(); k++) {
                const ServeJob &J = sj[k];
                k_resolve<<<grid_for(J.F.npix), kBlock, 0, st>>>(J.F, J.fb, rgb.p + off * 3);
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
        const Frame F = make_frame(p);
        // guided: the network inside the persistent kernel unless the
        // caller asks for one launch per bounce (tail_paths -1)
        pr->fused = guide && params->tail_paths != -1;
        long long slots = params->path_slots > 0 ? params->path_slots
                          : pr->fused ? resident_threads(k_persist_guided)
                          : guide     ? (1LL << 20)
                                      : resident_persist_threads();
        slots = std::max<long long>(1, std::min<long long>(slots, 1LL << 26));
        pr->R = std::make_unique<Renderer>(*scene, F, int(slots));
        if (guide) {
            const size_t ns = size_t(slots);
            pr->q_slot.alloc(ns);
            pr->q_feat.alloc(ns * kGuideScalars);
            pr->q_out.alloc(ns * kLobeOut);
            pr->grec.alloc(ns * size_t(p.depth + 1) * 3);
            pr->train.alloc(size_t(pr->train_cap) * kGuideSample);
            pr->Q = GuideQueue{pr->q_slot.p, pr->q_feat.p, pr->q_out.p};
            pr->R->P.grec = pr->grec.p;
        }
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
        HIP_OK(hipEventRecord(ev.a, R.st));
        const bool clear = !keep || pr->total == 0;
        if (pr->net && pr->fused)
            R.run(F, kAutoTailPaths, nullptr, nullptr, clear, pr->net, &pr->Q, true);
        else if (pr->net && pr->guide_bounces <= pr->params.depth)
            R.run_guided_batches(F, nullptr, nullptr, clear, pr->net, &pr->Q);
        else
            R.run(F, pr->net ? 0 : kAutoTailPaths, nullptr, nullptr, clear, pr->net, pr->net ? &pr->Q : nullptr);
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
            for (int e = 0; e < pr->epochs; e++) net.train(pr->train.p, int(collected), pr->batch, R.st);
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
        k_resolve<<<grid_for(F.npix), kBlock, 0, R.st>>>(F, R.P.fb, R.B.rgb.p);
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
