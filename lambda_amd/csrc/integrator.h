// Device side of the path tracer shared by every kernel translation unit:
// path state (structure of arrays), frame settings, the per-vertex shading
// step (render/renderer.h:102-192), sample retirement into the FP64 XYZ
// framebuffer, traversal-stack set-up and wave-wide queue pushes.  Each
// kernel file includes it and defines its kernels plus the host functions
// that launch them (declared in launch.h).
#pragma once

#include "dscene.h"
#include "guide.h"
#include "guide_mlp.h"
#include "runtime.h"

namespace lam {

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
    float guide_frac;         // probability of sampling the guide at a guided vertex
    int rr_start;             // bounce index from which Russian roulette applies
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
constexpr int kPersistWaves = 4;
// Waves per SIMD of the table-guided persistent kernel and of the serving
// kernel (measured on gfx950: 4 and 3 best).
constexpr int kPersistTableWaves = 4;
constexpr int kServeWaves = 3;
// Automatic schedule: live paths below which a render switches to the
// persistent schedule (above any slot count: persistent from the first pass
// on).  Slots default to the number of persistent threads resident at once.
constexpr long long kAutoTailPaths = 1LL << 40;
using Stack = TraversalStack<kRing>;
// No-allocation closest-hit queries: a fixed grid of this many threads, and
// scratch for the traversal-stack entries beyond the LDS ring of scenes
// whose stacks are at most kNoAllocMaxStack entries deep.
constexpr int kNoAllocThreads = 65536;
constexpr int kNoAllocMaxStack = 48;
constexpr int kNoAllocSpill = kNoAllocMaxStack - kRing;
constexpr int kNoAllocMaxDevices = 64;

__device__ __forceinline__ Stack make_stack(int *s_code, float *s_dist, int2 *spill) {
    Stack s;
    s.lcode = (LAM_LDS int *)(s_code + threadIdx.x);
    s.ldist = (LAM_LDS float *)(s_dist + threadIdx.x);
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

// Default probability of sampling the guide at a guided vertex (Frame::
// guide_frac); the cosine lobe takes the rest, and both MIS densities use
// this one-sample mixture.
constexpr float kGuideFrac = 0.5f;
// Largest accepted guide fraction: the cosine lobe keeps at least 5 % so that
// directions where every vMF lobe's density underflows are still sampled.
constexpr float kGuideFracMax = 0.95f;

// One iteration body of Renderer::trace (render/renderer.h:102-192) for the
// path in slot s after its closest-hit query (P.hit[s]).  Updates the slot's
// state; returns true when the sample's path has ended.  *push_shadow tells
// whether next-event estimation produced an occlusion ray.
//
// kGuided with gout set: a Lambertian vertex samples the next direction from
// the one-sample mixture (1 - a) cos/pi + a q (a = F.guide_frac, q: the guiding distribution in
// the compact form of guide_compact, features gfeat); the mixture density is the
// scatter pdf of both MIS weights (next-event estimation and emitter hits),
// so the estimator stays unbiased.  The vertex is recorded for training in
// P.grec.  Without gout the vertex is shaded as in the plain tracer.  With
// a table instead, a Lambertian vertex with bounce index below
// F.guide_bounces takes its mixture from its table entry, and its features
// are computed only when the vertex is recorded.
template <bool kGuided = false>
__device__ __forceinline__ bool shade_slot(const DeviceScene &S, const Frame &F, const Paths &P, int s,
                                           bool &push_shadow, ShadowRay &sh, const float *gout = nullptr,
                                           const float *gfeat = nullptr, const GuideTableDev *table = nullptr,
                                           const int2 *hit_now = nullptr) {
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
        // the caller may hand over the hit it has just found (and stored)
        int2 hit = hit_now ? *hit_now : P.hit[s];
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
            // the material indices are requested with the hit's records
            const int inst_m = S.inst_i[hit.x].w, obj_m = S.obj_mat[hit.y];
            Hit h;
            const bool rec = hit_record(S, hit.x, hit.y, o, d, h);
            keep(inst_m);
            keep(obj_m);
            if (!rec) {
                atomicOr(&P.ctr[kCtrError], 1);
                finished = true;
            } else {
                const int mat = inst_m >= 0 ? inst_m : obj_m;   // material_of
                const int4 mt = S.mat[mat];
                const int4 mtex = S.mat_tex[mat];
                if (mt.x == LAMBDA_MATERIAL_EMISSIVE) {
                    // Emitter hit (render/renderer.h:119-131).
                    float w = 1;
                    if (!specular && total > 0) {
                        float lp = instance_area(S, hit.x, hit.y) * S.tex_avg[mt.y];
                        w = power_heuristic(prev_pdf, instance_pdf(S, hit.x, hit.y, o, d) * lp / total);
                    }
                    S4 e = spectrum_texture4_rec(S, mtex, h, lam);
                    S4 c = beta * e * w;
#pragma unroll
                    for (int k = 0; k < kHero; k++) L.v[k] += c.v[k];
                    finished = true;
                } else {
                    const bool lambertian = mt.x == LAMBDA_MATERIAL_LAMBERTIAN;
                    S4 albedo = s4(0);
                    if (lambertian) albedo = spectrum_texture4_rec(S, mtex, h, lam);
                    if constexpr (kGuided) {
                        if (table && lambertian && bounce < F.guide_bounces)
                            gout = table->rows + guide_table_entry(*table, h.p, h.n) * kLobeOut;
                    }
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
                                    if (guided) sp = (1 - F.guide_frac) * sp + F.guide_frac * guide_pdf_c(gout, ldir);
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
                            if (guided && rng.next() >= 1 - F.guide_frac) {
                                float u0 = rng.next(), u1 = rng.next(), u2 = rng.next();
                                nd = guide_sample_c(gout, u0, u1, u2);
                                from_guide = true;
                            }
                        }
                        if (!from_guide) nd = random_in_hemisphere(h.n, rng);
                        att = dot(h.n, nd) <= 0 ? s4(0) : albedo / kPi;
                    } else if (mt.x == LAMBDA_MATERIAL_MIRROR) {
                        nd = reflected(d, h.n);
                        att = spectrum_texture4_rec(S, mtex, h, lam);
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
                            prev_pdf = (1 - F.guide_frac) * prev_pdf + F.guide_frac * guide_pdf_c(gout, d);
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
                    if (bounce >= F.rr_start) {
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
                            GuideFeatures tf;
                            if (table) {
                                tf = guide_features(h.p, v3(F.box_lo.x, F.box_lo.y, F.box_lo.z),
                                                    v3(F.box_hi.x, F.box_hi.y, F.box_hi.z), h.n, lam.v[0]);
                                gfeat = tf.f;
                            }
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

// Paths queued for the guiding network in a guided pass.
struct GuideQueue {
    int *slot;       // [n] slot of each queued path
    float *feat;     // [n][6] features
    float *out;      // [n][40] network outputs
};

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

// Waves per SIMD of the guided persistent kernel (its LDS, stacks plus the
// network's staging, allows two blocks per CU).  Staging in the traversal
// stacks' LDS instead, at 3 waves (168 VGPRs), spilled 158 registers and
// ran 4 % slower on the guiding room.
constexpr int kPersistGuidedWaves = 2;

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

}  // namespace lam
