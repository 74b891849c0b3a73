// Shading passes of the per-bounce schedule, plain and guided.
#include "integrator.h"
#include "launch.h"

namespace lam {
namespace {

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


}  // namespace

void launch_shade(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int count, const int *counts) {
    k_shade<<<grid, block, 0, st>>>(S, F, P, count, counts);
}

void launch_shade_guided_queue(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int count, GuideQueue Q) {
    k_shade_guided_queue<<<grid, block, 0, st>>>(S, F, P, count, Q);
}

void launch_shade_guided(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, GuideQueue Q) {
    k_shade_guided<<<grid, block, 0, st>>>(S, F, P, Q);
}

}  // namespace lam
