// The persistent per-thread schedule, compiled for kWaves waves per SIMD
// (its registers are capped so that this many fit): k_persist.hip builds the
// default variant, k_persist_occ.hip the others.
#pragma once

#include "integrator.h"

namespace lam {
namespace {

// Persistent per-thread schedule: each thread takes one live slot and runs
// it -- closest hit, shading, the occlusion ray at once, retirement, the
// slot's next sample, then further work units -- without waiting for other
// paths.  Threads [0, n_in) take q_in (paths that need their closest hit);
// the next P.ctr[kCtrNext] threads take q_out (samples just started by
// k_finalize).  No occlusion ray or ended sample may be pending when it
// starts.  Threads exit when no work unit is left.
template <int kWaves>
__global__ void __launch_bounds__(kBlock, kWaves) k_persist(DeviceScene S, Frame F, Paths P, int n_in, int2 *spill) {
    __shared__ int s_code[kRing * kBlock];
    __shared__ float s_dist[kRing * kBlock];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_new = P.ctr[kCtrNext];
    if (t >= n_in + n_new) return;
    const int s = t < n_in ? P.q_in[t] : P.q_out[t - n_in];
    Stack stk = make_stack(s_code, s_dist, spill);
    unsigned n_closest = 0, n_shadow = 0;
    if (F.cost) F.unit_t0[s] = wall_clock64();
    // Each iteration traces the occlusion ray of the previous vertex (if
    // any) and then the closest hit of the path's next ray, in one
    // traversal loop; a path that ended with an occlusion ray pending skips
    // its closest hit for one iteration and retires after the occlusion
    // ray.  Radiance is summed in the same order as the reference's trace
    // loop: the occlusion ray's contribution lands before the next vertex's.
    bool pend = false;   // an occlusion ray of the previous vertex waits
    bool alive = true;   // the path has a closest-hit ray to trace
    ShadowRay sh{};
    while (true) {
        int2 hit = make_int2(-1, -1);
        {
            bool blocked;
            float best;
            int hi, ho;
            traverse_pair<kRing>(S, pend, v3(sh.o), v3(sh.d), sh.o.w, sh.i.y, sh.i.z, blocked, alive, P.o + s,
                                 P.d + s, best, hi, ho, stk);
            if (pend) {
                n_shadow++;
                if (!blocked) {
                    float4 L = P.L[s];
                    L.x += sh.c.x;
                    L.y += sh.c.y;
                    L.z += sh.c.z;
                    L.w += sh.c.w;
                    P.L[s] = L;
                }
                pend = false;
            }
            if (alive) {
                n_closest++;
                hit = make_int2(hi, ho);
                P.hit[s] = hit;
            }
        }
        if (alive) {
            bool push = false;
            const bool ended = shade_slot(S, F, P, s, push, sh, nullptr, nullptr, nullptr, &hit);
            pend = push;
            if (!ended) continue;
            if (pend) {
                alive = false;
                continue;
            }
        }
        alive = true;
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

}  // namespace
}  // namespace lam
