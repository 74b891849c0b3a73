// The persistent per-thread schedule, compiled for kWaves waves per SIMD
// (its registers are capped so that this many fit): k_persist.hip builds the
// default variant, k_persist_occ.hip the others, k_persist_vote.hip the
// voting variants (kVote).
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
//
// kVote: the wave alternates between traversal steps and shading.  Each
// lane walks its rays one wide node per step (RayWalk): first the
// occlusion ray of the previous vertex (if any), then the closest hit of
// the path's next ray.  A lane whose closest hit is found waits; the wave
// leaves the traversal phase when at least half of its lanes wait (or none
// traverses), and the waiting lanes then shade, retire and start their
// next rays together, so lanes with short traversals do not idle through
// the wave's longest one.  Same results; faster with independent streams,
// slower with per-pixel sample chains (a slow chain's lane waits through
// every shading phase of the others).
template <int kWaves, bool kVote = false>
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
    if constexpr (kVote) {
        RayWalk wk;
        {
            const float4 o = P.o[s], d = P.d[s];
            wk.begin(false, v3(o), v3(d), kMax, -1, -2);
        }
        stk.sp = stk.lo = 0;
        bool trav = true;    // the lane has a ray in its walk
        bool alive = true;   // the path's next ray still needs its closest hit
        bool pend = false;   // an occlusion ray's result waits to be added
        bool blocked = false;
        float4 sh_c = make_float4(0, 0, 0, 0);
        while (true) {
            // traversal steps while fewer than half of the wave's lanes wait
            while (true) {
                const int nt = __popcll(__ballot(trav));
                if (nt == 0) break;
                const int na = __popcll(__ballot(1));
                if (na - nt > 0 && na - nt >= (na >> 1)) break;
                if (trav && wk.step<kRing>(S, stk)) {
                    if (wk.any) {
                        blocked = wk.blocked;
                        if (alive) {
                            const float4 o = P.o[s], d = P.d[s];
                            wk.begin(false, v3(o), v3(d), kMax, -1, -2);
                            stk.sp = stk.lo = 0;
                        } else {
                            trav = false;
                        }
                    } else {
                        trav = false;
                    }
                }
            }
            if (trav) continue;
            if (pend) {
                n_shadow++;
                if (!blocked) {
                    float4 L = P.L[s];
                    L.x += sh_c.x;
                    L.y += sh_c.y;
                    L.z += sh_c.z;
                    L.w += sh_c.w;
                    P.L[s] = L;
                }
                pend = false;
            }
            bool retire = !alive;
            if (alive) {
                n_closest++;
                const int2 hit = make_int2(wk.hi, wk.ho);
                P.hit[s] = hit;
                bool push = false;
                ShadowRay sh;
                const bool ended = shade_slot(S, F, P, s, push, sh, nullptr, nullptr, nullptr, &hit);
                if (push) {
                    // the occlusion ray first; a path that ended walks only it
                    pend = true;
                    sh_c = sh.c;
                    if (ended) alive = false;
                    wk.begin(true, v3(sh.o), v3(sh.d), sh.o.w, sh.i.y, sh.i.z);
                    stk.sp = stk.lo = 0;
                    trav = true;
                } else if (!ended) {
                    const float4 o = P.o[s], d = P.d[s];
                    wk.begin(false, v3(o), v3(d), kMax, -1, -2);
                    stk.sp = stk.lo = 0;
                    trav = true;
                } else {
                    retire = true;
                }
            }
            if (!retire) continue;
            alive = true;
            if (!retire_sample(S, F, P, s)) {
                if (F.cost) {
                    const unsigned long long now = wall_clock64();
                    const unsigned long long dt = now - F.unit_t0[s];
                    atomicAdd(F.cost + P.info[s].x, int(dt < (1ull << 30) ? dt : (1ull << 30)));
                    F.unit_t0[s] = now;
                }
                const int w = atomicAdd(&P.ctr[kCtrWork], 1);
                if (w >= work_units(F)) break;
                start_unit(S, F, P, s, w);
            }
            const float4 o = P.o[s], d = P.d[s];
            wk.begin(false, v3(o), v3(d), kMax, -1, -2);
            stk.sp = stk.lo = 0;
            trav = true;
        }
    } else {
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
    }
    atomicAdd(&P.rays[0], (unsigned long long)n_closest);
    atomicAdd(&P.rays[1], (unsigned long long)n_shadow);
}

}  // namespace
}  // namespace lam
