// Batched serving: many jobs in one persistent launch.
#include "integrator.h"
#include "launch.h"

namespace lam {
namespace {

__global__ void __launch_bounds__(kBlock, kServeWaves) k_serve(const DeviceScene *scenes, const ServeJob *jobs,
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
        // the pixel's sample chain, as k_persist runs it
        bool pend = false;
        bool alive = true;
        ShadowRay sh{};
        while (true) {
            int2 hit = make_int2(-1, -1);
            {
                bool blocked;
                float best;
                int hi, ho;
                traverse_pair<kRing>(S, pend, v3(sh.o), v3(sh.d), sh.o.w, sh.i.y, sh.i.z, blocked, alive, Pj.o + s,
                                     Pj.d + s, best, hi, ho, stk);
                if (pend) {
                    n_shadow++;
                    if (!blocked) {
                        float4 L = Pj.L[s];
                        L.x += sh.c.x;
                        L.y += sh.c.y;
                        L.z += sh.c.z;
                        L.w += sh.c.w;
                        Pj.L[s] = L;
                    }
                    pend = false;
                }
                if (alive) {
                    n_closest++;
                    hit = make_int2(hi, ho);
                    Pj.hit[s] = hit;
                }
            }
            if (alive) {
                bool push = false;
                const bool ended = shade_slot(S, F, Pj, s, push, sh, nullptr, nullptr, nullptr, &hit);
                pend = push;
                if (!ended) continue;
                if (pend) {
                    alive = false;
                    continue;
                }
            }
            alive = true;
            if (retire_sample(S, F, Pj, s)) continue;
            break;
        }
        if (atomicAdd(&done[j], 1) == J.units - 1) finish[j] = wall_clock64();
    }
    atomicAdd(&P.rays[0], (unsigned long long)n_closest);
    atomicAdd(&P.rays[1], (unsigned long long)n_shadow);
}


}  // namespace

void launch_serve(dim3 grid, dim3 block, hipStream_t st, const DeviceScene *scenes, const ServeJob *jobs, int njobs, long long total_units, Paths P, ServeCtl *ctl, int *done, unsigned long long *finish, int2 *spill) {
    k_serve<<<grid, block, 0, st>>>(scenes, jobs, njobs, total_units, P, ctl, done, finish, spill);
}

}  // namespace lam
