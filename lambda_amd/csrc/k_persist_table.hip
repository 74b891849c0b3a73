// The persistent schedule guided by the tabulated network (GuideTableDev):
// each guided vertex reads its mixture from the table instead of evaluating
// the network, so the kernel needs no LDS rows and no matrix-core code and
// keeps the plain kernel's occupancy.
#include "integrator.h"
#include "launch.h"

namespace lam {
namespace {

// k_persist with guided vertices: shading looks up the table entry of each
// guided vertex (Lambertian, bounce index below F.guide_bounces).
__global__ void __launch_bounds__(kBlock, kPersistTableWaves) k_persist_table(DeviceScene S, Frame F, Paths P, int n_in,
                                                                         int2 *spill, GuideTableDev T) {
    __shared__ int s_code[kRing * kBlock];
    __shared__ float s_dist[kRing * kBlock];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n_new = P.ctr[kCtrNext];
    if (t >= n_in + n_new) return;
    const int s = t < n_in ? P.q_in[t] : P.q_out[t - n_in];
    Stack stk = make_stack(s_code, s_dist, spill);
    unsigned n_closest = 0, n_shadow = 0;
    // the loop of k_persist: occlusion ray and next closest hit in one
    // traversal, radiance summed in the reference's order
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
                // a record left by an earlier sample in this slot is void
                // until shading writes this vertex's
                const int bounce = P.info[s].z;
                if (bounce < F.grec_stride && bounce < F.guide_bounces)
                    P.grec[(size_t(s) * F.grec_stride + bounce) * 3 + 2].w = kRecNone;
            }
        }
        if (alive) {
            bool push = false;
            const bool ended = shade_slot<true>(S, F, P, s, push, sh, nullptr, nullptr, &T, &hit);
            pend = push;
            if (!ended) continue;
            if (pend) {
                alive = false;
                continue;
            }
        }
        alive = true;
        if (retire_sample(S, F, P, s)) continue;
        const int w = atomicAdd(&P.ctr[kCtrWork], 1);
        if (w >= work_units(F)) break;
        start_unit(S, F, P, s, w);
    }
    atomicAdd(&P.rays[0], (unsigned long long)n_closest);
    atomicAdd(&P.rays[1], (unsigned long long)n_shadow);
}

}  // namespace

void launch_persist_table(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_in,
                          int2 *spill, GuideTableDev T) {
    k_persist_table<<<grid, block, 0, st>>>(S, F, P, n_in, spill, T);
}

int persist_table_blocks_per_cu() {
    int b = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_persist_table, kBlock, 0));
    return b;
}

}  // namespace lam
