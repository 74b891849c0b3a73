// The persistent schedule with the guiding network evaluated inside the kernel.
#include "integrator.h"
#include "launch.h"

namespace lam {
namespace {

// mlp_rows64 as a call of its own, for kernels whose other code is register
// bound (the persistent guided kernel): the caller's live registers are
// saved once around the call instead of the network's working set spilling
// into the caller's loops.
__device__ __attribute__((noinline)) void mlp_rows64_call(const GuideDev &G, const float *feat, int fstride,
                                                          int nvalid, float *out, int ostride) {
    mlp_rows64(G, feat, fstride, nvalid, out, ostride);
}


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


}  // namespace

void launch_persist_guided(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_in, int2 *spill, GuideDev G) {
    k_persist_guided<<<grid, block, 0, st>>>(S, F, P, n_in, spill, G);
}

int persist_guided_blocks_per_cu() {
    int b = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_persist_guided, kBlock, 0));
    return b;
}

}  // namespace lam
