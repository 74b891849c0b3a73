// Closest-hit, occlusion and thin-film queries (checks (a) and (b)).
#include <mutex>

#include "integrator.h"
#include "launch.h"

namespace lam {
namespace {

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

// The no-allocation closest-hit query: a fixed grid of kNoAllocThreads
// threads takes the rays grid-stride, and the traversal stacks' global
// entries go to `scratch` below instead of a buffer allocated per call.
__global__ void __launch_bounds__(kBlock) k_query_closest_fixed(DeviceScene S, const float *rays, int n,
                                                                lambda_hit *out, int2 *spill) {
    __shared__ int s_code[kRing * kBlock];
    __shared__ float s_dist[kRing * kBlock];
    Stack stk = make_stack(s_code, s_dist, spill);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        V3 o = v3(rays[6 * i + 0], rays[6 * i + 1], rays[6 * i + 2]);
        V3 d = v3(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
        float best = kMax;
        int hi = -1, ho = -1;
        stk.sp = stk.lo = 0;
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
}

// Global traversal-stack entries of k_query_closest_fixed, [entry][thread]
// (16 MiB per device, held until the process ends).  Host code reaches it
// only through the address hipGetSymbolAddress returns for the current
// device (noalloc_scratch), never through this name.
static __device__ __align__(256) int2 scratch[size_t(kNoAllocSpill) * kNoAllocThreads];

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


}  // namespace

int2 *noalloc_scratch() {
    static std::mutex lock;
    static int2 *addr[kNoAllocMaxDevices] = {};
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) throw Failure{LAMBDA_EINVAL, "device index beyond the supported range"};
    std::lock_guard<std::mutex> g(lock);
    if (!addr[dev]) {
        void *p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Failure{LAMBDA_EHIP, "no device address for the query scratch area"};
        addr[dev] = static_cast<int2 *>(p);
    }
    return addr[dev];
}

void launch_query_closest(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, const float *rays, int n, lambda_hit *out, int2 *spill) {
    k_query_closest<<<grid, block, 0, st>>>(S, rays, n, out, spill);
}

void launch_query_closest_fixed(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, const float *rays, int n, lambda_hit *out, int2 *spill) {
    k_query_closest_fixed<<<grid, block, 0, st>>>(S, rays, n, out, spill);
}

void launch_query_occluded(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, const float *rays, const int *li, const int *lo, int n, int *out, int2 *spill) {
    k_query_occluded<<<grid, block, 0, st>>>(S, rays, li, lo, n, out, spill);
}

void launch_query_thinfilm(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, int mat, const float *q, const float *pts, int n, float *out) {
    k_query_thinfilm<<<grid, block, 0, st>>>(S, mat, q, pts, n, out);
}

}  // namespace lam
