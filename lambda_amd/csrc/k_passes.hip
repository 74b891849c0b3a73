// One pass over the in-flight paths, launch per bounce: new samples, closest-hit and occlusion
// traversal, retirement; framebuffer resolve, fills and the cost-order tile kernels.
#include "integrator.h"
#include "launch.h"

namespace lam {
namespace {

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

__global__ void k_serve_clock(ServeCtl *ctl) {
    ctl->t0 = wall_clock64();
    ctl->next_unit = 0;
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

// Visiting order from measured pixel costs: 8x8 tiles (the unit of the
// default order, which keeps neighbouring paths in flight together) are
// summed on the device, sorted on the host (longest first) and expanded back
// to pixels, row order within a tile.  The order changes when a pixel runs,
// never what it computes.
__global__ void k_tile_cost(const int *cost, int w, int h, int tile, long long *tile_cost) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= w * h) return;
    const int tx = (p % w) / tile, ty = (p / w) / tile;
    atomicAdd(reinterpret_cast<unsigned long long *>(tile_cost + size_t(ty) * ((w + tile - 1) / tile) + tx),
              (unsigned long long)cost[p]);
}

// Thread (i, k): pixel k of the tile at position i of the sorted list.
__global__ void k_tile_expand(const int *tiles, const int *offsets, int ntiles, int w, int h, int tile, int *order) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int area = tile * tile;
    const int i = int(e / area), k = int(e % area);
    if (i >= ntiles) return;
    const int t = tiles[i], tiles_x = (w + tile - 1) / tile;
    const int tx = t % tiles_x, ty = t / tiles_x;
    const int tw = min(tile, w - tile * tx), th = min(tile, h - tile * ty);
    if (k >= tw * th) return;
    order[offsets[i] + k] = (tile * ty + k / tw) * w + tile * tx + k % tw;
}


}  // namespace

void launch_start(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int nslots, long long base) {
    k_start<<<grid, block, 0, st>>>(S, F, P, nslots, base);
}

void launch_pass_begin(dim3 grid, dim3 block, hipStream_t st, const int *ctr, int *counts) {
    k_pass_begin<<<grid, block, 0, st>>>(ctr, counts);
}

void launch_trace(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Paths P, int n_closest, int n_shadow, int2 *spill, const int *counts) {
    k_trace<<<grid, block, 0, st>>>(S, P, n_closest, n_shadow, spill, counts);
}

void launch_finalize(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_done, const int *counts) {
    k_finalize<<<grid, block, 0, st>>>(S, F, P, n_done, counts);
}

void launch_serve_clock(dim3 grid, dim3 block, hipStream_t st, ServeCtl *ctl) {
    k_serve_clock<<<grid, block, 0, st>>>(ctl);
}

void launch_resolve(dim3 grid, dim3 block, hipStream_t st, Frame F, const double *fb, float *rgb) {
    k_resolve<<<grid, block, 0, st>>>(F, fb, rgb);
}

void launch_set_i(dim3 grid, dim3 block, hipStream_t st, int *p, int v) {
    k_set_i<<<grid, block, 0, st>>>(p, v);
}

void launch_fill_d(dim3 grid, dim3 block, hipStream_t st, double *p, size_t n) {
    k_fill_d<<<grid, block, 0, st>>>(p, n);
}

void launch_fill_u64(dim3 grid, dim3 block, hipStream_t st, unsigned long long *p, size_t n) {
    k_fill_u64<<<grid, block, 0, st>>>(p, n);
}

void launch_fill_i(dim3 grid, dim3 block, hipStream_t st, int *p, int n) {
    k_fill_i<<<grid, block, 0, st>>>(p, n);
}

void launch_tile_cost(dim3 grid, dim3 block, hipStream_t st, const int *cost, int w, int h, int tile, long long *tile_cost) {
    k_tile_cost<<<grid, block, 0, st>>>(cost, w, h, tile, tile_cost);
}

void launch_tile_expand(dim3 grid, dim3 block, hipStream_t st, const int *tiles, const int *offsets, int ntiles, int w, int h, int tile, int *order) {
    k_tile_expand<<<grid, block, 0, st>>>(tiles, offsets, ntiles, w, h, tile, order);
}

}  // namespace lam
