// Host functions that launch the path kernels, one per kernel, each defined
// in the translation unit of its kernel (k_*.hip): they take the grid, the
// block and the stream, then the kernel's own arguments.
#pragma once

#include "integrator.h"

namespace lam {

void launch_start(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int nslots, long long base);
void launch_pass_begin(dim3 grid, dim3 block, hipStream_t st, const int *ctr, int *counts);
void launch_trace(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Paths P, int n_closest, int n_shadow, int2 *spill, const int *counts);
void launch_finalize(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_done, const int *counts);
void launch_serve_clock(dim3 grid, dim3 block, hipStream_t st, ServeCtl *ctl);
void launch_resolve(dim3 grid, dim3 block, hipStream_t st, Frame F, const double *fb, float *rgb);
void launch_set_i(dim3 grid, dim3 block, hipStream_t st, int *p, int v);
void launch_fill_d(dim3 grid, dim3 block, hipStream_t st, double *p, size_t n);
void launch_fill_u64(dim3 grid, dim3 block, hipStream_t st, unsigned long long *p, size_t n);
void launch_fill_i(dim3 grid, dim3 block, hipStream_t st, int *p, int n);
void launch_tile_cost(dim3 grid, dim3 block, hipStream_t st, const int *cost, int w, int h, int tile, long long *tile_cost);
void launch_tile_expand(dim3 grid, dim3 block, hipStream_t st, const int *tiles, const int *offsets, int ntiles, int w, int h, int tile, int *order);
void launch_shade(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int count, const int *counts);
void launch_shade_guided_queue(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int count, GuideQueue Q);
void launch_shade_guided(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, GuideQueue Q);
// The persistent kernel variant with `occupancy` waves per SIMD (0: the
// default, kPersistWaves); blocks of kBlock threads resident per compute unit.
void launch_persist(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_in, int2 *spill,
                    int occupancy);
int persist_blocks_per_cu(int occupancy);
void launch_persist_occ(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_in, int2 *spill,
                        int occupancy);
int persist_occ_blocks_per_cu(int occupancy);
// The voting variant of the persistent kernel (3 or 4 waves per SIMD).
void launch_persist_vote(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_in,
                         int2 *spill, int occupancy);
int persist_vote_blocks_per_cu(int occupancy);
// Bit of a Renderer's persistent-kernel variant code that selects the
// voting schedule (the low bits: waves per SIMD, 0 = default).
constexpr int kPersistVoteBit = 0x100;
void launch_persist_guided(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_in, int2 *spill, GuideDev G);
int persist_guided_blocks_per_cu();
void launch_persist_table(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, Frame F, Paths P, int n_in, int2 *spill,
                          GuideTableDev T);
int persist_table_blocks_per_cu();
void launch_serve(dim3 grid, dim3 block, hipStream_t st, const DeviceScene *scenes, const ServeJob *jobs, int njobs, long long total_units, Paths P, ServeCtl *ctl, int *done, unsigned long long *finish, int2 *spill);
void launch_query_closest(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, const float *rays, int n, lambda_hit *out, int2 *spill);
// Device address of the no-allocation query scratch on the current device
// (looked up once per device).
int2 *noalloc_scratch();
void launch_query_closest_fixed(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, const float *rays, int n, lambda_hit *out, int2 *spill);
void launch_query_occluded(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, const float *rays, const int *li, const int *lo, int n, int *out, int2 *spill);
void launch_query_thinfilm(dim3 grid, dim3 block, hipStream_t st, DeviceScene S, int mat, const float *q, const float *pts, int n, float *out);

}  // namespace lam
