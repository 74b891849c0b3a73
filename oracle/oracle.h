/*
 * Test infrastructure only: the CPU oracle of the spectral path tracer.
 *
 * A C++ restatement of Lambda's CPU integrator (render/renderer.h:42-195 and
 * everything it calls) that reads the same lambda_scene_view the HIP layer
 * reads.  It reproduces the reference's float RGB buffer bit for bit (same
 * operation order, no FMA, glibc libm, libstdc++ complex<double>), so the GPU
 * path can be checked against it on identical inputs.  Only tests, the smoke
 * check and the benchmark's CPU-baseline leg use it; the product never does.
 */
#ifndef LAMBDA_ORACLE_H
#define LAMBDA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "lambda_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Renders rows [row_begin, row_end) into out_rgb (the full W*H*3 buffer) with
 * an OpenMP guided schedule over rows (render/renderer.h:42-55), PCG32 stream
 * (seed, py*W+px) per pixel.  threads <= 0 uses the OpenMP default.
 * Returns 0, or nonzero for a malformed scene. */
int oracle_render(const lambda_scene_view *scene, const lambda_render_params *params, float *out_rgb,
                  int32_t row_begin, int32_t row_end, int32_t threads);

/* Closest-hit queries (scene/scene.h:51-124, closest-hit mode). */
int oracle_closest_hit(const lambda_scene_view *scene, const float *rays, int32_t n, lambda_hit *hits);

/* Occlusion queries toward light (instance, object); -1/-1 = environment. */
int oracle_occluded(const lambda_scene_view *scene, const float *rays, const int32_t *light_instance,
                    const int32_t *light_object, int32_t n, int32_t *occluded);

/* Thin-film reflectance of material `material` for queries[n][7] =
 * {cos_theta, lambda, front_facing, px, py, pz, unused}, normal (0,0,1) and
 * direction (sin, 0, -cos).  out_ref[i] is the reference's value (complex
 * double arithmetic, magnitudes rounded to float at the end,
 * scene/material.h:302-308); out_fp64[i] keeps every step in double.
 * out_ambiguous[i] (may be NULL) is 1 when a complex square root met its
 * branch cut up to a rounding residue (absorbing layers under total internal
 * reflection), so the reference's branch is decided by rounding noise. */
int oracle_thinfilm(const lambda_scene_view *scene, int32_t material, const float *queries, int32_t n,
                    float *out_ref, double *out_fp64, int32_t *out_ambiguous);

/* n PCG32 outputs of stream (seed, stream) (math/random.h:9-31). */
void oracle_pcg32(uint64_t seed, uint64_t stream, int32_t n, uint32_t *out);

/* Traversal work of the reference path over every pixel_stride-th pixel of
 * [pixel_begin, pixel_end), reference RNG streams, all samples:
 * counts[0] paths, [1] closest-hit rays, [2] shadow rays;
 * counts[3..7] closest-hit work and counts[8..12] shadow-ray work, each as
 * node tests, instance entries (instance-level leaves, plus the light
 * pre-intersection of shadow rays), sphere, quad and triangle tests. */
int oracle_path_stats(const lambda_scene_view *scene, const lambda_render_params *params,
                      int64_t pixel_begin, int64_t pixel_end, int64_t pixel_stride, int32_t threads,
                      int64_t *counts);

#ifdef __cplusplus
}
#endif

#endif /* LAMBDA_ORACLE_H */
