/*
 * lambda_hip.h - C ABI of the MI355X (gfx950) spectral path-tracing integrator.
 *
 * This is the drop-in boundary for Lambda's render call.  In the reference the
 * host parses a scene (core/parser.h:29), uploads the flat arrays
 * (src/main.cpp:84-109), launches renderKernel / Renderer::renderImage
 * (src/main.cpp:113-127, render/renderer.h:42-55,200-212), polls a progress
 * counter (src/main.cpp:129-146) and copies back a W*H*3 float buffer of
 * linear sRGB averaged over the samples (src/main.cpp:148-156,
 * render/renderer.h:87-91).  lambda_hip_render() replaces exactly that block.
 *
 * Everything crossing this boundary is plain old data: counts, pointers to
 * host arrays and sizes.  The arrays describe the scene after the host has
 * parsed it and built its BVH, in flat layouts of this library's own (no C++
 * unions).  The caller owns every array; the library only reads them.
 *
 * Return value of every int-returning entry point: 0 on success, otherwise a
 * LAMBDA_E* code, with a NUL-terminated message written to err[0..err_len).
 * Nothing throws across this boundary.
 */
#ifndef LAMBDA_HIP_H
#define LAMBDA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Wavelength grid of every dense spectrum: 1 nm bins over 360..830 nm
 * (core/constants.h:20-22). */
#define LAMBDA_CIE_LAMBDA_MIN 360
#define LAMBDA_CIE_LAMBDA_MAX 830
#define LAMBDA_CIE_BINS 471

/* Deepest BVH the traversal kernels accept (bvh.h:221 caps leaves at depth 20,
 * which bounds a depth-first stack at 20 entries, scene.h:52-57). */
#define LAMBDA_MAX_STACK 20

enum lambda_error {
    LAMBDA_OK = 0,
    LAMBDA_EINVAL = 1,    /* malformed scene or parameters */
    LAMBDA_EHIP = 2,      /* HIP runtime error */
    LAMBDA_ENOMEM = 3,    /* device or host allocation failed */
    LAMBDA_ENODEV = 4,    /* no gfx950 device visible */
    LAMBDA_EIO = 5,       /* file could not be read or written */
    LAMBDA_EPARSE = 6,    /* scene file rejected by the parser */
    LAMBDA_EBUSY = 7      /* batch queue full */
};

/* Object kinds (scene/object.h:13). */
enum lambda_object_type { LAMBDA_OBJECT_SPHERE = 0, LAMBDA_OBJECT_QUAD = 1, LAMBDA_OBJECT_TRI = 2 };

/* Material kinds (scene/material.h:15). */
enum lambda_material_type {
    LAMBDA_MATERIAL_LAMBERTIAN = 0,
    LAMBDA_MATERIAL_MIRROR = 1,
    LAMBDA_MATERIAL_DIELECTRIC = 2,
    LAMBDA_MATERIAL_EMISSIVE = 3,
    LAMBDA_MATERIAL_THINFILM = 4
};

/* Scalar texture kinds (scene/texture.h:8). */
enum lambda_scalar_texture_type {
    LAMBDA_SCALAR_CONSTANT = 0,
    LAMBDA_SCALAR_PERLIN = 1,
    LAMBDA_SCALAR_WORLEY = 2
};

/* Spectrum texture kinds (scene/texture.h:9). */
enum lambda_spectrum_texture_type {
    LAMBDA_SPECTRUM_CONSTANT = 0,
    LAMBDA_SPECTRUM_CHECKER = 1,
    LAMBDA_SPECTRUM_SCALAR = 2
};

/*
 * A parsed scene with its BVH, as flat arrays.  Field groups mirror the
 * arguments of Renderer::setScene (render/renderer.h:38) and Camera
 * (render/camera.h:10):
 *
 *  spectra            [num_spectra][471] float      DenseSpectrum<Float>
 *  complex_spectra    [num_complex_spectra][471][2] double (re, im)
 *  scalar_texture_*   type[n]; params[n][3] = {value,0,0} | {min,max,frequency}
 *  spectrum_texture_* type[n]; index[n][2] = {spectrum,0} | {spectrum1,spectrum2}
 *                     | {scalar texture,0}; scale[n] (checker only)
 *  material_*         type[n]; params[n][3]:
 *                       lambertian/mirror {albedo spectrum texture}
 *                       dielectric {n0 spectrum, n1 spectrum}
 *                       emissive {emission spectrum texture}
 *                       thinfilm {numLayers, offset of n, offset of d} into
 *                       material_properties (parser.h:183-210)
 *  object_*           type[n]; material[n]; geometry[n][12]:
 *                       sphere {cx,cy,cz, r, 0...}
 *                       quad/tri {corner[3], horizontal[3], vertical[3],
 *                                 normal[3] = cross(horizontal, vertical)}
 *                     shading[n][15] (triangles): {n0[3], n1[3], n2[3],
 *                                 u0,u1,u2, v0,v1,v2}
 *                     Objects are in BVH order (bvh.h:98-105).
 *  instance_*         ints[n][4] = {first object, object count, BLAS root,
 *                     material override or -1}; matrices[n][32] = object to
 *                     world 4x4 row-major, then its inverse (instance.h:114-115).
 *                     Instances are in BVH order (bvh.h:74-81).
 *  node_*             bounds[n][6] = {min xyz, max xyz}; ints[n][2] = {index,
 *                     count}: count > 0 is a leaf over [index, index+count);
 *                     count == 0 is an inner node with left child at i+1 and
 *                     right child at index (bvh.h:14-28).  Node 0 is the root of
 *                     the instance level; each instance's BLAS root is in
 *                     instance ints[.][2].
 *  light_*            num_lights entries; the last one is the environment with
 *                     instance and object -1 (parser.h:335-378).
 */
typedef struct lambda_scene_view {
    int32_t num_spectra;
    const float *spectra;
    int32_t num_complex_spectra;
    const double *complex_spectra;

    int32_t num_scalar_textures;
    const int32_t *scalar_texture_type;
    const float *scalar_texture_params;

    int32_t num_spectrum_textures;
    const int32_t *spectrum_texture_type;
    const int32_t *spectrum_texture_index;
    const float *spectrum_texture_scale;

    int32_t num_materials;
    const int32_t *material_type;
    const int32_t *material_params;
    int32_t num_material_properties;
    const int32_t *material_properties;

    int32_t num_objects;
    const int32_t *object_type;
    const int32_t *object_material;
    const float *object_geometry;
    const float *object_shading;

    int32_t num_instances;
    const int32_t *instance_ints;
    const float *instance_matrices;

    int32_t num_nodes;
    const float *node_bounds;
    const int32_t *node_ints;

    int32_t num_lights;
    const int32_t *light_instances;
    const int32_t *light_objects;
    const float *light_powers;
    float total_light_power;

    int32_t background_texture;
    float camera[12]; /* position, corner, horizontal, vertical */
} lambda_scene_view;

/* Random-number stream layouts. */
enum lambda_rng_mode {
    /* Each pixel owns PCG32 stream (seed, py*W+px) and its samples run one
     * after another, consuming the stream draw for draw as the reference does
     * (render/renderer.h:49,69-91).  Images match the reference CPU render up
     * to libm rounding. */
    LAMBDA_RNG_REFERENCE = 0,
    /* Each (pixel, sample) pair owns its own PCG32 stream, so every sample of
     * every pixel can be in flight at once.  Same estimator, different
     * random numbers. */
    LAMBDA_RNG_THROUGHPUT = 1
};

/* Render settings (Render command, parser.h:75-100; seed, main.cpp:122). */
typedef struct lambda_render_params {
    int32_t width;
    int32_t height;
    int32_t depth;         /* last bounce index; depth+1 segments */
    int32_t samples;       /* perfect square */
    int32_t sqrt_samples;
    float lambda_min;      /* nm, >= 360 */
    float lambda_max;      /* nm, <= 830 */
    uint64_t seed;         /* the reference always passes 42 */
    int32_t rng_mode;      /* enum lambda_rng_mode */
    int32_t path_slots;    /* paths in flight (throughput mode); 0 = automatic */
    int32_t tail_paths;    /* once fewer paths than this are live, one persistent
                              launch runs each remaining path to its end instead
                              of one launch per bounce; 0 = automatic, -1 = never */
    int32_t cost_order;    /* reference streams, persistent schedule: render
                              the first n samples of every pixel, timing each
                              pixel, then the rest with the slowest 8x8 tiles
                              first (the image does not change): 0 =
                              automatic (n = 4 when there are fewer slots
                              than pixels), -1 = off, n >= 1 = on with n
                              samples first (at most samples - 1) */
    int32_t graph_passes;  /* > 0: one launch per bounce throughout, the bounce
                              loop captured in a hipGraph of this many passes
                              (rounded up to even) and replayed until done;
                              0 = no graph */
    int32_t tile;          /* side of the square pixel tiles of the visiting
                              order (4, 8 or 16; 0 = 8) */
    int32_t occupancy;     /* persistent kernel variant: waves per SIMD its
                              registers are capped for (2, 3 or 4; 0 = 4) */
    int32_t vote;          /* persistent schedule: 0 = each lane traverses and
                              shades in turn; 1 = the wave alternates traversal
                              steps and shading, shading when at least half
                              of its lanes wait (3 or 4 waves per SIMD; same
                              image, faster with throughput streams) */
} lambda_render_params;

/*
 * One-shot render: uploads the scene, renders, writes out_rgb[W*H*3] (linear
 * sRGB, mean over samples, row-major from the top image row) and frees every
 * device allocation before returning.  progress, when non-NULL, receives the
 * number of finished pixels (monotone, readable from another thread while the
 * call runs; W*H on return).
 * Replaces src/main.cpp:105-156 (renderKernel launch, poll loop, copy-back).
 */
int lambda_hip_render(const lambda_scene_view *scene, const lambda_render_params *params,
                      float *out_rgb, volatile int32_t *progress, char *err, size_t err_len);

/* A scene resident in device memory (the uploads of src/main.cpp:84-103). */
typedef struct lambda_hip_scene lambda_hip_scene;

/* Validates the scene view and copies it into device memory. */
int lambda_hip_scene_create(const lambda_scene_view *scene, lambda_hip_scene **out,
                            char *err, size_t err_len);

/*
 * The same, with the BVHs built on the GPU (no reference counterpart: the
 * reference builds a SAH BVH on the host, scene/bvh.h:61-246; this is for
 * scenes whose geometry changes between frames).  The node arrays of the
 * view are not read: each object group (the object range of the instances)
 * and the instance level get a linear BVH (Morton codes, radix sort, binary
 * radix tree, one primitive per leaf).  Objects and instances are renumbered
 * in the new trees' leaf order: object_order[k] (O entries, may be NULL) and
 * instance_order[k] (I entries) give the view's index of the new object and
 * instance k; hit IDs are in the new numbering, hit distances are those of
 * the host-built BVH.  Instance object ranges must tile the object array.
 * *build_ms (may be NULL): the BVH build's wall time.
 */
int lambda_hip_scene_create_lbvh(const lambda_scene_view *scene, lambda_hip_scene **out, int32_t *object_order,
                                 int32_t *instance_order, double *build_ms, char *err, size_t err_len);

/* Releases every device allocation of the scene. */
void lambda_hip_scene_destroy(lambda_hip_scene *scene);

/*
 * Renders an uploaded scene; the same semantics as lambda_hip_render without
 * the upload.  Replaces src/main.cpp:113-156.
 */
int lambda_hip_scene_render(lambda_hip_scene *scene, const lambda_render_params *params,
                            float *out_rgb, volatile int32_t *progress, char *err, size_t err_len);

/* Measurements of one render (no reference counterpart: the reference has
 * no timers).  Kernel times come from HIP events recorded around every launch
 * on the render's stream. */
typedef struct lambda_render_stats {
    int32_t iterations;      /* bounce-loop passes */
    int32_t path_slots;      /* paths in flight at the start */
    int64_t closest_rays;    /* closest-hit rays traced */
    int64_t shadow_rays;     /* next-event occlusion rays traced */
    double kernel_ms[4];     /* total device time: trace (closest-hit and occlusion
                                rays), shade, finalize, persistent tail */
    int64_t launches[4];
    double render_ms;        /* wall time of the whole call */
} lambda_render_stats;

/* lambda_hip_scene_render plus per-kernel timing into *stats. */
int lambda_hip_scene_render_stats(lambda_hip_scene *scene, const lambda_render_params *params,
                                  float *out_rgb, lambda_render_stats *stats, char *err, size_t err_len);

/* What the library made of an uploaded scene (no reference counterpart). */
typedef struct lambda_scene_info {
    int32_t stack_need;     /* deepest traversal stack of the scene (entries) */
    int32_t stack_lds;      /* entries per lane kept in LDS; deeper entries go to
                               a global spill area */
    int32_t wide_nodes;     /* 4-wide nodes built from the scene's BVH nodes */
    int32_t num_nodes;      /* the scene's BVH nodes (node_bounds rows) */
    int32_t num_objects;
    int32_t num_instances;
    float bounds[6];        /* instance-level root box: min xyz, max xyz */
} lambda_scene_info;

int lambda_hip_scene_info(const lambda_hip_scene *scene, lambda_scene_info *info, char *err, size_t err_len);

/* Closest-hit record for one ray (math/intersection.h:8-17). */
typedef struct lambda_hit {
    int32_t hit;        /* 1 if the ray hit an object */
    int32_t instance;   /* index into the (BVH-ordered) instance array */
    int32_t object;     /* index into the (BVH-ordered) object array */
    float t;
    float point[3];     /* world space */
    float normal[3];    /* world space, facing the ray */
    float u, v;
    int32_t front_facing;
} lambda_hit;

/*
 * Closest-hit queries for rays[n][6] = {origin xyz, direction xyz} in world
 * space: Scene::hit in closest-hit mode (scene/scene.h:51-124 with
 * occludedObject = -2).
 */
int lambda_hip_closest_hit(lambda_hip_scene *scene, const float *rays, int32_t n,
                           lambda_hit *hits, char *err, size_t err_len);

/*
 * Closest-hit queries that allocate nothing (no reference counterpart: for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops).  rays[n][6] and hits[n] are device memory
 * of the current device, which must hold the scene; the kernel is queued on
 * `stream` (a hipStream_t, NULL for the null stream) and the call returns
 * without waiting.  Results equal lambda_hip_closest_hit's.  Traversal
 * stacks deeper than the LDS part use a fixed 16 MiB scratch area per device
 * (allocated with the library, held until the process ends): scenes whose
 * traversal needs more than 48 stack entries (lambda_scene_info.stack_need)
 * are rejected with LAMBDA_EINVAL.  The scratch needs no clearing; calls
 * that overlap on one device share it and race.
 */
int lambda_hip_closest_hit_device(lambda_hip_scene *scene, const float *rays, int32_t n, lambda_hit *hits,
                                  void *stream, char *err, size_t err_len);

/*
 * Shadow queries: Scene::hit in occlusion mode (scene/scene.h:61-70,99,106)
 * for rays[n][6] toward light (light_instance[i], light_object[i]); -1/-1 is
 * the environment.  occluded[i] = 1 if Scene::hit returned true.
 */
int lambda_hip_occluded(lambda_hip_scene *scene, const float *rays, const int32_t *light_instance,
                        const int32_t *light_object, int32_t n, int32_t *occluded,
                        char *err, size_t err_len);

/*
 * Thin-film reflectance (scene/material.h:263-309) of material `material` for
 * n queries[i][4] = {cos_theta, lambda_nm, front_facing (0/1), unused} on a
 * surface with normal (0,0,1) hit by direction (sin_theta, 0, -cos_theta); the
 * layer thicknesses come from the material's scalar textures evaluated at
 * points[i][3].  Computed by the same FP32 device code the integrator uses.
 * out[i] = reflectance.
 */
int lambda_hip_thinfilm_reflectance(lambda_hip_scene *scene, int32_t material, const float *queries,
                                    const float *points, int32_t n, float *out,
                                    char *err, size_t err_len);

/*
 * Path-guiding network (no reference counterpart: the reference samples
 * directions from the BSDF only).  A mixture of 8 von Mises-Fisher lobes
 * over the next direction, whose parameters come from an MLP of (position,
 * direction, wavelength): each of 6 features in [0, 1] is encoded
 * as sin/cos of 2^b pi x, b = 0..7 (96 inputs), 4 hidden layers of 64 ReLU
 * units, 40 outputs = 8 x {mean xyz (normalised), kappa (exp, clamped to [1e-3, 2000]),
 * weight (softmax logit)}.  BF16 weights and activations on the matrix
 * cores with FP32 accumulation; FP32 master weights; Adam.
 *
 * Parameters: for each layer l = 0..4, W_l[K_l][64] row-major then
 * b_l[64]; K_0 = 96, K_1..4 = 64; output columns 40..63 are padding.
 * The renderer guides Lambertian vertices only, whose integrand does not
 * depend on the outgoing direction: the direction feature it passes is the
 * surface normal on the side of the incoming ray.
 * Training sample: {features[6], direction[3], weight}; the loss of a
 * sample is -weight * log q(direction).
 */
#define LAMBDA_GUIDE_PARAMS 22848
#define LAMBDA_GUIDE_FEATURES 6
#define LAMBDA_GUIDE_OUTPUTS 40
#define LAMBDA_GUIDE_SAMPLE 10
/* bfloat16 weights per copy: 96*64 + 4*64*64 */
#define LAMBDA_GUIDE_WEIGHTS_BF16 22528

typedef struct lambda_guide lambda_guide;

/* New network: He-uniform weights from PCG32 stream (seed, 1), zero biases. */
int lambda_hip_guide_create(uint64_t seed, lambda_guide **out, char *err, size_t err_len);
void lambda_hip_guide_destroy(lambda_guide *guide);
int lambda_hip_guide_get_params(lambda_guide *guide, float *params, char *err, size_t err_len);
/* Replaces the parameters and restarts Adam. */
int lambda_hip_guide_set_params(lambda_guide *guide, const float *params, char *err, size_t err_len);
/* out[i][0..39] = raw network outputs for features[i][0..5]. */
int lambda_hip_guide_forward(lambda_guide *guide, const float *features, int32_t n, float *out,
                             char *err, size_t err_len);
/* Sum over the n samples of the gradient of the loss (grads[PARAMS]) and of
 * the loss itself; the parameters do not change. */
int lambda_hip_guide_gradient(lambda_guide *guide, const float *samples, int32_t n, float *grads, double *loss,
                              char *err, size_t err_len);
/* Adam steps (rate lr) over consecutive batches of `batch` samples, each
 * sample's weight capped at 20 x the mean weight of the n samples (heavy-
 * tailed radiance weights then cannot steer a whole step); *loss receives
 * the summed loss of the samples seen. */
int lambda_hip_guide_train(lambda_guide *guide, const float *samples, int32_t n, int32_t batch, float lr,
                           double *loss, char *err, size_t err_len);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected by the
 * step count since the last set_params) with the gradient grads[PARAMS] x
 * scale and rate lr: the step a training call takes after each batch.  It
 * updates the FP32 master weights and moments and rewrites both bfloat16
 * weight copies. */
int lambda_hip_guide_adam_step(lambda_guide *guide, const float *grads, float scale, float lr, char *err,
                               size_t err_len);
/* Adam state: moments m and v [PARAMS], the step count, and the two bfloat16
 * weight copies as raw 16-bit words, [LAMBDA_GUIDE_WEIGHTS_BF16] each:
 * w_nk holds W_l^T (n-major, [64][K_l]) and w_kn holds W_l ([K_l][64]),
 * layer after layer.  Any pointer may be NULL. */
int lambda_hip_guide_get_state(lambda_guide *guide, float *m, float *v, int32_t *step, uint16_t *w_nk,
                               uint16_t *w_kn, char *err, size_t err_len);

/* Times the network's kernels on `rows` device-resident samples: mean
 * milliseconds per forward pass and per gradient pass over `iters` runs
 * (HIP events on the null stream).  The parameters do not change. */
int lambda_hip_guide_benchmark(lambda_guide *guide, int32_t rows, int32_t iters, double *forward_ms,
                               double *gradient_ms, char *err, size_t err_len);

/*
 * Batched serving (no reference counterpart: the reference renders one
 * scene per process).  Many render jobs share one persistent launch: each
 * job is a scene and render settings with reference streams (its image is
 * bit-identical to lambda_hip_scene_render of the same settings), admitted
 * at its arrival time (milliseconds after the first batch starts, on the
 * device clock).  Jobs are admitted in arrival order while their buffers
 * (36 bytes per pixel) fit hbm_budget (0 = no limit); the rest form later
 * launches.  Latency = the time the job's last pixel finished - arrival.
 */
typedef struct lambda_job {
    int32_t scene;                  /* index into the scenes array */
    lambda_render_params params;    /* rng_mode is ignored (reference streams) */
    double arrival_ms;
    float *out_rgb;                 /* W*H*3 linear sRGB, written on return */
} lambda_job;

typedef struct lambda_job_result {
    double latency_ms;
    double finish_ms;               /* after the first batch started */
    int32_t batch;                  /* launch that served the job */
} lambda_job_result;

typedef struct lambda_serve_stats {
    int32_t batches;
    int64_t paths;                  /* sum over jobs of W*H*spp */
    int64_t closest_rays, shadow_rays;
    double wall_ms;                 /* first batch start to the last result */
} lambda_serve_stats;

int lambda_hip_serve(lambda_hip_scene *const *scenes, int32_t num_scenes, const lambda_job *jobs, int32_t num_jobs,
                     int64_t hbm_budget, lambda_job_result *results, lambda_serve_stats *stats, char *err,
                     size_t err_len);

/*
 * Live serving (no reference counterpart): jobs are submitted while earlier
 * ones render, as a serving front end receives them.  A scheduler thread
 * admits queued jobs in submission order while the buffers of the jobs in
 * flight (36 bytes per pixel) fit hbm_budget (0 = no limit; a job larger
 * than the budget runs alone), and renders each admitted group as one
 * persistent launch on one of max_launches streams (0 = 2), so that a new
 * group starts on the compute units a draining launch frees.  Within a
 * launch the jobs with the longest per-pixel sample chains (samples x
 * (depth + 1)) start first.  Every job's image is bit-identical to
 * lambda_hip_scene_render of the same settings with reference streams.
 * Latency = the device time the job's last pixel finished - submission.
 */
typedef struct lambda_server lambda_server;

enum lambda_job_state { LAMBDA_JOB_QUEUED = 0, LAMBDA_JOB_RUNNING = 1, LAMBDA_JOB_DONE = 2, LAMBDA_JOB_FAILED = 3 };

/* The scenes stay owned by the caller and must outlive the server. */
int lambda_hip_server_create(lambda_hip_scene *const *scenes, int32_t num_scenes, int64_t hbm_budget,
                             int32_t max_launches, lambda_server **out, char *err, size_t err_len);
/* Queues a job (arrival_ms is ignored: it arrives now) and returns its
 * ticket; job->out_rgb must stay valid until the job is done. */
int lambda_hip_server_submit(lambda_server *server, const lambda_job *job, int64_t *ticket, char *err,
                             size_t err_len);
/* Non-blocking: *state = LAMBDA_JOB_*; with LAMBDA_JOB_DONE the image is in
 * out_rgb and *result (may be NULL) holds latency, finish time (ms after
 * the server's creation) and launch index.  A failed job returns its error. */
int lambda_hip_server_poll(lambda_server *server, int64_t ticket, int32_t *state, lambda_job_result *result,
                           char *err, size_t err_len);
/* Waits until every submitted job has finished; *stats (may be NULL): launches,
 * paths, rays, and the time from the first submission to the last finish. */
int lambda_hip_server_drain(lambda_server *server, lambda_serve_stats *stats, char *err, size_t err_len);
/* Finishes the queued jobs, then releases the server. */
void lambda_hip_server_destroy(lambda_server *server);

/*
 * Progressive rendering (no reference counterpart): a session accumulates
 * samples over steps into one FP64 framebuffer, with throughput streams
 * (fresh PCG32 streams per step).  With a guide, Lambertian vertices sample
 * the one-sample mixture 0.5 cos/pi + 0.5 q(guide), whose density is the
 * scatter pdf of both MIS weights, so every step is unbiased; a step can
 * then train the guide (Adam, batches of 2^14) on up to 2^18 samples of the
 * radiance its own paths gathered at Lambertian vertices.  Without a guide
 * the session is the plain tracer.
 */
typedef struct lambda_hip_progressive lambda_hip_progressive;

typedef struct lambda_progressive_stats {
    int64_t samples;        /* samples per pixel accumulated so far */
    int64_t train_samples;  /* training samples collected by the last step */
    double render_ms;       /* device time of the last step's rendering */
    double train_ms;        /* device time of its training (0 if none) */
    double loss;            /* mean training loss of that training */
    int64_t closest_rays;   /* closest-hit rays traced by the last step */
    int64_t shadow_rays;    /* its next-event occlusion rays */
} lambda_progressive_stats;

/* params: size, depth, wavelength range, seed, path_slots and tail_paths
 * are used; rng_mode and samples are ignored.  A guided session evaluates
 * the network inside its persistent kernel (each wave runs it for its own
 * 64 paths on the matrix cores); tail_paths = -1 selects one launch per
 * bounce with the network as a kernel of its own instead.  guide may be
 * NULL; the session borrows it (keep it alive until the session is
 * destroyed). */
int lambda_hip_progressive_create(lambda_hip_scene *scene, const lambda_render_params *params, lambda_guide *guide,
                                  lambda_hip_progressive **out, char *err, size_t err_len);
/* Step flags. */
#define LAMBDA_STEP_TRAIN 1    /* train the guide on this step's paths afterwards */
#define LAMBDA_STEP_DISCARD 2  /* do not add this step's samples to the image */

/* Renders `samples` (a perfect square) more samples per pixel.  Training
 * samples are a uniform subsample (sized from the previous step) of the
 * guided vertices whose continuation was traced. */
int lambda_hip_progressive_step(lambda_hip_progressive *session, int32_t samples, int32_t flags,
                                lambda_progressive_stats *stats, char *err, size_t err_len);
/* out_rgb[W*H*3]: linear sRGB mean over all samples so far. */
int lambda_hip_progressive_image(lambda_hip_progressive *session, float *out_rgb, char *err, size_t err_len);
/* Training per step: `epochs` passes of Adam (rate lr) over the step's
 * samples in batches of `batch` (defaults 1, 16384, 1e-3).  Vertices with
 * bounce index below guided_bounces are guided (0 = the default of 8; a
 * value above the path depth guides every vertex); with fewer guided
 * bounces than the depth, each batch of paths runs its guided bounces as
 * passes and the rest in one persistent launch. */
int lambda_hip_progressive_set_training(lambda_hip_progressive *session, int32_t epochs, int32_t batch, float lr,
                                        int32_t guided_bounces, char *err, size_t err_len);
/* Tabulated guide (guided persistent sessions, before the first step): after
 * each training, the network is evaluated once at the centres of a grid of
 * cubic cells over the scene's bounding box (`cells` along its longest
 * side) for each of the six axis directions as the surface normal, at the
 * middle of the session's wavelength range; a guided vertex then samples the
 * mixture of its cell and of the axis nearest its normal instead of running
 * the network, in a kernel with the plain kernel's occupancy.  Steps stay
 * unbiased (sampling and both MIS densities use the same entry).  cells = 0
 * returns to the network inside the kernel; at most 1024 cells and 2^24
 * entries (cells x 6), 160 bytes of device memory each. */
int lambda_hip_progressive_set_guide_table(lambda_hip_progressive *session, int32_t cells, char *err, size_t err_len);
/* Probability of sampling the guide at a guided vertex (default 0.5, the
 * one-sample mixture 0.5 cos/pi + 0.5 q); frac in (0, 0.95], the cosine
 * lobe takes the rest and both MIS densities use the same mixture. */
int lambda_hip_progressive_set_guide_mix(lambda_hip_progressive *session, float frac, char *err, size_t err_len);
/* Bounce index from which Russian roulette applies: plain sessions default
 * to 5, the reference's (render/renderer.h:185-191); guided sessions default
 * to off (depth + 1: every path runs to the depth or to a miss or an
 * emitter), since roulette on the throughput ends most of the paths the
 * guide sends toward the light (their throughput is small where the guide's
 * density is high).  Any value keeps the estimator unbiased.
 * start_bounce >= 1. */
int lambda_hip_progressive_set_roulette(lambda_hip_progressive *session, int32_t start_bounce, char *err,
                                        size_t err_len);
void lambda_hip_progressive_destroy(lambda_hip_progressive *session);

/* Device name and compute-unit count of the current HIP device. */
int lambda_hip_device_info(char *name, size_t name_len, int32_t *compute_units,
                           char *err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* LAMBDA_HIP_H */
