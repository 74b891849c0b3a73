/*
 * lambda_host.h - C ABI of the host front end: the .lrd/.obj scene parser,
 * the binned-SAH BVH builder, the light list and the PPM writer.
 *
 * The front end is a C++ restatement of Lambda's host code (core/parser.h,
 * scene/bvh.h, core/writer.h).  Its output arrays are bit-identical to the
 * reference's: same object, instance and node order, same float values.  It
 * exists so that lambda_hip_render() can be driven without the reference
 * sources; a maintainer integrating the HIP layer into Lambda itself would
 * flatten Lambda's own vectors instead (see the integration notes).
 */
#ifndef LAMBDA_HOST_H
#define LAMBDA_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "lambda_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lambda_scene lambda_scene;

/*
 * Parses a .lrd file (and the .obj files it names), builds the two-level BVH
 * and the light list.  Replaces Parser::parseLRD (core/parser.h:29-379).
 * On a parse error returns LAMBDA_EPARSE with the reference's message
 * ("<file>:<line>: <message>").
 */
int lambda_scene_load(const char *lrd_path, lambda_scene **out, char *err, size_t err_len);

/* Same as lambda_scene_load on the text of a .lrd file; `name` is used in
 * error messages and relative .obj paths resolve against the current
 * directory, as in the reference. */
int lambda_scene_parse_string(const char *text, const char *name, lambda_scene **out,
                              char *err, size_t err_len);

/* Fills `view` with pointers into the scene; valid until lambda_scene_free. */
int lambda_scene_get_view(const lambda_scene *scene, lambda_scene_view *view);

/* Fills width..lambda_max from the Render command; seed 42, reference RNG. */
int lambda_scene_get_params(const lambda_scene *scene, lambda_render_params *params);

void lambda_scene_free(lambda_scene *scene);

/*
 * Writes a plain PPM (P3): clamp to [0,1], sRGB transfer curve, dithered
 * 8-bit quantization with PCG32 stream (0,0).  Replaces Writer::write
 * (core/writer.h:17-41).
 */
int lambda_write_ppm(const char *path, int32_t width, int32_t height, const float *rgb,
                     char *err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* LAMBDA_HOST_H */
