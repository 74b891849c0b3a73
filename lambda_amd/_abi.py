"""ctypes mirror of include/lambda_hip.h and include/lambda_host.h."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HOST_LIB = os.path.join(LIB_DIR, "liblambda_host.so")
HIP_LIB = os.path.join(LIB_DIR, "liblambda_hip.so")

CIE_BINS = 471

LAMBDA_RNG_REFERENCE = 0
LAMBDA_RNG_THROUGHPUT = 1

_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)
_f64p = C.POINTER(C.c_double)


class SceneView(C.Structure):
    _fields_ = [
        ("num_spectra", C.c_int32), ("spectra", _f32p),
        ("num_complex_spectra", C.c_int32), ("complex_spectra", _f64p),
        ("num_scalar_textures", C.c_int32), ("scalar_texture_type", _i32p), ("scalar_texture_params", _f32p),
        ("num_spectrum_textures", C.c_int32), ("spectrum_texture_type", _i32p),
        ("spectrum_texture_index", _i32p), ("spectrum_texture_scale", _f32p),
        ("num_materials", C.c_int32), ("material_type", _i32p), ("material_params", _i32p),
        ("num_material_properties", C.c_int32), ("material_properties", _i32p),
        ("num_objects", C.c_int32), ("object_type", _i32p), ("object_material", _i32p),
        ("object_geometry", _f32p), ("object_shading", _f32p),
        ("num_instances", C.c_int32), ("instance_ints", _i32p), ("instance_matrices", _f32p),
        ("num_nodes", C.c_int32), ("node_bounds", _f32p), ("node_ints", _i32p),
        ("num_lights", C.c_int32), ("light_instances", _i32p), ("light_objects", _i32p),
        ("light_powers", _f32p), ("total_light_power", C.c_float),
        ("background_texture", C.c_int32),
        ("camera", C.c_float * 12),
    ]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("depth", C.c_int32),
        ("samples", C.c_int32), ("sqrt_samples", C.c_int32),
        ("lambda_min", C.c_float), ("lambda_max", C.c_float),
        ("seed", C.c_uint64), ("rng_mode", C.c_int32), ("path_slots", C.c_int32),
        ("tail_paths", C.c_int32), ("cost_order", C.c_int32), ("graph_passes", C.c_int32),
        ("tile", C.c_int32), ("occupancy", C.c_int32),
        ("vote", C.c_int32),
    ]


class Hit(C.Structure):
    _fields_ = [
        ("hit", C.c_int32), ("instance", C.c_int32), ("object", C.c_int32), ("t", C.c_float),
        ("point", C.c_float * 3), ("normal", C.c_float * 3), ("u", C.c_float), ("v", C.c_float),
        ("front_facing", C.c_int32),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("path_slots", C.c_int32), ("closest_rays", C.c_int64),
        ("shadow_rays", C.c_int64), ("kernel_ms", C.c_double * 4), ("launches", C.c_int64 * 4),
        ("render_ms", C.c_double),
    ]


class SceneInfo(C.Structure):
    _fields_ = [("stack_need", C.c_int32), ("stack_lds", C.c_int32), ("wide_nodes", C.c_int32),
                ("num_nodes", C.c_int32), ("num_objects", C.c_int32), ("num_instances", C.c_int32),
                ("bounds", C.c_float * 6)]


GUIDE_WEIGHTS_BF16 = 22528


class Job(C.Structure):
    _fields_ = [("scene", C.c_int32), ("params", RenderParams), ("arrival_ms", C.c_double),
                ("out_rgb", C.POINTER(C.c_float))]


class JobResult(C.Structure):
    _fields_ = [("latency_ms", C.c_double), ("finish_ms", C.c_double), ("batch", C.c_int32)]


class ServeStats(C.Structure):
    _fields_ = [("batches", C.c_int32), ("paths", C.c_int64), ("closest_rays", C.c_int64),
                ("shadow_rays", C.c_int64), ("wall_ms", C.c_double)]


class ProgressiveStats(C.Structure):
    _fields_ = [
        ("samples", C.c_int64), ("train_samples", C.c_int64), ("render_ms", C.c_double),
        ("train_ms", C.c_double), ("loss", C.c_double), ("closest_rays", C.c_int64), ("shadow_rays", C.c_int64),
    ]


KERNELS = ("trace", "shade", "finalize", "persist")


HIT_DTYPE = np.dtype([
    ("hit", np.int32), ("instance", np.int32), ("object", np.int32), ("t", np.float32),
    ("point", np.float32, 3), ("normal", np.float32, 3), ("u", np.float32), ("v", np.float32),
    ("front_facing", np.int32),
])
assert HIT_DTYPE.itemsize == C.sizeof(Hit)


def _load(path: str, what: str):
    if not os.path.exists(path):
        raise RuntimeError(
            "%s is not built (%s missing); run __graft_entry__.build() first" % (what, path))
    return C.CDLL(path)


_host = None


def host():
    """The host front end library (parser, BVH builder, PPM writer)."""
    global _host
    if _host is None:
        lib = _load(HOST_LIB, "the host front end")
        vp = C.c_void_p
        lib.lambda_scene_load.argtypes = [C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_size_t]
        lib.lambda_scene_parse_string.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(vp), C.c_char_p, C.c_size_t]
        lib.lambda_scene_get_view.argtypes = [vp, C.POINTER(SceneView)]
        lib.lambda_scene_get_params.argtypes = [vp, C.POINTER(RenderParams)]
        lib.lambda_scene_free.argtypes = [vp]
        lib.lambda_scene_free.restype = None
        lib.lambda_write_ppm.argtypes = [C.c_char_p, C.c_int32, C.c_int32, _f32p, C.c_char_p, C.c_size_t]
        _host = lib
    return _host


_hip = None


def hip():
    """The HIP integrator library; raises if it is not built."""
    global _hip
    if _hip is None:
        lib = _load(HIP_LIB, "the HIP integrator")
        vp = C.c_void_p
        sv = C.POINTER(SceneView)
        rp = C.POINTER(RenderParams)
        lib.lambda_hip_render.argtypes = [sv, rp, _f32p, _i32p, C.c_char_p, C.c_size_t]
        lib.lambda_hip_scene_create.argtypes = [sv, C.POINTER(vp), C.c_char_p, C.c_size_t]
        lib.lambda_hip_scene_create_lbvh.argtypes = [sv, C.POINTER(vp), _i32p, _i32p, C.POINTER(C.c_double),
                                                     C.c_char_p, C.c_size_t]
        lib.lambda_hip_scene_destroy.argtypes = [vp]
        lib.lambda_hip_scene_destroy.restype = None
        lib.lambda_hip_scene_render.argtypes = [vp, rp, _f32p, _i32p, C.c_char_p, C.c_size_t]
        lib.lambda_hip_scene_render_stats.argtypes = [vp, rp, _f32p, C.POINTER(RenderStats), C.c_char_p, C.c_size_t]
        lib.lambda_hip_closest_hit.argtypes = [vp, _f32p, C.c_int32, C.POINTER(Hit), C.c_char_p, C.c_size_t]
        lib.lambda_hip_closest_hit_device.argtypes = [vp, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                                      C.c_char_p, C.c_size_t]
        lib.lambda_hip_occluded.argtypes = [vp, _f32p, _i32p, _i32p, C.c_int32, _i32p, C.c_char_p, C.c_size_t]
        lib.lambda_hip_thinfilm_reflectance.argtypes = [vp, C.c_int32, _f32p, _f32p, C.c_int32, _f32p,
                                                        C.c_char_p, C.c_size_t]
        lib.lambda_hip_device_info.argtypes = [C.c_char_p, C.c_size_t, _i32p, C.c_char_p, C.c_size_t]
        f64p = C.POINTER(C.c_double)
        lib.lambda_hip_guide_create.argtypes = [C.c_uint64, C.POINTER(vp), C.c_char_p, C.c_size_t]
        lib.lambda_hip_guide_destroy.argtypes = [vp]
        lib.lambda_hip_guide_destroy.restype = None
        lib.lambda_hip_guide_get_params.argtypes = [vp, _f32p, C.c_char_p, C.c_size_t]
        lib.lambda_hip_guide_set_params.argtypes = [vp, _f32p, C.c_char_p, C.c_size_t]
        lib.lambda_hip_guide_forward.argtypes = [vp, _f32p, C.c_int32, _f32p, C.c_char_p, C.c_size_t]
        lib.lambda_hip_guide_gradient.argtypes = [vp, _f32p, C.c_int32, _f32p, f64p, C.c_char_p, C.c_size_t]
        lib.lambda_hip_guide_train.argtypes = [vp, _f32p, C.c_int32, C.c_int32, C.c_float, f64p, C.c_char_p,
                                               C.c_size_t]
        lib.lambda_hip_guide_adam_step.argtypes = [vp, _f32p, C.c_float, C.c_float, C.c_char_p, C.c_size_t]
        u16p = C.POINTER(C.c_uint16)
        lib.lambda_hip_guide_get_state.argtypes = [vp, _f32p, _f32p, _i32p, u16p, u16p, C.c_char_p, C.c_size_t]
        lib.lambda_hip_scene_info.argtypes = [vp, C.POINTER(SceneInfo), C.c_char_p, C.c_size_t]
        lib.lambda_hip_guide_benchmark.argtypes = [vp, C.c_int32, C.c_int32, f64p, f64p, C.c_char_p, C.c_size_t]
        lib.lambda_hip_serve.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(Job), C.c_int32, C.c_int64,
                                         C.POINTER(JobResult), C.POINTER(ServeStats), C.c_char_p, C.c_size_t]
        lib.lambda_hip_server_create.argtypes = [C.POINTER(vp), C.c_int32, C.c_int64, C.c_int32, C.POINTER(vp),
                                                 C.c_char_p, C.c_size_t]
        lib.lambda_hip_server_submit.argtypes = [vp, C.POINTER(Job), C.POINTER(C.c_int64), C.c_char_p, C.c_size_t]
        lib.lambda_hip_server_poll.argtypes = [vp, C.c_int64, _i32p, C.POINTER(JobResult), C.c_char_p, C.c_size_t]
        lib.lambda_hip_server_drain.argtypes = [vp, C.POINTER(ServeStats), C.c_char_p, C.c_size_t]
        lib.lambda_hip_server_destroy.argtypes = [vp]
        lib.lambda_hip_server_destroy.restype = None
        lib.lambda_hip_progressive_create.argtypes = [vp, rp, vp, C.POINTER(vp), C.c_char_p, C.c_size_t]
        lib.lambda_hip_progressive_step.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(ProgressiveStats),
                                                    C.c_char_p, C.c_size_t]
        lib.lambda_hip_progressive_image.argtypes = [vp, _f32p, C.c_char_p, C.c_size_t]
        lib.lambda_hip_progressive_set_training.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, C.c_int32,
                                                            C.c_char_p, C.c_size_t]
        lib.lambda_hip_progressive_set_guide_table.argtypes = [vp, C.c_int32, C.c_char_p, C.c_size_t]
        lib.lambda_hip_progressive_set_guide_mix.argtypes = [vp, C.c_float, C.c_char_p, C.c_size_t]
        lib.lambda_hip_progressive_set_roulette.argtypes = [vp, C.c_int32, C.c_char_p, C.c_size_t]
        lib.lambda_hip_progressive_destroy.argtypes = [vp]
        lib.lambda_hip_progressive_destroy.restype = None
        _hip = lib
    return _hip


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
