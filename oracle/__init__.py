"""Test infrastructure only: the CPU oracle of the spectral path tracer.

``oracle/oracle.cpp`` restates Lambda's CPU integrator (render/renderer.h and
everything it calls) over the same ``lambda_scene_view`` the HIP layer
consumes, and reproduces the reference's float image bit for bit.  This module
loads ``oracle/lib/liblambda_oracle.so`` with ctypes.  Only tests, the smoke
check and the benchmark's CPU-baseline leg import it; the product never does.

Pins (each checked by tests/):
  * PCG32 against pcg-c's published demo output (seed 42, stream 54).
  * Everything else against golden vectors emitted by the reference itself
    (oracle/_ref/ref_harness, built from the reference sources): parsed scene
    arrays, rendered images, hit records, thin-film reflectances and
    randomUnitVector outputs (tests/golden/).
  * Thin film additionally against closed forms (Fresnel, Airy, quarter-wave).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "liblambda_oracle.so")

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built (%s); run `make -C oracle`" % LIB)
        import lambda_amd._abi as abi
        lib = C.CDLL(LIB)
        sv = C.POINTER(abi.SceneView)
        rp = C.POINTER(abi.RenderParams)
        f32 = C.POINTER(C.c_float)
        i32 = C.POINTER(C.c_int32)
        lib.oracle_render.argtypes = [sv, rp, f32, C.c_int32, C.c_int32, C.c_int32]
        lib.oracle_closest_hit.argtypes = [sv, f32, C.c_int32, C.POINTER(abi.Hit)]
        lib.oracle_occluded.argtypes = [sv, f32, i32, i32, C.c_int32, i32]
        lib.oracle_thinfilm.argtypes = [sv, C.c_int32, f32, C.c_int32, f32, C.POINTER(C.c_double), i32]
        lib.oracle_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_int32, C.POINTER(C.c_uint32)]
        lib.oracle_pcg32.restype = None
        lib.oracle_path_stats.argtypes = [sv, rp, C.c_int64, C.c_int64, C.c_int64, C.c_int32, C.POINTER(C.c_int64)]
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def render(scene, params=None, threads: int = 0, rows=None) -> np.ndarray:
    """Reference-exact CPU render; returns float32 [H, W, 3]."""
    lib = _load()
    p = params if params is not None else scene.params
    out = np.zeros((p.height, p.width, 3), np.float32)
    r0, r1 = (0, p.height) if rows is None else rows
    rc = lib.oracle_render(C.byref(scene.view), C.byref(p), _p(out, C.c_float), r0, r1, threads)
    if rc:
        raise RuntimeError("oracle_render rejected the scene")
    return out


def closest_hit(scene, rays: np.ndarray) -> np.ndarray:
    import lambda_amd._abi as abi
    lib = _load()
    rays = np.ascontiguousarray(rays, np.float32)
    out = np.zeros(len(rays), abi.HIT_DTYPE)
    if lib.oracle_closest_hit(C.byref(scene.view), _p(rays, C.c_float), len(rays), out.ctypes.data_as(C.POINTER(abi.Hit))):
        raise RuntimeError("oracle_closest_hit failed")
    return out


def occluded(scene, rays: np.ndarray, light_instance: np.ndarray, light_object: np.ndarray) -> np.ndarray:
    lib = _load()
    rays = np.ascontiguousarray(rays, np.float32)
    li = np.ascontiguousarray(light_instance, np.int32)
    lo = np.ascontiguousarray(light_object, np.int32)
    out = np.zeros(len(rays), np.int32)
    if lib.oracle_occluded(C.byref(scene.view), _p(rays, C.c_float), _p(li, C.c_int32), _p(lo, C.c_int32),
                           len(rays), _p(out, C.c_int32)):
        raise RuntimeError("oracle_occluded failed")
    return out


def thinfilm(scene, material: int, queries: np.ndarray, with_flags: bool = False):
    """Returns (reference float value, full-FP64 value) per query row
    {cos_theta, lambda, front, px, py, pz, unused}; with_flags adds the
    per-query flag of a square root whose branch is decided by a rounding
    residue."""
    lib = _load()
    q = np.ascontiguousarray(queries, np.float32)
    ref = np.zeros(len(q), np.float32)
    f64 = np.zeros(len(q), np.float64)
    amb = np.zeros(len(q), np.int32)
    if lib.oracle_thinfilm(C.byref(scene.view), material, _p(q, C.c_float), len(q), _p(ref, C.c_float),
                           _p(f64, C.c_double), _p(amb, C.c_int32)):
        raise RuntimeError("oracle_thinfilm: material %d is not a thin film" % material)
    return (ref, f64, amb.astype(bool)) if with_flags else (ref, f64)


def pcg32(seed: int, stream: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    _load().oracle_pcg32(seed, stream, n, _p(out, C.c_uint32))
    return out


STAT_NAMES = ("paths", "closest_rays", "shadow_rays",
              "closest_nodes", "closest_instances", "closest_spheres", "closest_quads", "closest_triangles",
              "shadow_nodes", "shadow_instances", "shadow_spheres", "shadow_quads", "shadow_triangles")


def path_stats(scene, params=None, pixels=None, stride: int = 1, threads: int = 0) -> dict:
    """Traversal work counted on the reference path over every stride-th
    pixel of a pixel range (all samples of each pixel)."""
    lib = _load()
    p = params if params is not None else scene.params
    a, b = (0, p.width * p.height) if pixels is None else pixels
    out = np.zeros(13, np.int64)
    if lib.oracle_path_stats(C.byref(scene.view), C.byref(p), a, b, stride, threads, _p(out, C.c_int64)):
        raise RuntimeError("oracle_path_stats failed")
    return dict(zip(STAT_NAMES, (int(x) for x in out)))


def ray_bytes(stats: dict, kind: str) -> float:
    """Algorithmic bytes per ray of one kind ("closest" or "shadow"): 32 per
    node test, 64 per instance entry, 16/48/36 per sphere/quad/triangle test,
    plus the 128-byte queue record read and written."""
    rays = stats[kind + "_rays"]
    if rays == 0:
        return 0.0
    b = (32.0 * stats[kind + "_nodes"] + 64.0 * stats[kind + "_instances"] + 16.0 * stats[kind + "_spheres"]
         + 48.0 * stats[kind + "_quads"] + 36.0 * stats[kind + "_triangles"])
    return b / rays + 128.0


def algorithmic_bytes(stats: dict) -> float:
    """Total traversal bytes of the counted paths (both ray kinds)."""
    return ray_bytes(stats, "closest") * stats["closest_rays"] + ray_bytes(stats, "shadow") * stats["shadow_rays"]
