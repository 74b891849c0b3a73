"""lambda_amd: MI355X-native spectral path-tracing integrator for Lambda scenes.

Python mirror of the reference's host interface:

* ``Scene.load(path)`` is ``Parser::parseLRD`` (core/parser.h:29): it parses a
  .lrd file with the C++ front end (liblambda_host.so) and builds the BVH.
* ``render(scene, ...)`` is the render call of src/main.cpp:113-156: it calls
  ``lambda_hip_render`` in liblambda_hip.so and returns the W x H x 3 float
  linear-sRGB image averaged over samples.
* ``write_ppm(path, image)`` is ``Writer::write`` (core/writer.h:17).

Errors from the C ABI are raised as ``LambdaError`` with the library's
message, the way the reference prints its exceptions (src/main.cpp:76-79).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import LAMBDA_RNG_REFERENCE, LAMBDA_RNG_THROUGHPUT, RenderParams, SceneView

__all__ = ["LambdaError", "Scene", "DeviceScene", "device_info", "render", "write_ppm", "LAMBDA_RNG_REFERENCE", "LAMBDA_RNG_THROUGHPUT"]


class LambdaError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def _check(code: int, buf) -> None:
    if code != 0:
        raise LambdaError(code, buf.value.decode(errors="replace"))


class Scene:
    """A parsed scene (host arrays) with its BVH; owns the C++ object."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        self.view = SceneView()
        _abi.host().lambda_scene_get_view(self._h, C.byref(self.view))
        self.params = RenderParams()
        _abi.host().lambda_scene_get_params(self._h, C.byref(self.params))

    @classmethod
    def load(cls, path: str) -> "Scene":
        h = C.c_void_p()
        err = C.create_string_buffer(4096)
        _check(_abi.host().lambda_scene_load(path.encode(), C.byref(h), err, 4096), err)
        return cls(h.value)

    @classmethod
    def parse(cls, text: str, name: str = "<string>") -> "Scene":
        h = C.c_void_p()
        err = C.create_string_buffer(4096)
        _check(_abi.host().lambda_scene_parse_string(text.encode(), name.encode(), C.byref(h), err, 4096), err)
        return cls(h.value)

    def close(self) -> None:
        if self._h is not None and self._h.value:
            _abi.host().lambda_scene_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- numpy copies of the flat arrays (for tests and tools) ----
    def arrays(self) -> dict:
        v = self.view

        def arr(p, n, dtype, shape):
            if n == 0:
                return np.zeros((0,) + shape[1:], dtype)
            ct = np.ctypeslib.as_array(p, shape=(int(np.prod(shape)),))
            return np.array(ct, dtype=dtype).reshape(shape)

        B = _abi.CIE_BINS
        return {
            "spectra": arr(v.spectra, v.num_spectra, np.float32, (v.num_spectra, B)),
            "complex_spectra": arr(v.complex_spectra, v.num_complex_spectra, np.float64, (v.num_complex_spectra, B, 2)),
            "scalar_texture_type": arr(v.scalar_texture_type, v.num_scalar_textures, np.int32, (v.num_scalar_textures,)),
            "scalar_texture_params": arr(v.scalar_texture_params, v.num_scalar_textures, np.float32, (v.num_scalar_textures, 3)),
            "spectrum_texture_type": arr(v.spectrum_texture_type, v.num_spectrum_textures, np.int32, (v.num_spectrum_textures,)),
            "spectrum_texture_index": arr(v.spectrum_texture_index, v.num_spectrum_textures, np.int32, (v.num_spectrum_textures, 2)),
            "spectrum_texture_scale": arr(v.spectrum_texture_scale, v.num_spectrum_textures, np.float32, (v.num_spectrum_textures,)),
            "material_type": arr(v.material_type, v.num_materials, np.int32, (v.num_materials,)),
            "material_params": arr(v.material_params, v.num_materials, np.int32, (v.num_materials, 3)),
            "material_properties": arr(v.material_properties, v.num_material_properties, np.int32, (v.num_material_properties,)),
            "object_type": arr(v.object_type, v.num_objects, np.int32, (v.num_objects,)),
            "object_material": arr(v.object_material, v.num_objects, np.int32, (v.num_objects,)),
            "object_geometry": arr(v.object_geometry, v.num_objects, np.float32, (v.num_objects, 12)),
            "object_shading": arr(v.object_shading, v.num_objects, np.float32, (v.num_objects, 15)),
            "instance_ints": arr(v.instance_ints, v.num_instances, np.int32, (v.num_instances, 4)),
            "instance_matrices": arr(v.instance_matrices, v.num_instances, np.float32, (v.num_instances, 32)),
            "node_bounds": arr(v.node_bounds, v.num_nodes, np.float32, (v.num_nodes, 6)),
            "node_ints": arr(v.node_ints, v.num_nodes, np.int32, (v.num_nodes, 2)),
            "light_instances": arr(v.light_instances, v.num_lights, np.int32, (v.num_lights,)),
            "light_objects": arr(v.light_objects, v.num_lights, np.int32, (v.num_lights,)),
            "light_powers": arr(v.light_powers, v.num_lights, np.float32, (v.num_lights,)),
            "total_light_power": np.float32(v.total_light_power),
            "background_texture": int(v.background_texture),
            "camera": np.array(v.camera[:], np.float32),
        }

    def render_params(self, **overrides) -> RenderParams:
        p = RenderParams()
        C.memmove(C.byref(p), C.byref(self.params), C.sizeof(RenderParams))
        for k, val in overrides.items():
            setattr(p, k, val)
        return p


def render(scene: Scene, rng_mode: int = LAMBDA_RNG_REFERENCE, seed: int = 42, path_slots: int = 0,
           progress: np.ndarray | None = None) -> np.ndarray:
    """Renders on the GPU through lambda_hip_render; returns float32 [H, W, 3]."""
    p = scene.render_params(rng_mode=rng_mode, seed=seed, path_slots=path_slots)
    out = np.zeros((p.height, p.width, 3), np.float32)
    err = C.create_string_buffer(4096)
    prog = None if progress is None else _abi.ptr(progress, C.c_int32)
    _check(_abi.hip().lambda_hip_render(C.byref(scene.view), C.byref(p), _abi.ptr(out, C.c_float), prog, err, 4096), err)
    return out


def write_ppm(path: str, image: np.ndarray) -> None:
    img = np.ascontiguousarray(image, np.float32)
    err = C.create_string_buffer(4096)
    _check(_abi.host().lambda_write_ppm(path.encode(), img.shape[1], img.shape[0], _abi.ptr(img, C.c_float), err, 4096), err)


class DeviceScene:
    """A scene uploaded to the GPU (lambda_hip_scene_create); the device-side
    counterpart of the arrays main.cpp copies to device memory
    (src/main.cpp:84-103)."""

    def __init__(self, scene: Scene, gpu_bvh: bool = False):
        """gpu_bvh: build the BVHs on the GPU (lambda_hip_scene_create_lbvh)
        instead of using the host-built ones; object_order and
        instance_order then map the renumbered objects and instances to the
        scene's, and bvh_build_ms is the build's time."""
        self._lib = _abi.hip()
        self._h = C.c_void_p()
        self.scene = scene
        self.object_order = self.instance_order = None
        self.bvh_build_ms = None
        err = C.create_string_buffer(4096)
        if gpu_bvh:
            v = scene.view
            self.object_order = np.zeros(v.num_objects, np.int32)
            self.instance_order = np.zeros(v.num_instances, np.int32)
            ms = C.c_double()
            _check(self._lib.lambda_hip_scene_create_lbvh(
                C.byref(v), C.byref(self._h), _abi.ptr(self.object_order, C.c_int32),
                _abi.ptr(self.instance_order, C.c_int32), C.byref(ms), err, 4096), err)
            self.bvh_build_ms = ms.value
        else:
            _check(self._lib.lambda_hip_scene_create(C.byref(scene.view), C.byref(self._h), err, 4096), err)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.lambda_hip_scene_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, rng_mode: int = LAMBDA_RNG_REFERENCE, seed: int = 42, path_slots: int = 0,
               progress: np.ndarray | None = None, out: np.ndarray | None = None, **overrides) -> np.ndarray:
        """Renders into out (float32 [H, W, 3], C order) when given, as main.cpp
        renders into the h_output it allocates once, else into a new array."""
        p = self.scene.render_params(rng_mode=rng_mode, seed=seed, path_slots=path_slots, **overrides)
        if out is None:
            out = np.zeros((p.height, p.width, 3), np.float32)
        elif out.shape != (p.height, p.width, 3) or out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-ordered float32 array of shape (%d, %d, 3)" % (p.height, p.width))
        err = C.create_string_buffer(4096)
        prog = None if progress is None else _abi.ptr(progress, C.c_int32)
        _check(self._lib.lambda_hip_scene_render(self._h, C.byref(p), _abi.ptr(out, C.c_float), prog, err, 4096), err)
        return out

    def render_stats(self, rng_mode: int = LAMBDA_RNG_REFERENCE, seed: int = 42, path_slots: int = 0,
                     **overrides):
        """Renders with per-kernel HIP-event timing; returns (image, stats dict)."""
        p = self.scene.render_params(rng_mode=rng_mode, seed=seed, path_slots=path_slots, **overrides)
        out = np.zeros((p.height, p.width, 3), np.float32)
        st = _abi.RenderStats()
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_scene_render_stats(self._h, C.byref(p), _abi.ptr(out, C.c_float), C.byref(st),
                                                       err, 4096), err)
        stats = {"iterations": st.iterations, "path_slots": st.path_slots, "closest_rays": st.closest_rays,
                 "shadow_rays": st.shadow_rays, "render_ms": st.render_ms}
        for k, name in enumerate(_abi.KERNELS):
            stats[name + "_ms"] = st.kernel_ms[k]
            stats[name + "_launches"] = st.launches[k]
        return out, stats

    def info(self) -> dict:
        """What the library built from the scene: traversal stack depth (and
        the part of it kept in LDS), 4-wide node count, counts and bounds."""
        st = _abi.SceneInfo()
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_scene_info(self._h, C.byref(st), err, 4096), err)
        return {"stack_need": st.stack_need, "stack_lds": st.stack_lds, "wide_nodes": st.wide_nodes,
                "num_nodes": st.num_nodes, "num_objects": st.num_objects, "num_instances": st.num_instances,
                "bounds": list(st.bounds)}

    def closest_hit(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32)
        out = np.zeros(len(rays), _abi.HIT_DTYPE)
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_closest_hit(self._h, _abi.ptr(rays, C.c_float), len(rays),
                                                out.ctypes.data_as(C.POINTER(_abi.Hit)), err, 4096), err)
        return out

    def occluded(self, rays: np.ndarray, light_instance: np.ndarray, light_object: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32)
        li = np.ascontiguousarray(light_instance, np.int32)
        lo = np.ascontiguousarray(light_object, np.int32)
        out = np.zeros(len(rays), np.int32)
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_occluded(self._h, _abi.ptr(rays, C.c_float), _abi.ptr(li, C.c_int32),
                                             _abi.ptr(lo, C.c_int32), len(rays), _abi.ptr(out, C.c_int32),
                                             err, 4096), err)
        return out

    def thinfilm(self, material: int, queries: np.ndarray, points: np.ndarray) -> np.ndarray:
        q = np.ascontiguousarray(queries, np.float32)
        pts = np.ascontiguousarray(points, np.float32)
        out = np.zeros(len(q), np.float32)
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_thinfilm_reflectance(self._h, material, _abi.ptr(q, C.c_float),
                                                         _abi.ptr(pts, C.c_float), len(q),
                                                         _abi.ptr(out, C.c_float), err, 4096), err)
        return out


def device_info() -> tuple[str, int]:
    name = C.create_string_buffer(256)
    cu = C.c_int32(0)
    err = C.create_string_buffer(4096)
    _check(_abi.hip().lambda_hip_device_info(name, 256, C.byref(cu), err, 4096), err)
    return name.value.decode(), cu.value


GUIDE_PARAMS = 22848
GUIDE_OUTPUTS = 40


class Guide:
    """The path-guiding network on the GPU (lambda_hip_guide_*)."""

    def __init__(self, seed: int = 1):
        self._lib = _abi.hip()
        self._h = C.c_void_p()
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_guide_create(seed, C.byref(self._h), err, 4096), err)

    def close(self):
        if self._h:
            self._lib.lambda_hip_guide_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(GUIDE_PARAMS, np.float32)
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_guide_get_params(self._h, _abi.ptr(out, C.c_float), err, 4096), err)
        return out

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        assert p.shape == (GUIDE_PARAMS,)
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_guide_set_params(self._h, _abi.ptr(p, C.c_float), err, 4096), err)

    def forward(self, features: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(features, np.float32).reshape(-1, 6)
        out = np.zeros((len(f), GUIDE_OUTPUTS), np.float32)
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_guide_forward(self._h, _abi.ptr(f, C.c_float), len(f),
                                                  _abi.ptr(out, C.c_float), err, 4096), err)
        return out

    def gradient(self, samples: np.ndarray):
        s = np.ascontiguousarray(samples, np.float32).reshape(-1, 10)
        g = np.zeros(GUIDE_PARAMS, np.float32)
        loss = C.c_double()
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_guide_gradient(self._h, _abi.ptr(s, C.c_float), len(s), _abi.ptr(g, C.c_float),
                                                   C.byref(loss), err, 4096), err)
        return g, loss.value

    def adam_step(self, grads: np.ndarray, scale: float = 1.0, lr: float = 1e-3) -> None:
        """One Adam step with gradient grads * scale (lambda_hip_guide_adam_step)."""
        g = np.ascontiguousarray(grads, np.float32)
        assert g.shape == (GUIDE_PARAMS,)
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_guide_adam_step(self._h, _abi.ptr(g, C.c_float), scale, lr, err, 4096), err)

    def state(self) -> dict:
        """Adam moments m, v, the step count and the two bfloat16 weight
        copies (as float32): w_nk = W_l^T per layer, w_kn = W_l per layer."""
        m = np.zeros(GUIDE_PARAMS, np.float32)
        v = np.zeros(GUIDE_PARAMS, np.float32)
        step = C.c_int32()
        nk = np.zeros(_abi.GUIDE_WEIGHTS_BF16, np.uint16)
        kn = np.zeros(_abi.GUIDE_WEIGHTS_BF16, np.uint16)
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_guide_get_state(self._h, _abi.ptr(m, C.c_float), _abi.ptr(v, C.c_float),
                                                    C.byref(step), _abi.ptr(nk, C.c_uint16),
                                                    _abi.ptr(kn, C.c_uint16), err, 4096), err)

        def widen(h):
            return (h.astype(np.uint32) << 16).view(np.float32)

        return {"m": m, "v": v, "step": step.value, "w_nk": widen(nk), "w_kn": widen(kn)}

    def benchmark(self, rows: int = 1 << 20, iters: int = 10):
        """(forward ms, gradient ms) per pass over `rows` device-resident rows."""
        f, g = C.c_double(), C.c_double()
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_guide_benchmark(self._h, rows, iters, C.byref(f), C.byref(g), err, 4096), err)
        return f.value, g.value

    def train(self, samples: np.ndarray, batch: int, lr: float = 1e-3) -> float:
        s = np.ascontiguousarray(samples, np.float32).reshape(-1, 10)
        loss = C.c_double()
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_guide_train(self._h, _abi.ptr(s, C.c_float), len(s), batch, lr,
                                                C.byref(loss), err, 4096), err)
        return loss.value


class Progressive:
    """Progressive rendering session (lambda_hip_progressive_*): plain when
    guide is None, guided otherwise.  The scene and guide must outlive it.
    Guided sessions run the network inside the persistent kernel; with
    tail_paths=-1 they run one launch per bounce with the network as its own
    kernel instead."""

    def __init__(self, dev: "DeviceScene", guide: "Guide | None" = None, path_slots: int = 0, seed: int = 42,
                 **overrides):
        self._lib = _abi.hip()
        self.dev, self.guide = dev, guide
        p = dev.scene.render_params(path_slots=path_slots, seed=seed, **overrides)
        self.width, self.height = p.width, p.height
        self._h = C.c_void_p()
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_progressive_create(dev._h, C.byref(p), guide._h if guide else None,
                                                       C.byref(self._h), err, 4096), err)

    def step(self, samples: int = 1, train: bool = True, keep: bool = True) -> dict:
        """Renders `samples` more samples per pixel; train: train the guide
        on them afterwards; keep: add them to the image."""
        st = _abi.ProgressiveStats()
        err = C.create_string_buffer(4096)
        flags = (1 if train else 0) | (0 if keep else 2)
        _check(self._lib.lambda_hip_progressive_step(self._h, samples, flags, C.byref(st), err, 4096), err)
        return {"samples": st.samples, "train_samples": st.train_samples, "render_ms": st.render_ms,
                "train_ms": st.train_ms, "loss": st.loss, "closest_rays": st.closest_rays,
                "shadow_rays": st.shadow_rays}

    def set_training(self, epochs: int = 1, batch: int = 1 << 14, lr: float = 1e-3, guided_bounces: int = 0) -> None:
        """Adam epochs per step, batch, rate, and the vertices guided: those
        with a bounce index below guided_bounces (0 = the default of 8; a
        value above the depth guides every vertex)."""
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_progressive_set_training(self._h, epochs, batch, lr, guided_bounces, err, 4096),
               err)

    def set_guide_table(self, cells: int) -> None:
        """Guide from a table of the network's mixtures, re-evaluated after
        every training, over cubic cells (`cells` along the scene box's
        longest side) and the six axis normals, instead of running the
        network at every guided vertex; 0 = the network.  Before the first
        step."""
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_progressive_set_guide_table(self._h, cells, err, 4096), err)

    def set_guide_mix(self, frac: float) -> None:
        """Probability of sampling the guide at a guided vertex (default
        0.5; at most 0.95, the cosine lobe takes the rest)."""
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_progressive_set_guide_mix(self._h, frac, err, 4096), err)

    def set_roulette(self, start_bounce: int) -> None:
        """Bounce index from which Russian roulette applies (plain sessions:
        5, the reference's; guided sessions: off); above the path depth it
        is off."""
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_progressive_set_roulette(self._h, start_bounce, err, 4096), err)

    def image(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 3), np.float32)
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_progressive_image(self._h, _abi.ptr(out, C.c_float), err, 4096), err)
        return out

    def close(self):
        if self._h:
            self._lib.lambda_hip_progressive_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def finite_pixels(*imgs: np.ndarray) -> np.ndarray:
    """Mask of the pixels whose channels are finite in every image (the
    reference estimator yields 0/0 in rare pixels when randomFloat() returns
    1.0; render/renderer.h:179-183)."""
    m = np.ones(imgs[0].shape[:2], bool)
    for im in imgs:
        m &= np.isfinite(im).all(-1)
    return m


def rel_mse(img: np.ndarray, ref: np.ndarray) -> float:
    """Mean over pixels and channels of (I - R)^2 / (R^2 + 0.01), over the
    pixels finite in both images."""
    m = finite_pixels(img, ref)
    a = img[m].astype(np.float64)
    r = ref[m].astype(np.float64)
    return float(np.mean((a - r) ** 2 / (r * r + 0.01)))


def serve(devices: "list[DeviceScene]", jobs: "list[dict]", hbm_budget: int = 0):
    """Renders many jobs in batched persistent launches (lambda_hip_serve).
    jobs: dicts {scene: index into devices, arrival_ms, and render-parameter
    overrides (width, height, depth, samples, sqrt_samples, lambda_min,
    lambda_max, seed)}.  Returns (images, results, stats)."""
    lib = _abi.hip()
    handles = (C.c_void_p * len(devices))(*[d._h for d in devices])
    arr = (_abi.Job * len(jobs))()
    images = []
    for k, jd in enumerate(jobs):
        jd = dict(jd)
        sc = jd.pop("scene")
        arrival = float(jd.pop("arrival_ms", 0.0))
        p = devices[sc].scene.render_params(**jd)
        img = np.zeros((p.height, p.width, 3), np.float32)
        images.append(img)
        arr[k].scene = sc
        arr[k].params = p
        arr[k].arrival_ms = arrival
        arr[k].out_rgb = _abi.ptr(img, C.c_float)
    res = (_abi.JobResult * len(jobs))()
    st = _abi.ServeStats()
    err = C.create_string_buffer(4096)
    _check(lib.lambda_hip_serve(handles, len(devices), arr, len(jobs), hbm_budget, res, C.byref(st), err, 4096), err)
    results = [{"latency_ms": r.latency_ms, "finish_ms": r.finish_ms, "batch": r.batch} for r in res]
    stats = {"batches": st.batches, "paths": st.paths, "closest_rays": st.closest_rays,
             "shadow_rays": st.shadow_rays, "wall_ms": st.wall_ms}
    return images, results, stats


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3


class Server:
    """Live serving (lambda_hip_server_*): submit jobs while earlier ones
    render; admission by HBM budget, overlapping persistent launches."""

    def __init__(self, devices: "list[DeviceScene]", hbm_budget: int = 0, max_launches: int = 0):
        self._lib = _abi.hip()
        self.devices = devices
        self._handles = (C.c_void_p * len(devices))(*[d._h for d in devices])
        self._h = C.c_void_p()
        self._keep = {}
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_server_create(self._handles, len(devices), hbm_budget, max_launches,
                                                  C.byref(self._h), err, 4096), err)

    def submit(self, scene: int, **params) -> "tuple[int, np.ndarray]":
        """Queues a job on devices[scene] with render-parameter overrides;
        returns (ticket, image buffer filled when the job is done)."""
        # settings from the job's scene (the library rejects an index out of range)
        p = self.devices[scene if 0 <= scene < len(self.devices) else 0].scene.render_params(**params)
        img = np.zeros((p.height, p.width, 3), np.float32)
        job = _abi.Job()
        job.scene = scene
        job.params = p
        job.arrival_ms = 0.0
        job.out_rgb = _abi.ptr(img, C.c_float)
        t = C.c_int64()
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_server_submit(self._h, C.byref(job), C.byref(t), err, 4096), err)
        self._keep[t.value] = img
        return t.value, img

    def poll(self, ticket: int) -> "tuple[int, dict | None]":
        st = C.c_int32()
        res = _abi.JobResult()
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_server_poll(self._h, ticket, C.byref(st), C.byref(res), err, 4096), err)
        if st.value != JOB_DONE:
            return st.value, None
        return st.value, {"latency_ms": res.latency_ms, "finish_ms": res.finish_ms, "batch": res.batch}

    def drain(self) -> dict:
        st = _abi.ServeStats()
        err = C.create_string_buffer(4096)
        _check(self._lib.lambda_hip_server_drain(self._h, C.byref(st), err, 4096), err)
        return {"batches": st.batches, "paths": st.paths, "closest_rays": st.closest_rays,
                "shadow_rays": st.shadow_rays, "wall_ms": st.wall_ms}

    def close(self):
        if self._h:
            self._lib.lambda_hip_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

