"""Schedule autotuner: picks, for a scene and a frame, the schedule of the
integrator -- persistent launch, one launch per bounce with a persistent
tail, or the bounce loop captured in a hipGraph -- with its path-slot
count (queue size), the tile size and cost order of the pixel visiting
order, and the compiled occupancy variant of the persistent kernel (2, 3 or
4 waves per SIMD: its registers capped for that many), by timing short
renders of the same frame at reduced samples per pixel.  The choice is
cached in a JSON file keyed by the scene's shape, the frame's settings and
the device.

The block size (256 threads) and the LDS stack ring (16 entries per lane)
are fixed at compile time; the occupancy variants are the compiled axis.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import LAMBDA_RNG_REFERENCE, DeviceScene

CACHE = os.environ.get("LAMBDA_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "lambda_amd",
                                                             "autotune.json"))


def candidates(npix: int, rng_mode: int) -> list[dict]:
    """The search space: one dict of render-parameter overrides each."""
    base = {"tail_paths": 0, "path_slots": 0, "cost_order": 4, "tile": 8, "occupancy": 4}
    c = [
        base,                                       # persistent, resident slots, slowest 8x8 tiles first
        dict(base, cost_order=8),                   # ... timed over the first 8 samples
        dict(base, cost_order=-1),                  # persistent, tile order
        dict(base, tile=4),                         # 4x4 tiles
        dict(base, tile=16),                        # 16x16 tiles
        dict(base, occupancy=2),                    # kernel variant for 2 waves per SIMD
        dict(base, occupancy=3),                    # ... for 3
        {"tail_paths": 65536, "path_slots": npix, "cost_order": -1, "tile": 8, "occupancy": 3},  # passes + tail
        {"tail_paths": -1, "path_slots": npix, "cost_order": -1, "graph_passes": 8, "tile": 8, "occupancy": 3},
    ]
    if rng_mode != LAMBDA_RNG_REFERENCE:
        c = [dict(x, cost_order=0) for x in c if x["cost_order"] in (4, -1)]
        c.append(dict(base, cost_order=0, path_slots=1 << 20))
        # the voting schedule (wave alternates traversal steps and shading);
        # only with independent streams: with per-pixel sample chains a
        # short trial can favour it while the full frame, bound by its
        # slowest chains, is slower
        c.append(dict(base, cost_order=0, vote=1))
        c.append(dict(base, cost_order=0, vote=1, occupancy=3))
    return c


def _device_name() -> str:
    try:
        from . import device_info
        return device_info()[0]
    except Exception:
        return "no device"


def key(dev: DeviceScene, **params) -> str:
    v = dev.scene.view
    p = dev.scene.render_params(**params)
    shape = (v.num_nodes, v.num_objects, v.num_instances, v.num_materials, p.width, p.height, p.depth,
             p.samples, p.rng_mode, tuple(np.round(dev.scene.arrays()["node_bounds"][0], 3).tolist()),
             _device_name())
    return hashlib.sha1(repr(shape).encode()).hexdigest()[:16]


def tune(dev: DeviceScene, trial_spp: int = 64, repeats: int = 2, cache: str | None = CACHE, **params) -> dict:
    """Returns {"best": overrides, "table": [(overrides, ms per render)],
    "trial_spp": s}; params are the render parameters the choice is for."""
    k = key(dev, **params)
    if cache and os.path.exists(cache):
        with open(cache) as fh:
            data = json.load(fh)
        if k in data:
            return data[k]
    p = dev.scene.render_params(**params)
    root = int(np.sqrt(min(trial_spp, p.samples)))
    trial = dict(params, samples=root * root, sqrt_samples=root)
    table = []
    for cand in candidates(p.width * p.height, p.rng_mode):
        dev.render(**trial, **cand)
        t0 = time.perf_counter()
        for _ in range(repeats):
            dev.render(**trial, **cand)
        table.append((cand, (time.perf_counter() - t0) / repeats * 1e3))
    best = min(table, key=lambda e: e[1])[0]
    out = {"best": best, "table": table, "trial_spp": root * root}
    if cache:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        data = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                data = json.load(fh)
        data[k] = out
        with open(cache, "w") as fh:
            json.dump(data, fh, indent=1)
    return out
