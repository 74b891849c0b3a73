"""Test infrastructure: the CPU oracle's render speed against the reference
harness (oracle/_ref/ref_harness, built by `make -C oracle ref` from the
reference sources) on the same scene, threads and samples, with the images
compared bit for bit.  The oracle is the CPU baseline of the benchmark, so
its speed must match the reference CPU integrator's (within 5 %) for the
GPU/CPU ratio to mean what it says.

Runs where the reference sources are present (not on a machine without
them).  Usage: python oracle/speed_check.py [--threads 8] [--runs 3]
Prints one JSON line per scene: median render seconds of each, their ratio,
and whether the images are identical.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from oracle import ref_dump  # noqa: E402


def check(lrd: str, threads: int, runs: int) -> dict:
    import lambda_amd
    cwd = os.getcwd()
    os.chdir(os.path.dirname(lrd))
    try:
        scene = lambda_amd.Scene.load(os.path.basename(lrd))
    finally:
        os.chdir(cwd)
    ref_s, ora_s = [], []
    ref_img = ora_img = None
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "ref.f32")
        for _ in range(runs):
            img, info = ref_dump.render(lrd, out, threads)
            ref_s.append(info["render_s"])
            ref_img = img
            t0 = time.perf_counter()
            ora_img = oracle.render(scene, threads=threads)
            ora_s.append(time.perf_counter() - t0)
    p = scene.params
    same = bool(np.array_equal(ref_img.view(np.uint32), ora_img.reshape(-1).view(np.uint32)))
    ref_m, ora_m = float(np.median(ref_s)), float(np.median(ora_s))
    return {"scene": os.path.basename(lrd), "width": p.width, "height": p.height, "spp": p.samples,
            "depth": p.depth, "threads": threads, "reference_s": ref_s, "oracle_s": ora_s,
            "reference_median_s": ref_m, "oracle_median_s": ora_m, "oracle_over_reference": ora_m / ref_m,
            "images_bit_identical": same}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--runs", type=int, default=3)
    args = ap.parse_args()
    if not ref_dump.available():
        print("the reference harness is not built (make -C oracle ref)", file=sys.stderr)
        return 2
    from lambda_amd import scenes
    with tempfile.TemporaryDirectory() as d:
        cases = [scenes.write_closed_box(os.path.join(d, "box"), samples=64),
                 scenes.write_thinfilm_grid(os.path.join(d, "film"), samples=16),
                 # the benchmark's instanced mesh at icosphere level 5 (20,480
                 # triangles per instance), where the reference's BVH build is short
                 scenes.write_instanced_mesh(os.path.join(d, "mesh"), width=480, height=270, samples=16, level=5)]
        for lrd in cases:
            print(json.dumps(check(lrd, args.threads, args.runs)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
