#!/usr/bin/env python3
"""Benchmark of the plain GPU path tracer: paths/s at 1080p on one MI355X.

Workload (default): the instanced-mesh scene -- a level-8 perlin-displaced
icosphere (1,310,720 triangles) instanced 64 times with seeded rotation and
scale 0.5-1.5, round-robin Lambertian / mirror / dielectric / thin-film
materials, 4 quad emitters and a constant environment; 1920x1080, depth 10,
64 spp, reference PCG32 streams.  One "step" is one full render of the frame
(132.7M camera paths) with the scene already resident in HBM; the timed
region of each step ends when the RGB image is back in host memory.

Printed JSON line (rank 0, one GPU):
  value        paths/s over the K timed renders (reference PCG32 streams)
  roofline     the kernel that traces the frame's rays (k_persist in the
               default persistent schedule, k_trace with one launch per
               bounce): algorithmic bytes per launch (oracle-counted node,
               instance and primitive tests per ray of each kind on a strided
               pixel sample, times the rays of each kind the launch traced)
               over its mean launch time from HIP events, against 8 TB/s
  cpu_baseline the CPU oracle (bit-identical restatement of the reference
               CPU path) on a band of rows, all host threads
  throughput_streams  the same frame with independent streams per sample
  guided_vs_plain     wall time to relMSE 1e-3 against a 65536-spp plain
               reference on the room lit through a door gap (320x180, depth
               16): the plain tracer, and the median of 3 guided runs (the
               guide's mixtures tabulated over 48 cells along the room's
               longest side, one-sample mixture 0.5 cos + 0.5 guide)
  guide_mlp_roofline  the guiding network's forward and gradient kernels,
               algorithmic FLOP per second against 2.5 PFLOP/s dense BF16
  batched_serving     32 seeded jobs with Poisson arrivals in persistent
               launches: aggregate paths/s and latency p50 / p99

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "paths/s @1080p; wall time to relMSE 1e-3, guided vs plain, 1 MI355X"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_workload(name: str, scene_dir: str):
    from lambda_amd import scenes
    if name == "instanced_mesh":
        path = scenes.write_instanced_mesh(os.path.join(scene_dir, "instanced_mesh"))
        cfg = {"workload": "instanced_mesh_1080p", "width": 1920, "height": 1080, "depth": 10, "spp": 64,
               "mesh": "icosphere level 8, perlin-displaced (amplitude 0.05, frequency 6)",
               "triangles_per_instance": 1310720, "instances": 64, "emitters": 4}
    elif name == "closed_box":
        path = scenes.write_closed_box(os.path.join(scene_dir, "closed_box"))
        cfg = {"workload": "closed_box_256", "width": 256, "height": 256, "depth": 8, "spp": 16}
    elif name == "thinfilm_grid":
        path = scenes.write_thinfilm_grid(os.path.join(scene_dir, "thinfilm_grid"))
        cfg = {"workload": "thinfilm_grid_512", "width": 512, "height": 512, "depth": 12, "spp": 64}
    else:
        raise SystemExit("unknown workload " + name)
    return path, cfg


def cpu_baseline(scene, seconds: float, threads: int) -> dict:
    """Times the CPU oracle on the full frame at the largest square spp (up
    to the workload's) whose estimated time fits `seconds`, from a 1-spp
    calibration render; below 1 spp it renders a band of rows.  paths/s of
    this integrator does not depend on spp to first order."""
    import oracle
    p = scene.params
    q = scene.render_params(samples=1, sqrt_samples=1)
    t0 = time.perf_counter()
    oracle.render(scene, q, threads=threads)
    t1 = time.perf_counter() - t0
    root = 1
    while (root + 1) ** 2 <= p.samples and t1 * (root + 1) ** 2 <= seconds:
        root += 1
    rows = p.height
    if t1 > seconds:
        rows = max(1, int(p.height * seconds / t1))
    spp = root * root
    q = scene.render_params(samples=spp, sqrt_samples=root)
    r0 = (p.height - rows) // 2
    t0 = time.perf_counter()
    oracle.render(scene, q, threads=threads, rows=(r0, r0 + rows))
    t = time.perf_counter() - t0
    paths = rows * p.width * spp
    return {"value": paths / t, "unit": "paths/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "host_cpus": os.cpu_count(),
            "sample": "rows %d-%d of %d (%d pixels) at %d spp, %.1f s; CPU oracle, bit-identical restatement of "
                      "the reference CPU integrator, OpenMP guided schedule over rows" % (
                          r0, r0 + rows, p.height, rows * p.width, spp, t)}


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_to_target(dev, guide, ref, target: float, step_spp: int, max_spp: int, warmup: int = 0,
                   train_steps: int = 1 << 30, epochs: int = 1, batch: int = 1 << 14, guided_bounces: int = 0,
                   table_cells: int = 0, mix: float = 0.5, rr_start: int = 0):
    """Progressive render until relMSE against ref <= target; returns the
    render-and-train wall time at which the target was crossed (log-linear
    interpolation between steps), the samples per pixel then, and the curve.
    With a guide: `warmup` steps train without adding to the image, then
    the first `train_steps` kept steps train too; table_cells > 0 guides
    from the tabulated network.  relMSE evaluation is not timed."""
    import lambda_amd
    p = lambda_amd.Progressive(dev, guide)
    if rr_start:
        p.set_roulette(rr_start)
    if guide is not None:
        p.set_training(epochs, batch, 1e-3, guided_bounces)
        if table_cells:
            p.set_guide_table(table_cells)
        p.set_guide_mix(mix)
    elapsed, spp, curve = 0.0, 0, []
    prev = None
    hit = None
    if guide is not None:
        for _ in range(warmup):
            t0 = time.perf_counter()
            p.step(step_spp, train=True, keep=False)
            elapsed += time.perf_counter() - t0
    kept = 0
    train_ms = []
    while spp < max_spp:
        t0 = time.perf_counter()
        st = p.step(step_spp, train=guide is not None and kept < train_steps)
        if st["train_ms"] > 0:
            train_ms.append(st["train_ms"])
        elapsed += time.perf_counter() - t0
        kept += 1
        spp += step_spp
        err = lambda_amd.rel_mse(p.image(), ref)
        curve.append((round(elapsed, 4), spp, err))
        if kept % 50 == 0:
            log("  %s: %d spp, relMSE %.3g, %.1f s" % ("guided" if guide is not None else "plain", spp, err, elapsed))
        if err <= target:
            if prev is None:
                hit = (elapsed, spp)
            else:
                import math
                (t1, e1), (t2, e2) = (prev[0], prev[2]), (elapsed, err)
                f = (math.log(e1) - math.log(target)) / max(math.log(e1) - math.log(e2), 1e-12)
                hit = (t1 + f * (t2 - t1), spp)
            break
        prev = (elapsed, spp, err)
    p.close()
    if train_ms:
        curve.append(("train_ms_per_training_step", sum(train_ms) / len(train_ms), len(train_ms)))
    return hit, curve


def guided_leg(args, scene_dir: str) -> dict:
    """Wall time to relMSE 1e-3 against a plain reference of 65536 spp on the
    room lit through a door gap, plain tracer vs guided (the guide trains on
    the paths of the first steps); the guided figure is the median of
    args.guided_runs runs."""
    import lambda_amd
    from lambda_amd import scenes
    w, h = (int(x) for x in args.guided_res.split("x"))
    lrd = scenes.write_guiding_room(os.path.join(scene_dir, "guiding_room"), width=w, height=h, depth=16, samples=1)
    scene = lambda_amd.Scene.load(lrd)
    dev = lambda_amd.DeviceScene(scene)
    t0 = time.perf_counter()
    if args.guided_ref_load:
        import numpy as np
        ref = np.load(args.guided_ref_load)
        if ref.shape != (h, w, 3):
            raise SystemExit("%s holds a %s image, not %dx%d" % (args.guided_ref_load, ref.shape, w, h))
        log("guided leg: %dx%d reference loaded from %s" % (w, h, args.guided_ref_load))
    else:
        # the reference's streams are independent of the measured runs
        ref_session = lambda_amd.Progressive(dev, None, seed=0x5EED0F2EF)
        steps = args.guided_ref_spp // 256
        for k in range(steps):
            ref_session.step(256, train=False)
            if (k + 1) % 32 == 0:
                log("guided leg: reference %d of %d spp" % (256 * (k + 1), args.guided_ref_spp))
        ref = ref_session.image()
        ref_session.close()
        t_ref = time.perf_counter() - t0
        log("guided leg: %dx%d reference at %d spp in %.1f s" % (w, h, args.guided_ref_spp, t_ref))
        if args.guided_ref_save:
            import numpy as np
            np.save(args.guided_ref_save, ref)
            if args.guided_ref_only:
                dev.close()
                return {"reference_saved": args.guided_ref_save, "width": w, "height": h,
                        "reference_spp": args.guided_ref_spp, "seconds": t_ref}
    plain, plain_curve = time_to_target(dev, None, ref, 1e-3, args.guided_step, args.guided_max_spp)
    log("guided leg: plain reaches relMSE 1e-3 at %s" % (plain,))
    # the plain tracer without Russian roulette (as the guided sessions run):
    # the share of the guided gain that comes from the roulette setting alone
    plain_nr, _ = time_to_target(dev, None, ref, 1e-3, args.guided_step, args.guided_max_spp, rr_start=17)
    log("guided leg: plain without roulette reaches relMSE 1e-3 at %s" % (plain_nr,))
    # Training is not deterministic (the order of the gradient atomics and of
    # the training-sample buffer), so the guided time varies from run to run:
    # several runs with their own guide seeds, the median reported.
    runs = []
    for k in range(max(1, args.guided_runs)):
        guide = lambda_amd.Guide(seed=args.guided_seed + k)
        hit, curve = time_to_target(dev, guide, ref, 1e-3, args.guided_step, args.guided_max_spp,
                                    warmup=args.guided_warmup, train_steps=args.guided_train_steps,
                                    epochs=args.guided_epochs, batch=args.guided_batch,
                                    guided_bounces=args.guided_bounces, table_cells=args.guided_table_cells,
                                    mix=args.guided_mix)
        guide.close()
        log("guided leg: guided run %d reaches relMSE 1e-3 at %s" % (k, hit))
        runs.append((hit, curve))
    dev.close()
    done = sorted((r for r in runs if r[0]), key=lambda r: r[0][0])
    # the median run; none when a run did not reach the target
    guided, guided_curve = done[(len(done) - 1) // 2] if len(done) == len(runs) else (None, runs[0][1])
    train_ms = [c[-1][1] for _, c in runs if c and c[-1][0] == "train_ms_per_training_step"]
    curves = [[x for x in c if x[0] != "train_ms_per_training_step"] for c in (plain_curve, guided_curve)]
    plain_curve, guided_curve = curves
    out = {"workload": "guiding_room", "width": w, "height": h, "depth": 16,
           "reference_spp": args.guided_ref_spp, "target_rel_mse": 1e-3, "step_spp": args.guided_step,
           "guided_warmup_steps": args.guided_warmup, "guided_train_steps": min(args.guided_train_steps, 1 << 30),
           "guided_epochs": args.guided_epochs, "guided_batch": args.guided_batch, "adam_lr": 1e-3,
           "guided_bounces": args.guided_bounces or 8, "guide_fraction": args.guided_mix,
           "guided_schedule": ("persistent kernel, mixtures looked up in a table of the network's outputs "
                               "(%d cells along the longest side x 6 axis normals), rebuilt after each training"
                               % args.guided_table_cells) if args.guided_table_cells else
                              "persistent kernel, network evaluated per wave on the matrix cores",
           "plain_s": plain[0] if plain else None, "plain_spp": plain[1] if plain else None,
           "plain_no_roulette_s": plain_nr[0] if plain_nr else None,
           "plain_no_roulette_spp": plain_nr[1] if plain_nr else None,
           "roulette": "plain: from bounce 5 (the reference's); guided: off",
           "guided_runs": len(runs), "guided_s_runs": [r[0][0] if r[0] else None for r in runs],
           "guided_spp_runs": [r[0][1] if r[0] else None for r in runs],
           "guided_s": guided[0] if guided else None, "guided_spp": guided[1] if guided else None,
           "train_ms_per_training_step": sum(train_ms) / len(train_ms) if train_ms else None,
           "plain_curve": plain_curve[::max(1, len(plain_curve) // 12)],
           "guided_curve": guided_curve[::max(1, len(guided_curve) // 12)]}
    if plain and guided:
        out["guided_over_plain_time"] = guided[0] / plain[0]
    if plain_nr and guided:
        out["guided_over_plain_no_roulette_time"] = guided[0] / plain_nr[0]
    return out


def serve_leg(args, mesh_dev, scene_dir: str) -> dict:
    """32 seeded jobs, random variants of the closed box, the thin-film grid
    and the instanced mesh (side 128-1024 pixels, depth 4-16, 16-256 spp in
    perfect squares, 380-780 or 360-830 nm), submitted to the live server
    by a host thread with Poisson arrivals (mean gap 50 ms) and admitted by
    HBM budget into overlapping persistent launches."""
    import threading
    import numpy as np
    import lambda_amd
    from lambda_amd import scenes
    box = lambda_amd.Scene.load(scenes.write_closed_box(os.path.join(scene_dir, "closed_box")))
    film = lambda_amd.Scene.load(scenes.write_thinfilm_grid(os.path.join(scene_dir, "thinfilm_grid")))
    devs = [lambda_amd.DeviceScene(box), lambda_amd.DeviceScene(film), mesh_dev]
    rng = np.random.default_rng(args.serve_seed)
    t, jobs = 0.0, []
    for _ in range(args.serve_jobs):
        t += rng.exponential(args.serve_gap_ms)
        side = int(rng.integers(128, 1025))
        root = int(rng.integers(4, 17))
        lo, hi = (380.0, 780.0) if rng.random() < 0.5 else (360.0, 830.0)
        jobs.append(dict(scene=int(rng.integers(0, 3)), width=side, height=side, depth=int(rng.integers(4, 17)),
                         samples=root * root, sqrt_samples=root, lambda_min=lo, lambda_max=hi, arrival_ms=t))
    srv = lambda_amd.Server(devs, hbm_budget=args.serve_budget, max_launches=args.serve_launches)
    tickets = []

    def feed():
        t0 = time.perf_counter()
        for jd in jobs:
            jd = dict(jd)
            due = jd.pop("arrival_ms") / 1e3
            delay = due - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(jd.pop("scene"), **jd)[0])

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    st = srv.drain()
    results = [srv.poll(tk)[1] for tk in tickets]
    srv.close()
    for d in devs[:2]:
        d.close()
    lat = np.array([r["latency_ms"] for r in results])
    batch_px = {}
    for jd, r in zip(jobs, results):
        batch_px[r["batch"]] = batch_px.get(r["batch"], 0) + jd["width"] * jd["height"]
    return {"jobs": len(jobs), "batches": st["batches"], "paths": st["paths"], "wall_ms": st["wall_ms"],
            "paths_per_s": st["paths"] / (st["wall_ms"] / 1e3),
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(lat.max())},
            "last_arrival_ms": jobs[-1]["arrival_ms"], "hbm_budget": args.serve_budget,
            "max_launches": args.serve_launches, "largest_batch_bytes": 36 * max(batch_px.values()),
            "mode": "live server: host-timed Poisson submissions, HBM-budget admission, overlapping launches",
            "seed": args.serve_seed}


# FLOP per row of the guiding network (2 per multiply-add).  Forward: 96x64
# + 3 x 64x64 + 64x40 (the 40 outputs used; the inference kernel computes 48
# output columns, 43,008 FLOP on the matrix cores).  Gradient, as computed
# (output padded to 64): the padded forward + the backward data and weight
# products (2x the forward's, layer 0 without the data product).
GUIDE_FWD_FLOP = 2 * (96 * 64 + 3 * 64 * 64 + 64 * 40)
GUIDE_PADDED_FWD_FLOP = 2 * (96 * 64 + 4 * 64 * 64)
GUIDE_GRAD_FLOP = GUIDE_PADDED_FWD_FLOP + 2 * (2 * (96 * 64 + 4 * 64 * 64) - 96 * 64)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def guide_kernels(rows_fwd: int = 1 << 20, rows_grad: int = 1 << 16) -> dict:
    """Throughput of the guiding network's BF16 MFMA kernels against the
    dense BF16 matrix peak."""
    import lambda_amd
    g = lambda_amd.Guide(seed=1)
    f_ms, _ = g.benchmark(rows_fwd, 20)
    _, g_ms = g.benchmark(rows_grad, 20)
    g.close()
    f_tf = GUIDE_FWD_FLOP * rows_fwd / (f_ms * 1e-3) / 1e12
    g_tf = GUIDE_GRAD_FLOP * rows_grad / (g_ms * 1e-3) / 1e12
    return {"bound": "mfma", "unit": "TFLOP/s", "peak": MFMA_BF16_PEAK_TFLOPS,
            "forward": {"rows": rows_fwd, "ms": f_ms, "achieved": f_tf, "frac": f_tf / MFMA_BF16_PEAK_TFLOPS,
                        "flop_per_row": GUIDE_FWD_FLOP},
            "gradient": {"rows": rows_grad, "ms": g_ms, "achieved": g_tf, "frac": g_tf / MFMA_BF16_PEAK_TFLOPS,
                         "flop_per_row": GUIDE_GRAD_FLOP}}


def copy_probe(nbytes: int) -> None:
    """One device-to-device hipMemcpy of nbytes through the HIP runtime the
    integrator itself uses (the runtime's copy-buffer kernel reads and writes
    nbytes each): the known-size point that scales hardware byte counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)):
            raise RuntimeError("hipMalloc of %d bytes failed" % nbytes)
    try:
        if hip.hipMemset(src, 1, C.c_size_t(nbytes)) or hip.hipDeviceSynchronize():
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) or hip.hipDeviceSynchronize():  # 3: device to device
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    log("copy probe: %d bytes read, %d bytes written" % (nbytes, nbytes))


def kernel_source_hash() -> str:
    """SHA-256 over the device sources (lambda_amd/csrc): a traffic record is
    used only with the sources it was measured on."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "lambda_amd", "csrc")
    for fn in sorted(os.listdir(d)):
        with open(os.path.join(d, fn), "rb") as fh:
            h.update(fn.encode() + b"\0" + fh.read())
    return h.hexdigest()


def traffic_summary(fetch_csv: str, write_csv: str, kernel: str, copy_bytes: int, out: str) -> dict:
    """Writes the traffic record of `kernel` from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch) of the same command, each
    scaled by what the counter reads for the copy probe's known bytes."""
    import csv

    def read(path, counter):
        vals, copy = [], []
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            v = float(r["Counter_Value"]) * 1024.0
            name = r["Kernel_Name"]
            if kernel in name:
                vals.append(v)
            elif "copyBuffer" in name:
                copy.append(v)
        if not vals or not copy:
            raise SystemExit("%s: no %s rows for %s and the copy probe" % (path, counter, kernel))
        return vals, max(copy)

    f, fc = read(fetch_csv, "FETCH_SIZE")
    w, wc = read(write_csv, "WRITE_SIZE")
    fs, ws = copy_bytes / fc, copy_bytes / wc
    rec = {"kernel": kernel, "workload": "instanced_mesh_1080p",
           "bytes_per_launch": (sum(f) / len(f)) * fs + (sum(w) / len(w)) * ws,
           "fetch_bytes_per_launch": (sum(f) / len(f)) * fs, "write_bytes_per_launch": (sum(w) / len(w)) * ws,
           "fetch_scale": fs, "write_scale": ws, "dispatches_averaged": len(f),
           "calibration": "one %d-byte device-to-device hipMemcpy (__amd_rocclr_copyBuffer) in the same process: "
                          "FETCH_SIZE %.1f KiB, WRITE_SIZE %.1f KiB" % (copy_bytes, fc / 1024, wc / 1024),
           "source_sha256": kernel_source_hash()}
    with open(out, "w") as fh:
        json.dump(rec, fh, indent=1)
    return rec


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="instanced_mesh", choices=["instanced_mesh", "closed_box", "thinfilm_grid"])
    ap.add_argument("--rng", default="reference", choices=["reference", "throughput"])
    ap.add_argument("--path-slots", type=int, default=0)
    ap.add_argument("--cost-order", type=int, default=0,
                    help="reference streams: time the first N samples of every pixel, then run the slowest tiles "
                         "first (0 automatic: N = 4; -1 off)")
    ap.add_argument("--graph-passes", type=int, default=0, help="capture the bounce loop in a hipGraph of this many passes")
    ap.add_argument("--occupancy", type=int, default=0,
                    help="with --no-autotune: compiled waves per SIMD of the persistent kernel (2, 3 or 4; 0: default)")
    ap.add_argument("--tile", type=int, default=0, help="with --no-autotune: tile side of the cost order (0: default)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the library's default schedule instead of timing short renders first")
    ap.add_argument("--tune-cache", default=os.environ.get("LAMBDA_AUTOTUNE_CACHE", ""),
                    help="autotuner cache file (default: the tuner's own)")
    ap.add_argument("--tail-paths", type=int, default=0,
                    help="live paths below which the render finishes in one persistent launch (0 automatic, -1 never)")
    ap.add_argument("--scene-dir", default=os.environ.get("LAMBDA_SCENE_DIR", "/tmp/lambda_bench"))
    ap.add_argument("--cpu-seconds", type=float, default=25.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-guided", action="store_true", help="skip the guided-vs-plain leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the batched-serving leg")
    ap.add_argument("--serve-jobs", type=int, default=32)
    ap.add_argument("--serve-gap-ms", type=float, default=50.0)
    ap.add_argument("--serve-budget", type=int, default=128 << 20,
                    help="HBM bytes of the jobs in flight (0 = no limit)")
    ap.add_argument("--serve-launches", type=int, default=2, help="persistent launches in flight")
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--guided-res", default="320x180")
    ap.add_argument("--guided-ref-spp", type=int, default=65536)
    ap.add_argument("--guided-step", type=int, default=64)
    ap.add_argument("--guided-max-spp", type=int, default=65536)
    ap.add_argument("--guided-seed", type=int, default=1)
    ap.add_argument("--guided-runs", type=int, default=3,
                    help="guided runs (guide seeds guided-seed, +1, ...); the median time is reported")
    ap.add_argument("--guided-warmup", type=int, default=0, help="guided steps that train without adding samples")
    ap.add_argument("--guided-train-steps", type=int, default=48,
                    help="first kept guided steps that also train (their samples are unbiased and count)")
    ap.add_argument("--guided-epochs", type=int, default=2, help="Adam epochs over each step's samples")
    ap.add_argument("--guided-batch", type=int, default=16384)
    ap.add_argument("--guided-bounces", type=int, default=0,
                    help="vertices guided per path (bounce index below this; 0 = the library default, 8)")
    ap.add_argument("--guided-ref-save", default="", help="save the guided leg's reference image (.npy)")
    ap.add_argument("--guided-ref-load", default="", help="use a reference image saved by --guided-ref-save")
    ap.add_argument("--guided-ref-only", action="store_true",
                    help="with --guided-ref-save: render and save the reference, then stop")
    ap.add_argument("--guided-only", action="store_true",
                    help="run only the guided-vs-plain leg and print its record (no paths/s line)")
    ap.add_argument("--guided-table-cells", type=int, default=48,
                    help="guide from a table of the network's mixtures with this many cells along the scene's "
                         "longest side (0: the network inside the persistent kernel)")
    ap.add_argument("--guided-mix", type=float, default=0.5,
                    help="probability of sampling the guide at a guided vertex (the cosine lobe takes the rest)")
    ap.add_argument("--throughput-steps", type=int, default=2,
                    help="also time this many renders with throughput streams (0 to skip)")
    ap.add_argument("--stats-pixels", type=int, default=20000, help="pixels sampled to count traversal work")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES (a plain copy kernel of known traffic, "
                         "the calibration point for hardware byte counters)")
    ap.add_argument("--traffic-summary", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_BYTES"),
                    help="write the k_persist traffic record from two rocprofv3 --pmc runs and exit")
    args = ap.parse_args()
    if args.traffic_summary:
        fc, wc, nb = args.traffic_summary
        rec = traffic_summary(fc, wc, "k_persist", int(nb),
                              os.path.join(ROOT, "profiles", "persist_traffic_instanced_mesh_1080p.json"))
        print(json.dumps(rec))
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the renderer has no multi-GPU path (frames shard by pixel bands; "
              "see the design notes). Run with --gpus 1.", file=sys.stderr)
        return 2

    import numpy as np
    import lambda_amd
    import oracle

    name, cus = lambda_amd.device_info()
    log("device %s, %d CUs" % (name, cus))
    if args.guided_only:
        rec = guided_leg(args, args.scene_dir)
        rec.update({"device": name, "metric": "wall time to relMSE 1e-3, guided vs plain"})
        print(json.dumps(rec), flush=True)
        return 0
    if args.copy_probe:
        copy_probe(args.copy_probe)
    t0 = time.perf_counter()
    path, cfg = make_workload(args.workload, args.scene_dir)
    t_gen = time.perf_counter() - t0
    log("scene files ready (%.1f s)" % t_gen)
    t0 = time.perf_counter()
    cwd = os.getcwd()
    os.chdir(os.path.dirname(path))
    try:
        scene = lambda_amd.Scene.load(os.path.basename(path))
    finally:
        os.chdir(cwd)
    t_parse = time.perf_counter() - t0
    arrs_nodes = scene.view.num_nodes
    log("parsed + BVH built in %.1f s (%d nodes, %d objects, %d instances)" % (
        t_parse, arrs_nodes, scene.view.num_objects, scene.view.num_instances))
    t0 = time.perf_counter()
    dev = lambda_amd.DeviceScene(scene)
    t_upload = time.perf_counter() - t0
    mode = lambda_amd.LAMBDA_RNG_REFERENCE if args.rng == "reference" else lambda_amd.LAMBDA_RNG_THROUGHPUT
    p = scene.params
    tuned = None
    sched = {"path_slots": args.path_slots, "tail_paths": args.tail_paths, "cost_order": args.cost_order,
             "graph_passes": args.graph_passes}
    if args.occupancy:
        sched["occupancy"] = args.occupancy
    if args.tile:
        sched["tile"] = args.tile
    if not args.no_autotune:
        from lambda_amd import autotune
        t0 = time.perf_counter()
        kw = {"cache": args.tune_cache} if args.tune_cache else {}
        tuned = autotune.tune(dev, trial_spp=16, rng_mode=mode, **kw)
        tuned["seconds"] = time.perf_counter() - t0
        log("autotune (%.1f s): %s" % (tuned["seconds"], tuned["best"]))
        sched = dict(tuned["best"])
    paths_per_step = p.width * p.height * p.samples

    # the output buffer is allocated once, as main.cpp allocates h_output
    img = np.zeros((p.height, p.width, 3), np.float32)
    for i in range(args.warmup):
        t0 = time.perf_counter()
        dev.render(rng_mode=mode, out=img, **sched)
        log("warmup %d: %.3f s" % (i, time.perf_counter() - t0))

    t0 = time.perf_counter()
    for i in range(args.steps):
        dev.render(rng_mode=mode, out=img, **sched)
    elapsed = time.perf_counter() - t0
    value = paths_per_step * args.steps / elapsed
    log("timed %d steps: %.3f s -> %.4g paths/s" % (args.steps, elapsed, value))
    finite = bool(np.all(np.isfinite(img)))

    # Kernel timing from HIP events on the render stream.
    _, st = dev.render_stats(rng_mode=mode, **sched)
    log("kernel ms: " + ", ".join("%s %.1f" % (k, st[k + "_ms"]) for k in ("trace", "shade", "finalize", "persist")))

    # The same frame with independent streams per (pixel, sample): same
    # estimator, different random numbers, no per-pixel sample chains.
    thr = None
    if mode == lambda_amd.LAMBDA_RNG_REFERENCE and args.throughput_steps > 0:
        thr_img = np.zeros_like(img)
        thr_sched = {}
        if not args.no_autotune:
            from lambda_amd import autotune
            kw = {"cache": args.tune_cache} if args.tune_cache else {}
            thr_sched = dict(autotune.tune(dev, trial_spp=16, rng_mode=lambda_amd.LAMBDA_RNG_THROUGHPUT, **kw)["best"])
        dev.render(rng_mode=lambda_amd.LAMBDA_RNG_THROUGHPUT, out=thr_img, **thr_sched)
        t0 = time.perf_counter()
        for _ in range(args.throughput_steps):
            dev.render(rng_mode=lambda_amd.LAMBDA_RNG_THROUGHPUT, out=thr_img, **thr_sched)
        t_thr = (time.perf_counter() - t0) / args.throughput_steps
        thr = {"value": paths_per_step / t_thr, "unit": "paths/s", "ms_per_step": t_thr * 1e3,
               "steps": args.throughput_steps, "schedule": thr_sched}
        log("throughput streams: %.4g paths/s" % thr["value"])

    # Algorithmic bytes per closest-hit ray, counted by the oracle on a
    # strided pixel sample (all samples of each sampled pixel).
    npix = p.width * p.height
    stride = max(1, npix // max(1, args.stats_pixels))
    t0 = time.perf_counter()
    ps = oracle.path_stats(scene, stride=stride)
    t_stats = time.perf_counter() - t0
    b_closest = oracle.ray_bytes(ps, "closest")
    b_shadow = oracle.ray_bytes(ps, "shadow")
    # The dominant kernel traces every ray of the frame: k_persist (the
    # persistent schedule, one launch) or k_trace (one launch per bounce);
    # each handles both ray kinds.
    kern = "persist" if st["persist_ms"] >= st["trace_ms"] else "trace"
    launches = max(1, st[kern + "_launches"])
    bytes_total = b_closest * st["closest_rays"] + b_shadow * st["shadow_rays"]
    bytes_per_launch = bytes_total / launches
    avg_launch_s = st[kern + "_ms"] / 1e3 / launches
    achieved = bytes_per_launch / avg_launch_s / 1e9
    log("traversal work from %d sampled pixels (%.1f s): %.0f B per closest ray, %.0f B per shadow ray" % (
        ps["paths"] // max(1, p.samples), t_stats, b_closest, b_shadow))

    # HBM bytes per launch from the copy-calibrated FETCH_SIZE + WRITE_SIZE
    # counters (profiles/), valid only for the kernel sources they were
    # measured on
    traffic, traffic_note = None, "no counter record for this kernel and workload"
    tpath = os.path.join(ROOT, "profiles", "%s_traffic_%s.json" % (kern, cfg["workload"]))
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tj = json.load(fh)
        if tj.get("source_sha256") == kernel_source_hash():
            traffic, traffic_note = tj.get("bytes_per_launch"), "counters of these kernel sources (%s)" % os.path.basename(tpath)
        else:
            traffic_note = "counter record is stale (kernel sources changed since it was measured)"

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        log("CPU baseline on %d threads ..." % threads)
        cpu = cpu_baseline(scene, args.cpu_seconds, threads)
        log("CPU baseline %.4g paths/s (%s)" % (cpu["value"], cpu["sample"]))

    guided = None if args.no_guided else guided_leg(args, args.scene_dir)
    mlp = None if args.no_guided else guide_kernels()
    if mlp:
        log("guide MLP: forward %.1f TFLOP/s, gradient %.1f TFLOP/s (BF16 MFMA)" % (
            mlp["forward"]["achieved"], mlp["gradient"]["achieved"]))
    serving = None
    if not args.no_serve:
        serving = serve_leg(args, dev, args.scene_dir)
        log("serving: %d jobs in %d launches, %.4g paths/s, latency p50 %.0f ms p99 %.0f ms" % (
            serving["jobs"], serving["batches"], serving["paths_per_s"], serving["latency_ms"]["p50"],
            serving["latency_ms"]["p99"]))

    if tuned:
        cfg["autotune"] = {"best": tuned["best"], "trial_spp": tuned["trial_spp"],
                           "seconds": round(tuned.get("seconds", 0.0), 2),
                           "table_ms": [[c, round(ms, 2)] for c, ms in tuned["table"]]}
    cfg.update({"rng": args.rng, "path_slots": st["path_slots"], "schedule": sched, "setup_s": {
        "generate": round(t_gen, 2), "parse_and_bvh": round(t_parse, 2), "upload": round(t_upload, 2)}})
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (generated scene, seeded)",
        "config": cfg,
        "roofline": {
            "bound": "hbm", "kernel": "k_" + kern, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": traffic_note,
            "bytes_per_closest_ray": b_closest, "bytes_per_shadow_ray": b_shadow, "launches": launches,
            "avg_launch_ms": avg_launch_s * 1e3, "algorithmic_bytes_per_launch": bytes_per_launch,
        },
        "cpu_baseline": cpu,
        "throughput_streams": thr,
        "guided_vs_plain": guided,
        "guide_mlp_roofline": mlp,
        "batched_serving": serving,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
        "kernels_ms": {k: st[k + "_ms"] for k in ("trace", "shade", "finalize", "persist")},
        "rays": {"closest": st["closest_rays"], "shadow": st["shadow_rays"], "iterations": st["iterations"]},
        "image_finite": finite,
        "device": name,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
