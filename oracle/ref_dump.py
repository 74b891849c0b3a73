"""Test infrastructure: readers for the reference harness's outputs.

oracle/_ref/ref_harness writes tagged binary records (16-byte name, u64
element size, u64 count, raw bytes) holding the reference's own in-memory
structs.  The struct layouts decoded here are those of the reference
(g++ 11.4, x86-64): Object 116 B (scene/object.h:157-164), Instance 144 B
(scene/instance.h:113-115), BVHNode 32 B (scene/bvh.h:54-56), Material 16 B
(scene/material.h:129-138), ScalarTexture 16 B (scene/texture.h:92-99),
SpectrumTexture 16 B (scene/texture.h:180-187), Camera 48 B
(render/camera.h:18).  Unused words of the reference's unions are
uninitialised there, so only the words a record's type uses are returned.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "ref_harness")

HIT_RECORD = np.dtype([
    ("hit", np.int32), ("instance", np.int32), ("object", np.int32), ("t", np.float32),
    ("point", np.float32, 3), ("normal", np.float32, 3), ("u", np.float32), ("v", np.float32),
    ("front_facing", np.int32),
])


def available() -> bool:
    return os.path.exists(HARNESS)


def read_records(path: str) -> dict:
    out = {}
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0
    while pos < len(data):
        name = data[pos:pos + 16].split(b"\0")[0].decode()
        elem = int(np.frombuffer(data, np.uint64, 1, pos + 16)[0])
        count = int(np.frombuffer(data, np.uint64, 1, pos + 24)[0])
        pos += 32
        out[name] = (elem, count, data[pos:pos + elem * count])
        pos += elem * count
    return out


def scene_arrays(path: str) -> dict:
    """Reference scene dump converted to the flat layout of lambda_scene_view."""
    r = read_records(path)

    def words(name, n_words, dtype):
        elem, count, raw = r[name]
        assert elem == n_words * 4, (name, elem)
        return np.frombuffer(raw, dtype).reshape(count, n_words) if count else np.zeros((0, n_words), dtype)

    out = {}
    ri = np.frombuffer(r["render_i"][2], np.int32)
    rf = np.frombuffer(r["render_f"][2], np.float32)
    out["render"] = dict(width=int(ri[0]), height=int(ri[1]), depth=int(ri[2]), samples=int(ri[3]),
                         sqrt_samples=int(ri[4]), lambda_min=float(rf[0]), lambda_max=float(rf[1]))
    out["camera"] = np.frombuffer(r["camera"][2], np.float32).copy()
    elem, count, raw = r["spectra"]
    out["spectra"] = np.frombuffer(raw, np.float32).reshape(count, 471).copy()
    elem, count, raw = r["cspectra"]
    out["complex_spectra"] = np.frombuffer(raw, np.float64).reshape(count, 471, 2).copy()
    out["background_texture"] = int(np.frombuffer(r["background"][2], np.int32)[1])

    ob_i = words("objects", 29, np.int32)
    ob_f = words("objects", 29, np.float32)
    n = len(ob_i)
    out["object_type"] = ob_i[:, 0].copy()
    out["object_material"] = ob_i[:, 1].copy()
    geom = np.zeros((n, 12), np.float32)
    shade = np.zeros((n, 15), np.float32)
    sph = out["object_type"] == 0
    geom[sph, :4] = ob_f[sph, 2:6]
    geom[~sph] = ob_f[~sph, 2:14]
    tri = out["object_type"] == 2
    shade[tri] = ob_f[tri, 14:29]
    out["object_geometry"] = geom
    out["object_shading"] = shade

    in_i = words("instances", 36, np.int32)
    in_f = words("instances", 36, np.float32)
    out["instance_ints"] = in_i[:, :4].copy()
    out["instance_matrices"] = in_f[:, 4:36].copy()

    nd_i = words("nodes", 8, np.int32)
    nd_f = words("nodes", 8, np.float32)
    out["node_bounds"] = nd_f[:, :6].copy()
    out["node_ints"] = nd_i[:, 6:8].copy()

    out["light_instances"] = np.frombuffer(r["light_inst"][2], np.int32).copy()
    out["light_objects"] = np.frombuffer(r["light_obj"][2], np.int32).copy()
    out["light_powers"] = np.frombuffer(r["light_pow"][2], np.float32).copy()
    out["total_light_power"] = np.frombuffer(r["light_total"][2], np.float32)[0]

    mt_i = words("materials", 4, np.int32)
    out["material_type"] = mt_i[:, 0].copy()
    mp = np.zeros((len(mt_i), 3), np.int32)
    for k, t in enumerate(mt_i[:, 0]):
        used = {0: 1, 1: 1, 2: 2, 3: 1, 4: 3}[int(t)]
        mp[k, :used] = mt_i[k, 1:1 + used]
    out["material_params"] = mp
    out["material_properties"] = np.frombuffer(r["mat_props"][2], np.int32).copy()

    st_i = words("scalar_tex", 4, np.int32)
    st_f = words("scalar_tex", 4, np.float32)
    out["scalar_texture_type"] = st_i[:, 0].copy()
    sp = np.zeros((len(st_i), 3), np.float32)
    for k, t in enumerate(st_i[:, 0]):
        used = 1 if t == 0 else 3
        sp[k, :used] = st_f[k, 1:1 + used]
    out["scalar_texture_params"] = sp

    xt_i = words("spectrum_tex", 4, np.int32)
    xt_f = words("spectrum_tex", 4, np.float32)
    out["spectrum_texture_type"] = xt_i[:, 0].copy()
    idx = np.zeros((len(xt_i), 2), np.int32)
    scale = np.zeros(len(xt_i), np.float32)
    for k, t in enumerate(xt_i[:, 0]):
        if t == 1:
            idx[k] = xt_i[k, 1:3]
            scale[k] = xt_f[k, 3]
        else:
            idx[k, 0] = xt_i[k, 1]
    out["spectrum_texture_index"] = idx
    out["spectrum_texture_scale"] = scale
    return out


def run(*args, timeout=3600, env=None) -> str:
    res = subprocess.run([HARNESS, *map(str, args)], capture_output=True, text=True, timeout=timeout, env=env)
    if res.returncode != 0:
        raise RuntimeError("ref_harness %s failed: %s" % (args[0], res.stderr.strip()))
    return res.stdout


def dump_scene(lrd: str, out: str) -> dict:
    run("scene", lrd, out)
    return scene_arrays(out)


def render(lrd: str, out: str, threads: int = 0):
    import json
    info = json.loads(run("render", lrd, out, threads).strip().splitlines()[-1])
    img = np.fromfile(out, np.float32)
    return img, info


def hits(lrd: str, rays: np.ndarray, occluders: np.ndarray | None, workdir: str) -> np.ndarray:
    rp = os.path.join(workdir, "rays.f32")
    np.ascontiguousarray(rays, np.float32).tofile(rp)
    op = "-"
    if occluders is not None:
        op = os.path.join(workdir, "occ.i32")
        np.ascontiguousarray(occluders, np.int32).tofile(op)
    out = os.path.join(workdir, "hits.bin")
    run("hits", lrd, rp, op, out)
    return np.fromfile(out, HIT_RECORD)


def thinfilm(lrd: str, material: int, queries: np.ndarray, workdir: str) -> np.ndarray:
    qp = os.path.join(workdir, "tf.f32")
    np.ascontiguousarray(queries, np.float32).tofile(qp)
    out = os.path.join(workdir, "tf_out.f32")
    run("thinfilm", lrd, material, qp, out)
    return np.fromfile(out, np.float32)


def pcg(seed: int, stream: int, n: int) -> np.ndarray:
    return np.array([int(x) for x in run("pcg", seed, stream, n).split()], np.uint32)


def unit_vectors(seed: int, stream: int, n: int) -> np.ndarray:
    rows = [[float.fromhex(t) for t in line.split()] for line in run("unitvec", seed, stream, n).splitlines()]
    return np.array(rows, np.float32)


def ppm(image: np.ndarray, workdir: str) -> bytes:
    """The bytes of the PPM the reference's Writer makes of a float [H, W, 3]
    image (core/writer.h:17-41)."""
    img = np.ascontiguousarray(image, np.float32)
    src = os.path.join(workdir, "ppm_in.f32")
    dst = os.path.join(workdir, "ppm_out.ppm")
    img.tofile(src)
    run("ppm", src, img.shape[1], img.shape[0], dst)
    with open(dst, "rb") as fh:
        return fh.read()
