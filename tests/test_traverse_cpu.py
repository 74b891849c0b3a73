"""The integrator's BVH traversal (lambda_amd/csrc/dscene.h over the 4-wide
BVH of lambda_amd/csrc/scene_prep.h) compiled for the host and checked
against the reference's golden hits and the oracle: check (a) without a GPU.

Built twice: with the 16-entry LDS ring the kernels use, and with a 2-entry
ring so that deep traversals go through the global spill area."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN_SCENES, ROOT, SCENES, golden

import lambda_amd
import oracle

LIB = os.path.join(ROOT, "lambda_amd", "lib")


def build_driver(out, ring):
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-ffp-contract=off", "--offload-arch=gfx950", "-DRING=%d" % ring,
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "lambda_amd", "csrc"),
           os.path.join(ROOT, "tests", "host_math", "traverse_host.hip"),
           "-L" + LIB, "-llambda_host", "-Wl,-rpath," + LIB, "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)


@pytest.fixture(scope="module", params=[16, 2], ids=["ring16", "ring2"])
def driver(request, tmp_path_factory):
    if not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not available")
    if not os.path.exists(os.path.join(LIB, "liblambda_host.so")):
        pytest.skip("host library not built")
    out = tmp_path_factory.mktemp("trav") / ("traverse_host_%d" % request.param)
    build_driver(out, request.param)
    return str(out)


def closest(driver, lrd, rays, tmp_path, cwd):
    rp, op = tmp_path / "r.f32", tmp_path / "o.i32"
    np.ascontiguousarray(rays, np.float32).tofile(rp)
    p = subprocess.run([driver, lrd, "closest", str(rp), str(op)], check=True, capture_output=True, cwd=cwd,
                       text=True)
    out = np.fromfile(op, np.int32).reshape(-1, 4)
    return out, int(p.stdout.split()[0])


def occluded(driver, lrd, rays, li, lo, tmp_path, cwd):
    rp, lp, op = tmp_path / "r.f32", tmp_path / "l.i32", tmp_path / "o.i32"
    np.ascontiguousarray(rays, np.float32).tofile(rp)
    np.ascontiguousarray(np.stack([li, lo], 1), np.int32).tofile(lp)
    subprocess.run([driver, lrd, "occluded", str(rp), str(lp), str(op)], check=True, capture_output=True, cwd=cwd)
    return np.fromfile(op, np.int32)


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_closest_hits_match_reference(driver, name, tmp_path):
    g = golden(name + "_hits.npz")
    out, need = closest(driver, name + ".lrd", g["rays"], tmp_path, SCENES)
    ref = g["hits"]
    assert np.array_equal(out[:, 0], ref["hit"])
    m = ref["hit"] == 1
    assert np.array_equal(out[m, 1], ref["instance"][m])
    assert np.array_equal(out[m, 2], ref["object"][m])
    assert np.array_equal(out[m, 3], ref["t"][m].view(np.int32))
    assert need >= 0


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_occlusion_matches_reference(driver, name, tmp_path):
    g = golden(name + "_hits.npz")
    out = occluded(driver, name + ".lrd", g["shadow_rays"], g["light_instance"], g["light_object"], tmp_path, SCENES)
    assert np.array_equal(out, g["occluded"])


def test_random_rays_on_instanced_mesh(driver, tmp_path):
    """100k rays between random points of a 64-instance mesh scene (rotated,
    scaled, translated instances): IDs and t bit-exact against the oracle."""
    from lambda_amd import scenes as gen
    lrd = gen.write_instanced_mesh(str(tmp_path), 64, 36, 10, 1, level=4, grid=8, name="m")
    s = lambda_amd.Scene.load(lrd)
    a = s.arrays()
    lo, hi = a["node_bounds"][0, :3], a["node_bounds"][0, 3:]
    rng = np.random.default_rng(1)
    n = 100000
    p = lo + (hi - lo) * rng.random((n, 3))
    q = lo + (hi - lo) * rng.random((n, 3))
    d = (q - p) / np.linalg.norm(q - p, axis=1, keepdims=True)
    rays = np.concatenate([p, d], 1).astype(np.float32)
    want = oracle.closest_hit(s, rays)
    got, need = closest(driver, lrd, rays, tmp_path, os.path.dirname(lrd))
    print("stack bound %d" % need)
    assert want["hit"].sum() > n // 4
    assert np.array_equal(got[:, 0], want["hit"])
    m = want["hit"] == 1
    assert np.array_equal(got[m, 1], want["instance"][m])
    assert np.array_equal(got[m, 2], want["object"][m])
    assert np.array_equal(got[m, 3], want["t"][m].view(np.int32))
    # occlusion toward random object points of the light list
    li = np.full(n, -1, np.int32)
    lo_ = np.full(n, -1, np.int32)
    occ = occluded(driver, lrd, rays, li, lo_, tmp_path, os.path.dirname(lrd))
    want_occ = oracle.occluded(s, rays, li, lo_)
    assert np.array_equal(occ, want_occ)
