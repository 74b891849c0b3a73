"""Parity on the full 1080p instanced-mesh scene: a level-8 perlin-displaced
icosphere (1,310,720 triangles) instanced 64 times with rotation and scale,
Lambertian / mirror / dielectric / thin-film instances, 4 quad emitters and
a constant environment, 1920x1080.  Its deepest traversal stack is longer
than the part of the stack each lane keeps in LDS, so traversals here spill
to global memory.  Checked against the oracle: closest-hit and occlusion
queries bit for bit, and a band of rows rendered with the reference's
streams (scene/scene.h:51-124, render/renderer.h:64-92)."""
import os
import time

import numpy as np
import pytest

import lambda_amd
import oracle
from lambda_amd import scenes as gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("full"))
    t0 = time.perf_counter()
    lrd = gen.write_instanced_mesh(d)
    cwd = os.getcwd()
    os.chdir(d)
    try:
        s = lambda_amd.Scene.load(os.path.basename(lrd))
    finally:
        os.chdir(cwd)
    dev = lambda_amd.DeviceScene(s)
    print("full-scale scene: %d objects, %d instances, %d nodes, ready in %.1f s" % (
        s.view.num_objects, s.view.num_instances, s.view.num_nodes, time.perf_counter() - t0), flush=True)
    yield s, dev
    dev.close()


def test_stack_spills_past_lds(full):
    s, dev = full
    info = dev.info()
    print("scene info:", info, flush=True)
    assert info["num_objects"] == 1310720 + 4 and info["num_instances"] == 65
    assert info["stack_need"] > info["stack_lds"]


def test_closest_hits_and_occlusion_bit_exact(full):
    s, dev = full
    lo, hi = np.array(dev.info()["bounds"][:3]), np.array(dev.info()["bounds"][3:])
    rng = np.random.default_rng(8)
    n = 200000
    p = lo + (hi - lo) * rng.random((n, 3))
    q = lo + (hi - lo) * rng.random((n, 3))
    d = (q - p) / np.linalg.norm(q - p, axis=1, keepdims=True)
    rays = np.concatenate([p, d], 1).astype(np.float32)
    t0 = time.perf_counter()
    want = oracle.closest_hit(s, rays)
    t1 = time.perf_counter()
    got = dev.closest_hit(rays)
    print("oracle %.1f s; %d of %d rays hit" % (t1 - t0, int(want["hit"].sum()), n), flush=True)
    assert want["hit"].sum() > n // 10
    assert np.array_equal(got["hit"], want["hit"])
    m = want["hit"] == 1
    assert np.array_equal(got["instance"][m], want["instance"][m])
    assert np.array_equal(got["object"][m], want["object"][m])
    assert np.array_equal(got["t"][m].view(np.uint32), want["t"][m].view(np.uint32))
    # occlusion rays from the hit points toward points on the four emitters
    a = s.arrays()
    li_all, lo_all = a["light_instances"][:-1], a["light_objects"][:-1]
    pts = want["point"][m][:50000].astype(np.float64)
    k = rng.integers(0, len(li_all), len(pts))
    g = a["object_geometry"][lo_all[k]].astype(np.float64)
    uv = rng.random((len(pts), 2))
    tgt = g[:, :3] + uv[:, :1] * g[:, 3:6] + uv[:, 1:] * g[:, 6:9]
    dd = (tgt - pts) / np.linalg.norm(tgt - pts, axis=1, keepdims=True)
    srays = np.concatenate([pts, dd], 1).astype(np.float32)
    li, lo_ = li_all[k].astype(np.int32), lo_all[k].astype(np.int32)
    occ_want = oracle.occluded(s, srays, li, lo_)
    occ_got = dev.occluded(srays, li, lo_)
    print("occluded %d of %d" % (int(occ_want.sum()), len(occ_want)), flush=True)
    assert 0 < occ_want.sum() < len(occ_want)
    assert np.array_equal(occ_got, occ_want)


@pytest.mark.parametrize("spp", [1, 4])
def test_row_band_render_matches_oracle(full, spp):
    """The whole frame on the GPU with the reference's streams; 12 rows
    through the middle of the instance grid by the oracle."""
    s, dev = full
    root = int(round(spp ** 0.5))
    p = s.render_params(samples=spp, sqrt_samples=root)
    r0, r1 = 560, 572
    t0 = time.perf_counter()
    want = oracle.render(s, p, rows=(r0, r1))[r0:r1]
    t1 = time.perf_counter()
    got = dev.render(samples=spp, sqrt_samples=root)[r0:r1]
    m = lambda_amd.finite_pixels(got, want)
    assert m.mean() > 0.99
    a, b = got[m].astype(np.float64), want[m].astype(np.float64)
    err = float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    print("%d spp rows %d-%d: oracle %.1f s, relRMSE %.3g, mean %.4g" % (spp, r0, r1, t1 - t0, err, b.mean()),
          flush=True)
    assert b.mean() > 0
    assert err < 1e-4


def test_schedules_bit_identical_at_full_scale(full):
    """With traversals that spill past the LDS ring: the default persistent
    kernel, its 3- and 2-wave variants, the voting schedule and the
    two-launch cost order render the frame (4 spp, reference streams) bit
    for bit alike."""
    s, dev = full
    kw = dict(samples=4, sqrt_samples=2)
    a = dev.render(cost_order=-1, **kw)
    for extra in (dict(occupancy=3, cost_order=-1), dict(occupancy=2, cost_order=-1), dict(vote=1, cost_order=-1),
                  dict(cost_order=2), dict(vote=1, occupancy=3, cost_order=2)):
        b = dev.render(**extra, **kw)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), extra


def test_gpu_bvh_on_full_scale_scene(full):
    """The BVHs of the 1.31M-triangle object group and the 65 instances
    built on the GPU: hit distances bit-identical to the host-built
    (reference) BVH on 200k random rays; the build time is printed."""
    s, dev = full
    g = lambda_amd.DeviceScene(s, gpu_bvh=True)
    info = dev.info()
    lo, hi = np.array(info["bounds"][:3]), np.array(info["bounds"][3:])
    rng = np.random.default_rng(9)
    n = 200000
    p = lo + (hi - lo) * rng.random((n, 3))
    q = lo + (hi - lo) * rng.random((n, 3))
    d = (q - p) / np.linalg.norm(q - p, axis=1, keepdims=True)
    rays = np.concatenate([p, d], 1).astype(np.float32)
    want = dev.closest_hit(rays)
    got = g.closest_hit(rays)
    gi = g.info()
    print("GPU BVH: %d nodes in %.1f ms (host-built: %d nodes), stack %d" % (
        gi["num_nodes"], g.bvh_build_ms, info["num_nodes"], gi["stack_need"]), flush=True)
    assert np.array_equal(got["hit"], want["hit"])
    m = want["hit"] == 1
    assert np.array_equal(got["t"][m].view(np.uint32), want["t"][m].view(np.uint32))
    same = g.object_order[got["object"][m]] == want["object"][m]
    assert same.mean() > 0.999
    g.close()
