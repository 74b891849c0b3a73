"""GPU-side BVH build (lambda_hip_scene_create_lbvh): linear BVHs built on
the device renumber objects and instances but must find the same surfaces
at the same distances as the reference's host-built BVH
(scene/bvh.h:61-246, scene/scene.h:51-124): closest hits and occlusion on
the golden ray sets, random rays on instanced meshes, and reference-stream
renders."""
import numpy as np
import pytest

from conftest import GOLDEN_SCENES, golden

import lambda_amd

pytestmark = pytest.mark.gpu


def compare_hits(dev_l, got, want):
    """Hit flags and distances equal; objects and instances equal through
    the renumbering (exact ties between primitives aside)."""
    assert np.array_equal(got["hit"], want["hit"])
    m = want["hit"] == 1
    assert np.array_equal(got["t"][m].view(np.uint32), want["t"][m].view(np.uint32))
    obj = dev_l.object_order[got["object"][m]]
    inst = dev_l.instance_order[got["instance"][m]]
    same = (obj == want["object"][m]) & (inst == want["instance"][m])
    return float(same.mean()) if m.any() else 1.0


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_gpu_bvh_hits_and_occlusion_match_reference(name, scenes):
    s = scenes[name]
    dev = lambda_amd.DeviceScene(s, gpu_bvh=True)
    g = golden(name + "_hits.npz")
    agree = compare_hits(dev, dev.closest_hit(g["rays"]), g["hits"])
    inv_obj = np.argsort(dev.object_order)
    inv_inst = np.argsort(dev.instance_order)
    li, lo = g["light_instance"], g["light_object"]
    li2 = np.where(li >= 0, inv_inst[np.maximum(li, 0)], -1).astype(np.int32)
    lo2 = np.where(lo >= 0, inv_obj[np.maximum(lo, 0)], -1).astype(np.int32)
    occ = dev.occluded(g["shadow_rays"], li2, lo2)
    info = dev.info()
    print("%s: GPU BVH %d nodes (%.2f ms), stack %d; same object %.4f" % (
        name, info["num_nodes"], dev.bvh_build_ms, info["stack_need"], agree), flush=True)
    dev.close()
    assert np.array_equal(occ, g["occluded"])
    assert agree > 0.995


@pytest.mark.parametrize("name", ["box", "mesh", "room"])
def test_gpu_bvh_render_matches_host_bvh_render(name, scenes):
    """Reference streams: the same image up to exact distance ties."""
    s = scenes[name]
    a = lambda_amd.DeviceScene(s)
    b = lambda_amd.DeviceScene(s, gpu_bvh=True)
    ia, ib = a.render(), b.render()
    a.close()
    b.close()
    m = lambda_amd.finite_pixels(ia, ib)
    err = float(np.sqrt(((ia[m] - ib[m]).astype(np.float64) ** 2).sum() / (ia[m].astype(np.float64) ** 2).sum()))
    print("%s: relRMSE host BVH vs GPU BVH %.3g" % (name, err), flush=True)
    assert err < 1e-5


def test_gpu_bvh_random_rays_on_instanced_mesh(tmp_path):
    """200k rays through a 64-instance mesh scene (rotated, scaled,
    translated instances): distances bit-identical to the host-built BVH."""
    from lambda_amd import scenes as gen
    lrd = gen.write_instanced_mesh(str(tmp_path), 64, 36, 10, 1, level=5, grid=8, name="m")
    s = lambda_amd.Scene.load(lrd)
    a = lambda_amd.DeviceScene(s)
    b = lambda_amd.DeviceScene(s, gpu_bvh=True)
    info = a.info()
    lo, hi = np.array(info["bounds"][:3]), np.array(info["bounds"][3:])
    rng = np.random.default_rng(3)
    n = 200000
    p = lo + (hi - lo) * rng.random((n, 3))
    q = lo + (hi - lo) * rng.random((n, 3))
    d = (q - p) / np.linalg.norm(q - p, axis=1, keepdims=True)
    rays = np.concatenate([p, d], 1).astype(np.float32)
    want = a.closest_hit(rays)
    agree = compare_hits(b, b.closest_hit(rays), want)
    print("mesh level 5 x 64: %d objects, GPU BVH build %.1f ms, stack %d (host-built %d); same object %.5f" % (
        s.view.num_objects, b.bvh_build_ms, b.info()["stack_need"], info["stack_need"], agree), flush=True)
    a.close()
    b.close()
    assert want["hit"].sum() > n // 10
    assert agree > 0.999
