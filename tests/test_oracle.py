"""Pins the CPU oracle before it is used to judge the GPU path.

* PCG32 against pcg-c's published demo output (seed 42, stream 54).
* Images, hit records, occlusion, thin film, randomUnitVector against golden
  vectors emitted by the reference itself (tests/golden/make_golden.py).
* Thin film against closed forms that fix the transfer-matrix algebra:
  single interface (Fresnel), absentee half-wave layer, quarter-wave
  anti-reflection, and the Airy formula for one layer.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN_SCENES, ROOT, SCENES, golden, load_golden_scene

import oracle
import lambda_amd


def test_pcg32_published_vectors():
    # pcg32-demo (pcg-c) prints these for pcg32_srandom_r(&rng, 42u, 54u).
    want = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    assert oracle.pcg32(42, 54, 6).tolist() == want
    assert golden("rng.npz")["pcg_42_54"].tolist() == want


def test_pcg32_matches_reference_draws():
    g = golden("rng.npz")
    assert np.array_equal(oracle.pcg32(7, 123456789, 64), g["pcg_7_123456789"])


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_render_bit_identical_to_reference(name, scenes):
    """The oracle's float RGB buffer equals the reference binary's, bit for bit."""
    img = oracle.render(scenes[name])
    ref = golden(name + "_render.npy")
    assert img.shape == ref.shape
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_closest_hits_match_reference(name, scenes):
    g = golden(name + "_hits.npz")
    h = oracle.closest_hit(scenes[name], g["rays"])
    ref = g["hits"]
    assert np.array_equal(h["hit"], ref["hit"])
    m = ref["hit"] == 1
    for f in ("instance", "object", "front_facing"):
        assert np.array_equal(h[f][m], ref[f][m]), f
    for f in ("t", "point", "normal", "u", "v"):
        assert np.array_equal(h[f][m].view(np.uint32), ref[f][m].view(np.uint32)), f


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_occlusion_matches_reference(name, scenes):
    g = golden(name + "_hits.npz")
    occ = oracle.occluded(scenes[name], g["shadow_rays"], g["light_instance"], g["light_object"])
    assert np.array_equal(occ, g["occluded"])


def test_thinfilm_matches_reference(scenes):
    g = golden("thinfilm.npz")
    for m in (1, 2):
        ref_f32, fp64 = oracle.thinfilm(scenes["film"], m, g["queries"])
        assert np.array_equal(ref_f32.view(np.uint32), g["R%d" % m].view(np.uint32))
        # the all-double value differs from the reference's float only by its
        # final rounding of |r| to float
        assert np.max(np.abs(fp64 - g["R%d" % m])) < 1e-6


def _film_scene(n_layers_text: str, d_text: str):
    text = """Render 4 4 2 4 360 830
Material f thinfilm [ %s ] [ %s ]
Object o sphere f (0 0 0) 1
Instance o default (0 0 0) (0 0 0) (1 1 1)
Background equirectangular 0
Camera (0 0 -3) (-1 1 -2) (2 0 0) (0 -2 0)
""" % (n_layers_text, d_text)
    return lambda_amd.Scene.parse(text, "film.lrd")


def _fresnel(n0, n1, c0):
    s1 = n0 / n1 * np.sqrt(1 - c0 * c0)
    c1 = np.sqrt(1 - s1 * s1 + 0j)
    rs = (n0 * c0 - n1 * c1) / (n0 * c0 + n1 * c1)
    rp = (n1 * c0 - n0 * c1) / (n1 * c0 + n0 * c1)
    return 0.5 * (abs(rs) ** 2 + abs(rp) ** 2)


def _airy(n0, n1, n2, d, lam, c0):
    """Single layer on a substrate: r = (r01 + r12 e^{2i delta}) / (1 + r01 r12 e^{2i delta})."""
    s0 = np.sqrt(1 - c0 * c0)
    c1 = np.sqrt(1 - (n0 / n1 * s0) ** 2 + 0j)
    c2 = np.sqrt(1 - (n0 / n2 * s0) ** 2 + 0j)
    delta = 2 * np.pi / lam * n1 * d * c1
    out = 0
    for pol in ("s", "p"):
        if pol == "s":
            r01 = (n0 * c0 - n1 * c1) / (n0 * c0 + n1 * c1)
            r12 = (n1 * c1 - n2 * c2) / (n1 * c1 + n2 * c2)
        else:
            r01 = (n1 * c0 - n0 * c1) / (n1 * c0 + n0 * c1)
            r12 = (n2 * c1 - n1 * c2) / (n2 * c1 + n1 * c2)
        e = np.exp(2j * delta)
        r = (r01 + r12 * e) / (1 + r01 * r12 * e)
        out += 0.5 * abs(r) ** 2
    return out


def _queries(c, lam, front=1.0):
    c = np.atleast_1d(np.asarray(c, np.float32))
    lam = np.broadcast_to(np.asarray(lam, np.float32), c.shape)
    q = np.zeros((len(c), 7), np.float32)
    q[:, 0] = c
    q[:, 1] = lam
    q[:, 2] = front
    return q


def test_thinfilm_closed_form_fresnel_zero_thickness():
    s = _film_scene("1 1.33 1.5", "0")
    c = np.linspace(0.05, 1, 40)
    _, r = oracle.thinfilm(s, 1, _queries(c, 550))
    want = _fresnel(1.0, 1.5, c.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(r - want)) < 1e-6


def test_thinfilm_closed_form_absentee_and_quarter_wave():
    lam = 550.0
    n1 = 1.38
    d_half = np.float32(lam / (2 * n1))
    s = _film_scene("1 %r 1.5" % n1, "%r" % float(d_half))
    _, r = oracle.thinfilm(s, 1, _queries([1.0], lam))
    assert abs(r[0] - _fresnel(1.0, 1.5, 1.0)) < 1e-6   # half-wave layer is absent
    nq = float(np.float32(np.sqrt(1.5)))
    d_q = np.float32(lam / (4 * nq))
    s = _film_scene("1 %r 1.5" % nq, "%r" % float(d_q))
    _, r = oracle.thinfilm(s, 1, _queries([1.0], lam))
    assert r[0] < 1e-6                                      # quarter-wave AR coating


def test_thinfilm_closed_form_airy_single_layer():
    s = _film_scene("1 1.33 1", "400")
    rng = np.random.default_rng(3)
    c = rng.uniform(0.05, 1, 200).astype(np.float32)
    lam = rng.uniform(380, 780, 200).astype(np.float32)
    _, r = oracle.thinfilm(s, 1, _queries(c, lam))
    n1 = float(np.float32(1.33))
    want = _airy(1.0, n1, 1.0, 400.0, lam.astype(np.float64), c.astype(np.float64))
    # the reference forms 2*pi/lambda in float before going to double
    assert np.max(np.abs(r - want)) < 2e-6


def test_random_unit_vector_draw_order():
    """The reference's randomUnitVector puts its third draw in x
    (g++ evaluates the constructor arguments last to first)."""
    g = golden("rng.npz")["unitvec_3_9"]
    draws = oracle.pcg32(3, 9, 48).astype(np.float64)
    f = (draws / 2.0 ** 32).astype(np.float32)
    c = f * np.float32(2) - np.float32(1)
    trip = c.reshape(16, 3)[:, ::-1]
    l = np.sqrt((trip.astype(np.float32) ** 2).sum(1, dtype=np.float32)).astype(np.float32)
    assert np.array_equal((trip / l[:, None]).astype(np.float32), g)


def test_path_stats_counts_work(scenes):
    st = oracle.path_stats(scenes["box"])
    p = scenes["box"].params
    assert st["paths"] == p.width * p.height * p.samples
    assert st["closest_rays"] >= st["paths"]
    assert st["closest_nodes"] > st["closest_rays"]
    assert st["shadow_rays"] > 0 and st["shadow_instances"] >= st["shadow_rays"] // 2
    assert oracle.ray_bytes(st, "closest") > 128 and oracle.algorithmic_bytes(st) > 0
    # a strided sample sees the same per-ray work up to sampling noise
    sub = oracle.path_stats(scenes["box"], stride=7)
    assert abs(oracle.ray_bytes(sub, "closest") / oracle.ray_bytes(st, "closest") - 1) < 0.05


def _camera_rays(scene, hit=None, draws_after_hit=0):
    """Every camera ray of a render, in float64, from the reference's draw
    order (u jitter, v jitter, wavelength; render/renderer.h:69-82) on the
    PCG32 streams (seed 42, pixel).  A sample whose camera ray satisfies
    hit(o, d) draws draws_after_hit more numbers before the next sample
    (Appendix A of the reference's per-sample draw order); others none."""
    p = scene.params
    cam = np.array(scene.view.camera[:], np.float64).reshape(4, 3)
    n = p.samples
    out = []
    for py in range(p.height):
        for px in range(p.width):
            r = oracle.pcg32(42, py * p.width + px, (3 + draws_after_hit) * n).astype(np.float64) / 2.0 ** 32
            pos = 0
            for k in range(n):
                i, j = k // p.sqrt_samples, k % p.sqrt_samples
                u = (px + (i + r[pos]) / p.sqrt_samples) / p.width
                v = (py + (j + r[pos + 1]) / p.sqrt_samples) / p.height
                pos += 3
                d = cam[1] + cam[2] * u + cam[3] * v - cam[0]
                d /= np.linalg.norm(d)
                out.append(d)
                if hit is not None and hit(cam[0][None], d[None])[0]:
                    pos += draws_after_hit
    return cam[0], np.array(out)


def _hits_box(o, d, lo, hi):
    with np.errstate(divide="ignore"):
        t0, t1 = (lo - o) / d, (hi - o) / d
    tn = np.minimum(t0, t1).max(1)
    tf = np.maximum(t0, t1).min(1)
    return (tn <= tf) & (tf >= 0)


def _hits_sphere(o, d, c, r):
    oc = np.broadcast_to(o - c, d.shape)
    b = (d * oc).sum(1)
    disc = b * b - ((oc * oc).sum(1) - r * r)
    t = -b - np.sqrt(np.maximum(disc, 0))
    return (disc > 0) & (t > 1e-4)


SPHERE_SCENE = """Render 24 16 %d 4 380 780
Material m %s
Object o sphere m (0 0 0) 1
Instance o default (0 0 0) (0 0 0) (1 1 1)
Background equirectangular %s
Camera (0.3 0.2 -3) (-1.2 1.1 -2) (2.4 0 0) (0 -2.2 0)
"""


def test_path_stats_brute_force_emitter_sphere():
    """One emissive sphere, one instance: every path is one camera ray that
    tests the instance-level root; the rays that pass its box (the sphere's
    box, found here by a slab test in float64) enter the instance, test its
    object-level root and the sphere; nothing is traced after an emitter or
    a miss, and there are no occlusion rays."""
    s = lambda_amd.Scene.parse(SPHERE_SCENE % (3, "emissive 2", "0"), "e.lrd")
    st = oracle.path_stats(s)
    o, d = _camera_rays(s)
    a = int(_hits_box(o, d, np.full(3, -1.0), np.full(3, 1.0)).sum())
    paths = 24 * 16 * 4
    assert 0 < a < paths
    assert st["paths"] == paths and st["closest_rays"] == paths and st["shadow_rays"] == 0
    assert st["closest_nodes"] == paths + a
    assert st["closest_instances"] == a and st["closest_spheres"] == a
    assert st["closest_quads"] == 0 and st["closest_triangles"] == 0
    assert oracle.ray_bytes(st, "closest") == pytest.approx((32 * (paths + a) + 64 * a + 16 * a) / paths + 128)


def test_path_stats_brute_force_lambertian_sphere_under_sky():
    """One Lambertian sphere under a constant sky, depth 1: camera rays as
    above (replaying the draws a hit consumes); each camera ray that hits the sphere (float64 quadratic) scatters
    once more from its surface, and that ray, starting inside the box,
    tests both roots and the sphere and escapes; each occlusion ray toward
    the sky does the same (2 nodes, 1 instance, 1 sphere)."""
    s = lambda_amd.Scene.parse(SPHERE_SCENE % (1, "lambertian 0.5", "1"), "l.lrd")
    st = oracle.path_stats(s)
    # a hit draws 1 (light choice) + 3 (sky direction) + 2 (cosine scatter);
    # the escaping second ray draws nothing
    o, d = _camera_rays(s, lambda o, d: _hits_sphere(o, d, np.zeros(3), 1.0), 6)
    a = int(_hits_box(o, d, np.full(3, -1.0), np.full(3, 1.0)).sum())
    h = int(_hits_sphere(o, d, np.zeros(3), 1.0).sum())
    paths = 24 * 16 * 4
    assert 0 < h < a < paths
    assert st["closest_rays"] == paths + h
    assert st["closest_nodes"] == paths + a + 2 * h
    assert st["closest_instances"] == a + h and st["closest_spheres"] == a + h
    sr = st["shadow_rays"]
    assert 0.35 * h < sr < 0.65 * h       # a uniform direction faces the normal half the time
    assert st["shadow_nodes"] == 2 * sr and st["shadow_instances"] == sr and st["shadow_spheres"] == sr


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "ref_harness")),
                    reason="reference harness not built (reference sources absent)")
def test_parser_messages_match_reference_binary(tmp_path):
    from test_host_parity import ERRORS
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    for text, msg in ERRORS:
        f = tmp_path / "t.lrd"
        f.write_text(text)
        r = subprocess.run([harness, "scene", "t.lrd", str(tmp_path / "o.bin")], capture_output=True, text=True,
                           cwd=tmp_path)
        assert r.stderr.strip() == msg
