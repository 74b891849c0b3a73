"""Host front end vs the reference: parsed arrays, BVH layout, light list and
parser error messages (core/parser.h, scene/bvh.h)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_SCENES, ROOT, golden, load_golden_scene

import lambda_amd

FIELDS = [
    "spectra", "complex_spectra", "scalar_texture_type", "scalar_texture_params", "spectrum_texture_type",
    "spectrum_texture_index", "spectrum_texture_scale", "material_type", "material_params",
    "material_properties", "object_type", "object_material", "object_geometry", "object_shading",
    "instance_ints", "instance_matrices", "node_bounds", "node_ints", "light_instances", "light_objects",
    "light_powers", "camera",
]


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint8) if a.size else a


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_scene_arrays_bit_identical(name, scenes):
    """Every flat array (objects, instances, nodes after the SAH build and
    reordering, light powers) matches the reference's bytes."""
    mine = scenes[name].arrays()
    ref = golden(name + "_scene.npz")
    for f in FIELDS:
        a, b = np.asarray(mine[f]), np.asarray(ref[f])
        assert a.shape == b.shape, f
        assert np.array_equal(_bits(a), _bits(b)), "%s differs from the reference" % f
    assert np.float32(mine["total_light_power"]).tobytes() == np.float32(ref["total_light_power"]).tobytes()
    assert mine["background_texture"] == int(ref["background_texture"])
    p = scenes[name].params
    ri = ref["render_int"]
    assert [p.width, p.height, p.depth, p.samples, p.sqrt_samples] == list(ri)
    assert np.float32(p.lambda_min) == ref["render_float"][0] and np.float32(p.lambda_max) == ref["render_float"][1]


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_bvh_invariants(name, scenes):
    """Every object sits in exactly one leaf of its group's tree, children
    nest in their parents' boxes, leaf depth <= 20 (scene/bvh.h:221)."""
    a = scenes[name].arrays()
    lo, hi = a["node_bounds"][:, :3], a["node_bounds"][:, 3:]
    idx, cnt = a["node_ints"][:, 0], a["node_ints"][:, 1]
    seen = np.zeros(len(a["object_type"]), np.int32)
    roots = sorted(set(a["instance_ints"][:, 2].tolist()))

    def walk(root, limit, mark):
        stack = [(root, 1)]
        while stack:
            n, depth = stack.pop()
            assert depth <= 20
            if cnt[n] > 0:
                assert idx[n] >= 0 and idx[n] + cnt[n] <= limit
                if mark is not None:
                    mark[idx[n]:idx[n] + cnt[n]] += 1
                continue
            for c in (n + 1, idx[n]):
                assert np.all(lo[c] >= lo[n]) and np.all(hi[c] <= hi[n])
                stack.append((c, depth + 1))

    walk(0, len(a["instance_ints"]), None)
    for r in roots:
        walk(r, len(a["object_type"]), seen)
    groups = {tuple(x) for x in a["instance_ints"][:, :2].tolist()}
    covered = np.zeros_like(seen)
    for first, count in groups:
        covered[first:first + count] = 1
    assert np.array_equal(seen, covered)


BASE = """Render 4 4 2 4 380 780
Material m lambertian 0.5
Object o sphere m (0 0 0) 1
Instance o default (0 0 0) (0 0 0) (1 1 1)
Background equirectangular 0
Camera (0 0 -3) (-1 1 -2) (2 0 0) (0 -2 0)
"""

# (text, expected message) pairs; messages as the reference words them
# (core/parser.h:73-382).
ERRORS = [
    ("", "t.lrd: expected \"Render\""),
    ("Camera (0 0 0) (0 0 0) (1 0 0) (0 1 0)\n", "t.lrd:1: expected \"Render\", got \"Camera\""),
    ("Render 0 4 2 4 380 780\n", "t.lrd:1: 'width' must be positive, got 0"),
    ("Render 4 4 -1 4 380 780\n", "t.lrd:1: 'depth' must be non-negative, got -1"),
    ("Render 4 4 2 5 380 780\n", "t.lrd:1: 'samples' must be a perfect square, got 5"),
    ("Render 4 4 2 4 300 780\n", "t.lrd:1: 'lambdaMin' must be at least 360, got 300.000000"),
    ("Render 4 4 2 4 380 900\n", "t.lrd:1: 'lambdaMax' must be at most 830, got 900.000000"),
    ("Render 4 4 2 4 700 600\n", "t.lrd:1: 'lambdaMin' must be at most 'lambdaMax', got 700.000000 and 600.000000"),
    ("Render 4 4 2 4 380\n", "t.lrd:1: 'lambdaMax' is missing or invalid"),
    ("Render 4 4 2 4 380 780 extra\n", "t.lrd:1: unexpected token \"extra\""),
    ("Render 4 4 2 4 380 780\nFoo\n",
     "t.lrd:2: expected \"Background\", \"Camera\", \"Instance\", \"Material\", \"Render\", \"Object\", or \"Texture\", got \"Foo\""),
    ("Render 4 4 2 4 380 780\nMaterial m lambertian 1.5\n", "t.lrd:2: 'albedo' must be at most 1, got 1.500000"),
    ("Render 4 4 2 4 380 780\nMaterial m lambertian nope\n", "t.lrd:2: 'albedo' is not defined"),
    ("Render 4 4 2 4 380 780\nMaterial m plastic 1\n",
     "t.lrd:2: 'type' must be \"dielectric\", \"emissive\", \"lambertian\", \"mirror\", or \"thinfilm\""),
    ("Render 4 4 2 4 380 780\nMaterial m thinfilm [ 1 ] [ ]\n", "t.lrd:2: 'n' must have at least 2 entries, got 1"),
    ("Render 4 4 2 4 380 780\nMaterial m thinfilm [ 1 1.3 1 ] [ ]\n", "t.lrd:2: 'd' must have 1 entries, got 0"),
    ("Render 4 4 2 4 380 780\nMaterial m thinfilm [ 1 1.3 1 ] [ dd ]\n", "t.lrd:2: 'd' is not defined"),
    ("Render 4 4 2 4 380 780\nMaterial m thinfilm [ 1 1.3x 1 ] [ 5 ]\n",
     "t.lrd:2: ''n' is missing or invalid' is missing or invalid"),
    ("Render 4 4 2 4 380 780\nObject o sphere default (0 0 0) -1\n", "t.lrd:2: 'radius' must be positive, got -1.000000"),
    ("Render 4 4 2 4 380 780\nObject o quad default (0 0 0) (1 0 0) (2 0 0)\n",
     "t.lrd:2: 'horizontal' and 'vertical' must be non-parallel"),
    ("Render 4 4 2 4 380 780\nObject o sphere nope (0 0 0) 1\n", "t.lrd:2: 'material' is not defined"),
    ("Render 4 4 2 4 380 780\nObject o sphere default (0 0 0) 1\nObject p sphere default (0 0 0) 1\nObject o sphere default (0 0 0) 1\n",
     "t.lrd:4: 'name' is already defined"),
    ("Render 4 4 2 4 380 780\nInstance o default (0 0 0) (0 0 0) (1 1 1)\n", "t.lrd:2: 'object' is not defined"),
    ("Render 4 4 2 4 380 780\nMaterial l emissive 1\nObject o sphere l (0 0 0) 1\nInstance o default (0 0 0) (0 0 0) (1 2 1)\n",
     "t.lrd:4: 'scale' must be uniform for emissive objects"),
    ("Render 4 4 2 4 380 780\nObject o sphere default (0 0 0) 1\nInstance o default (0 0 0) (0 0 0) (1 0 1)\n",
     "t.lrd:3: 'scale' must be non-zero"),
    ("Render 4 4 2 4 380 780\nTexture t scalar perlin 2 1 1\n", "t.lrd:2: 'min' must be at most 'max', got 2.000000 and 1.000000"),
    ("Render 4 4 2 4 380 780\nTexture t spectrum scalar nope\n", "t.lrd:2: 'texture' is not defined"),
    ("Render 4 4 2 4 380 780\nTexture t spectrum constant ( 300 1 )\n", "t.lrd:2: 'spectrum' is missing or invalid"),
    ("Render 4 4 2 4 380 780\nBackground sphere 0\n", "t.lrd:2: 'type' must be \"equirectangular\""),
    ("Render 4 4 2 4 380 780\nObject o sphere default (0 0 0) 1\nInstance o default (0 0 0) (0 0 0) (1 1 1)\n",
     "t.lrd: expected \"Background\""),
    (BASE.replace("Camera (0 0 -3) (-1 1 -2) (2 0 0) (0 -2 0)\n", ""), "t.lrd: expected \"Camera\""),
]


@pytest.mark.parametrize("text,message", ERRORS)
def test_parser_errors(text, message):
    with pytest.raises(lambda_amd.LambdaError) as e:
        lambda_amd.Scene.parse(text, "t.lrd")
    assert str(e.value) == message


def test_parser_accepts_grammar_variants():
    """Comments, commas, spectra, complex literals, checker and scalar
    textures, mesh-free scene."""
    text = BASE + """# comment line
Texture chk spectrum checker ( 400 0.2, 700 0.8 ) 0.1 0.25  # trailing comment
Texture noise scalar worley 0 1 3
Material w lambertian chk
Material g dielectric 1 ( 400 1.6 700 1.5 )
Material f thinfilm [ 1 1.4+0.01i 1.5-i i ] [ noise 120 ]
Object q quad w (0 -1 0) (1 0 0) (0 0 1)
Instance q f (0 0 0) (0 0.5 0) (2 2 2)
"""
    s = lambda_amd.Scene.parse(text, "ok.lrd")
    a = s.arrays()
    assert a["material_type"].tolist() == [0, 0, 0, 2, 4]
    cs = a["complex_spectra"]
    assert cs.shape[0] == 4 and cs[1, 0, 1] == np.float64(np.float32(0.01)) and cs[2, 0, 1] == -1.0
    assert cs[3, 0, 0] == 0.0 and cs[3, 0, 1] == 1.0
    # the numeric thickness became a new constant scalar texture
    assert a["scalar_texture_type"][-1] == 0 and a["scalar_texture_params"][-1, 0] == 120


def test_spectrum_resampling_matches_linear_interpolation():
    s = lambda_amd.Scene.parse(BASE + "Texture t spectrum constant ( 400 0.2 500 0.6 )\n", "r.lrd")
    sp = s.arrays()["spectra"][-1]
    lam = np.arange(360, 831)
    want = np.interp(lam, [400, 500], [np.float32(0.2), np.float32(0.6)]).astype(np.float32)
    assert np.max(np.abs(sp - want)) <= 1e-7
    assert sp[0] == np.float32(0.2) and sp[-1] == np.float32(0.6)


def test_missing_file_error(tmp_path):
    with pytest.raises(lambda_amd.LambdaError) as e:
        lambda_amd.Scene.load(str(tmp_path / "none.lrd"))
    assert "failed to open file (no such file or directory)" in str(e.value)


def test_obj_loader_fan_and_negative_indices(tmp_path):
    obj = tmp_path / "quad.obj"
    obj.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nvn 0 0 1\n"
                   "f -4/1/1 -3/2/1 -2/3/1 -1/4/1\n")
    text = BASE.replace("Object o sphere m (0 0 0) 1", "Object o mesh m %s" % obj)
    a = lambda_amd.Scene.parse(text, "m.lrd").arrays()
    assert len(a["object_type"]) == 2 and set(a["object_type"].tolist()) == {2}
    geo = a["object_geometry"]
    assert np.allclose(sorted(geo[:, 6].tolist()), [0.0, 1.0])  # vertical.x of the two fan triangles


def test_write_ppm(tmp_path):
    img = np.linspace(0, 1, 2 * 3 * 3, dtype=np.float32).reshape(2, 3, 3)
    out = tmp_path / "o.ppm"
    lambda_amd.write_ppm(str(out), img)
    lines = out.read_text().split("\n")
    assert lines[:3] == ["P3", "3 2", "255"]
    vals = np.array([int(x) for ln in lines[3:] if ln for x in ln.split()])
    assert vals.min() >= 0 and vals.max() <= 255 and vals[-1] == 255


@pytest.mark.parametrize("case", ["box_render", "ppm_cases"])
def test_write_ppm_byte_identical_to_reference_writer(tmp_path, case):
    """lambda_write_ppm against the reference's Writer (core/writer.h:23-41):
    the golden box render and an image of edge values (negative, NaN,
    infinities, the sRGB knee and its neighbours, values above 1) give the
    same bytes as the PPM files the reference harness wrote."""
    img = np.load(os.path.join(GOLDEN, case + ".npy"))
    out = tmp_path / "o.ppm"
    lambda_amd.write_ppm(str(out), img)
    with open(os.path.join(GOLDEN, case + ".ppm"), "rb") as fh:
        want = fh.read()
    assert out.read_bytes() == want



def test_cli_argument_and_file_errors(tmp_path):
    """The command-line renderer rejects what src/main.cpp:27-80 rejects,
    with the same messages and exit status, before any device work."""
    cli = os.path.join(ROOT, "lambda_amd", "lib", "lambda")
    if not os.path.exists(cli):
        pytest.skip("CLI not built")

    def run(*args):
        r = subprocess.run([cli, *args], cwd=tmp_path, capture_output=True, text=True, timeout=60)
        return r.returncode, r.stderr.strip()

    assert run() == (1, cli + ": invalid number of arguments")
    assert run("a.lrd", "b.ppm", "c") == (1, cli + ": invalid number of arguments")
    assert run("scene.txt") == (1, "scene.txt: invalid file extension")
    assert run("scene.lrd", "out.png") == (1, "out.png: invalid file extension")
    assert run("none.lrd") == (1, "none.lrd: failed to open file (no such file or directory)")
    (tmp_path / "bad.lrd").write_text("Camera (0 0 0) (0 0 0) (1 0 0) (0 1 0)\n")
    assert run("bad.lrd") == (1, 'bad.lrd:1: expected "Render", got "Camera"')


@pytest.mark.slow
def test_full_scale_arrays_match_reference_hashes(tmp_path):
    """The 1080p instanced-mesh scene (1,310,724 objects, 64 instances of a
    level-8 icosphere, 1,043,859 BVH nodes): every flat array the host front
    end builds -- the SAH node array with its std::partition order
    (scene/bvh.h:61-246), the reordered objects and instances, the light
    list -- hashes to the bytes the reference's own build produced
    (tests/golden/instanced_mesh_1080p_arrays.json, written by
    make_golden.py --full-scale from the reference harness)."""
    import hashlib
    import json
    import time
    from lambda_amd import scenes as gen
    want = json.load(open(os.path.join(GOLDEN, "instanced_mesh_1080p_arrays.json")))
    lrd = gen.write_instanced_mesh(str(tmp_path))
    for fn, h in want["inputs_sha256"].items():
        data = (tmp_path / fn).read_bytes()
        if fn.endswith(".lrd"):
            data = data.replace((str(tmp_path) + "/").encode(), b"")
        assert hashlib.sha256(data).hexdigest() == h, fn
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        t0 = time.perf_counter()
        s = lambda_amd.Scene.load(os.path.basename(lrd))
        dt = time.perf_counter() - t0
    finally:
        os.chdir(cwd)
    a = s.arrays()
    print("host parse + BVH %.1f s (reference %.0f s)" % (dt, want["reference_parse_and_bvh_s"]), flush=True)
    for f, rec in want["arrays"].items():
        x = np.ascontiguousarray(a[f])
        assert list(x.shape) == rec["shape"] and str(x.dtype) == rec["dtype"], f
        assert hashlib.sha256(x.tobytes()).hexdigest() == rec["sha256"], "%s differs from the reference" % f
    assert np.float32(a["total_light_power"]) == np.float32(want["total_light_power"])
