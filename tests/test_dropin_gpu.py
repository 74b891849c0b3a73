"""The drop-in exactly as Lambda's main calls its render (src/main.cpp:113-171):
the command-line renderer on a golden scene file, and lambda_hip_render
called directly on the parsed arrays, against the reference's own outputs
(its float image and the PPM its Writer made of it)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, GOLDEN_SCENES, ROOT, SCENES, golden

import lambda_amd

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "lambda_amd", "lib", "lambda")


def ppm_values(data: bytes):
    tok = data.split()
    assert tok[:4] == [b"P3"] + tok[1:3] + [b"255"]
    return int(tok[1]), int(tok[2]), np.array([int(t) for t in tok[4:]], np.int32)


def test_cli_renders_golden_box_like_the_reference(tmp_path):
    """`lambda box.lrd out.ppm`: progress lines up to 100%, and a PPM that
    matches the reference Writer's PPM of the reference's render wherever
    the GPU's libm rounding does not move a value across a quantisation
    step (dither identical: Random(0, 0))."""
    shutil.copy(os.path.join(SCENES, "box.lrd"), tmp_path / "box.lrd")
    r = subprocess.run([CLI, "box.lrd", "out.ppm"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-200:], r.stderr[-500:], flush=True)
    assert r.returncode == 0
    assert r.stdout.rstrip().endswith("100% complete")
    w, h, got = ppm_values((tmp_path / "out.ppm").read_bytes())
    w2, h2, want = ppm_values(open(os.path.join(GOLDEN, "box_render.ppm"), "rb").read())
    assert (w, h) == (w2, h2) == (64, 64)
    diff = np.abs(got - want)
    print("PPM values differing: %d of %d (max %d)" % (int((diff > 0).sum()), diff.size, int(diff.max())), flush=True)
    assert diff.max() <= 1 and (diff > 0).mean() <= 1e-3
    # default output name: the input's with .ppm
    r = subprocess.run([CLI, "box.lrd"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and (tmp_path / "box.ppm").read_bytes() == (tmp_path / "out.ppm").read_bytes()


@pytest.mark.parametrize("name", GOLDEN_SCENES)
def test_lambda_hip_render_matches_reference_render(name, scenes):
    """lambda_hip_render (upload, render, copy back, free) on the parsed
    arrays against the reference binary's float image of the same scene;
    the progress counter ends at W*H."""
    s = scenes[name]
    prog = np.zeros(1, np.int32)
    img = lambda_amd.render(s, progress=prog)
    ref = golden(name + "_render.npy")
    m = lambda_amd.finite_pixels(img, ref)
    a, b = img[m].astype(np.float64), ref[m].astype(np.float64)
    err = float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    print("%s: lambda_hip_render relRMSE vs reference render %.3g" % (name, err), flush=True)
    assert np.array_equal(m, lambda_amd.finite_pixels(ref))
    assert err < 1e-5
    assert prog[0] == s.params.width * s.params.height
