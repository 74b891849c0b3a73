"""The benchmark's contract pieces that need no GPU: one-GPU refusal, the
relMSE metric against closed forms, and the FLOP counts it reports."""
import os
import subprocess
import sys

import numpy as np

import lambda_amd
from conftest import ROOT


def test_more_than_one_gpu_is_refused_before_any_device_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_rel_mse_closed_forms():
    rng = np.random.default_rng(3)
    ref = rng.uniform(0.0, 2.0, (8, 9, 3)).astype(np.float32)
    assert lambda_amd.rel_mse(ref, ref) == 0.0
    # a constant offset d gives mean(d^2 / (R^2 + 0.01))
    d = 0.25
    want = float(np.mean(d * d / (ref.astype(np.float64) ** 2 + 0.01)))
    assert abs(lambda_amd.rel_mse(ref + np.float32(d), ref) - want) < 1e-6 * want
    # pixels non-finite in either image are left out of both
    bad = ref.copy()
    bad[2, 3, 1] = np.nan
    off = ref + np.float32(d)
    off[5, 5, 0] = np.inf
    m = lambda_amd.finite_pixels(bad, off)
    assert m.sum() == ref.shape[0] * ref.shape[1] - 2 and not m[2, 3] and not m[5, 5]
    r64 = ref.astype(np.float64)[m]
    want = float(np.mean(d * d / (r64 * r64 + 0.01)))
    assert abs(lambda_amd.rel_mse(off, bad) - want) < 1e-6 * want


def test_guide_flop_counts():
    sys.path.insert(0, ROOT)
    import bench
    # 96x64 + 3 x 64x64 + 64x40 multiply-adds per row, 2 FLOP each
    assert bench.GUIDE_FWD_FLOP == 41984
    # padded forward 45,056 + backward data (all layers but the first) and
    # weight products (all layers)
    assert bench.GUIDE_GRAD_FLOP == 45056 + 2 * (2 * (96 * 64 + 4 * 64 * 64) - 96 * 64) == 122880
