"""The integrator's thin-film arithmetic (float-float admittance recursion,
lambda_amd/csrc/dscene.h) compiled for the host and checked against the FP64
transfer-matrix oracle: check (b) without a GPU."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden
from thinfilm_cases import STACKS, queries, stack_array, stack_scene, within_tolerance

import oracle


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    out = tmp_path_factory.mktemp("tf") / "thinfilm_host"
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-ffp-contract=off", "--offload-arch=gfx950",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "lambda_amd", "csrc"),
           os.path.join(ROOT, "tests", "host_math", "thinfilm_host.hip"), "-o", str(out)]
    if not os.path.exists(cmd[0]):
        pytest.skip("hipcc not available")
    subprocess.run(cmd, check=True, capture_output=True)
    return str(out)


def run_driver(driver, tmp_path, stack, q):
    sp, qp, op = tmp_path / "s.f32", tmp_path / "q.f32", tmp_path / "o.f32"
    stack.tofile(sp)
    np.ascontiguousarray(q[:, :3]).tofile(qp)
    subprocess.run([driver, str(sp), str(qp), str(op)], check=True)
    return np.fromfile(op, np.float32)


@pytest.mark.parametrize("k", range(len(STACKS)))
def test_stack_within_1e5_of_fp64(driver, tmp_path, k):
    ns, ds = STACKS[k]
    q = queries(3000, seed=k)
    _, fp64, amb = oracle.thinfilm(stack_scene(ns, ds), 1, q, with_flags=True)
    got = run_driver(driver, tmp_path, stack_array(ns, ds), q)
    err, n_amb = within_tolerance(got, fp64, amb)
    print("stack %d: max error %.3g, %d noise-branch queries skipped" % (k, err, n_amb))
    assert err < 1e-5
    assert n_amb < len(q) // 2


def test_golden_film_queries(driver, tmp_path):
    """Lossless soap film [1, 1.33, 1] at the golden queries, thickness fixed."""
    g = golden("thinfilm.npz")
    q = g["queries"]
    for d in (200.0, 800.0):
        _, fp64 = oracle.thinfilm(stack_scene([1, 1.33, 1], [d]), 1, q)
        got = run_driver(driver, tmp_path, stack_array([1, 1.33, 1], [d]), q)
        assert np.max(np.abs(got - fp64)) < 1e-6
