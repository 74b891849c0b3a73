"""The guiding network's MFMA kernels against the FP64 oracle (with the
GPU's bfloat16 rounding points emulated), and training on a known target."""
import numpy as np
import pytest

import lambda_amd
import guide_oracle as go

from test_guide_cpu import random_dirs, random_params, random_samples

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def guide():
    g = lambda_amd.Guide(seed=7)
    yield g
    g.close()


def test_initial_parameters_keep_padding_zero(guide):
    p = guide.params().astype(np.float64)
    w4 = p[go.w_offset(4):go.w_offset(4) + 64 * 64].reshape(64, 64)
    assert np.all(w4[:, go.N_OUT:] == 0)
    assert np.all(p[go.b_offset(0):go.b_offset(0) + 64] == 0)


def test_forward_matches_oracle(guide):
    rng = np.random.default_rng(10)
    p = random_params(rng).astype(np.float32)
    guide.set_params(p)
    f = rng.random((1000, 6)).astype(np.float32)   # not a multiple of 64 rows
    got = guide.forward(f).astype(np.float64)
    ref, _ = go.forward(p.astype(np.float64), f.astype(np.float64), bf=True)
    ref = ref[:, :go.N_OUT]
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("forward relative L2 error %.3g, max abs %.3g" % (err, np.abs(got - ref).max()), flush=True)
    assert err < 5e-3
    np.testing.assert_allclose(got, ref, rtol=0.05, atol=0.02 * np.abs(ref).max())


def test_gradient_matches_oracle(guide):
    rng = np.random.default_rng(11)
    p = random_params(rng).astype(np.float32)
    guide.set_params(p)
    s = random_samples(rng, 1000).astype(np.float32)
    g, loss = guide.gradient(s)
    gref, lref = go.gradient(p.astype(np.float64), s.astype(np.float64), bf=True)
    print("loss gpu %.6g oracle %.6g" % (loss, lref), flush=True)
    assert abs(loss - lref) <= 2e-3 * abs(lref)
    for l in range(go.LAYERS):
        for off, size in ((go.w_offset(l), go.layer_in(l) * 64), (go.b_offset(l), 64)):
            a, b = g[off:off + size].astype(np.float64), gref[off:off + size]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
            cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
            print("layer %d %s: relative L2 %.3g, cosine %.6f" % (l, "W" if off == go.w_offset(l) else "b", rel, cos),
                  flush=True)
            assert rel < 3e-2 and cos > 0.9995


def test_gradient_is_additive_over_samples(guide):
    """Blocks accumulate with atomics: the gradient of a set equals the sum
    of the gradients of its halves (to float rounding)."""
    rng = np.random.default_rng(12)
    guide.set_params(random_params(rng).astype(np.float32))
    s = random_samples(rng, 640).astype(np.float32)
    g, _ = guide.gradient(s)
    g1, _ = guide.gradient(s[:320])
    g2, _ = guide.gradient(s[320:])
    np.testing.assert_allclose(g, g1 + g2, rtol=1e-4, atol=1e-4 * np.abs(g).max())


def target_samples(rng, n, kappa=20.0):
    """Directions from a vMF whose mean depends on the position feature."""
    s = np.zeros((n, 10), np.float32)
    s[:, :6] = rng.random((n, 6))
    mu = s[:, :3] - 0.5
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    u = rng.random(n)
    w = 1 + np.log1p(u * np.expm1(-2 * kappa)) / kappa
    v = np.sqrt(np.maximum(0, 1 - w * w))
    phi = 2 * np.pi * rng.random(n)
    t = np.cross(mu, np.where(np.abs(mu[:, :1]) < 0.9, [[1, 0, 0]], [[0, 1, 0]]))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(mu, t)
    d = t * (v * np.cos(phi))[:, None] + b * (v * np.sin(phi))[:, None] + mu * w[:, None]
    s[:, 6:9] = d
    s[:, 9] = 1.0
    return s


def test_training_learns_a_position_dependent_distribution():
    """Adam on the GPU: the mean loss -log q falls from log(4 pi) (uniform)
    toward the target's entropy (about -0.16 for kappa = 20)."""
    g = lambda_amd.Guide(seed=3)
    rng = np.random.default_rng(13)
    train = target_samples(rng, 1 << 16)
    test = target_samples(rng, 4096)
    _, l0 = g.gradient(test)
    for epoch in range(30):
        g.train(train, batch=1024, lr=2e-3)
    _, l1 = g.gradient(test)
    g.close()
    print("mean loss before %.3f after %.3f" % (l0 / len(test), l1 / len(test)), flush=True)
    assert l0 / len(test) > 2.0
    assert l1 / len(test) < 1.0
