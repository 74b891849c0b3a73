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


def test_forward_over_several_grid_sweeps(guide):
    """More rows than one sweep of the inference grid covers (768 blocks x 4
    waves x 64 rows = 196,608), with a ragged tail: every wave takes three
    chunks, the last one partly filled; the row count is read on the
    device."""
    rng = np.random.default_rng(14)
    p = random_params(rng).astype(np.float32)
    guide.set_params(p)
    n = 2 * 196608 + 37
    f = rng.random((n, 6)).astype(np.float32)
    got = guide.forward(f).astype(np.float64)
    ref, _ = go.forward(p.astype(np.float64), f.astype(np.float64), bf=True)
    ref = ref[:, :go.N_OUT]
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("forward over %d rows: relative L2 error %.3g" % (n, err), flush=True)
    assert err < 5e-3
    # per chunk: the first, a sweep boundary, the ragged tail
    for a, b in ((0, 64), (196608 - 64, 196608 + 64), (n - 101, n)):
        e = np.linalg.norm(got[a:b] - ref[a:b]) / np.linalg.norm(ref[a:b])
        assert e < 1e-2, (a, b, e)


def test_adam_steps_match_oracle(guide):
    """Ten Adam steps from set parameters with given gradients (zeros,
    tiny and large entries, varying scale): FP32 master weights, both
    moments and the step count against the FP64 oracle's Adam; both
    bfloat16 copies equal the master weights rounded to bfloat16, W_l and
    W_l^T."""
    rng = np.random.default_rng(15)
    p = random_params(rng).astype(np.float32)
    guide.set_params(p)
    st = guide.state()
    assert st["step"] == 0 and not st["m"].any() and not st["v"].any()
    pr = p.astype(np.float64)
    m = np.zeros(go.N_PARAMS)
    v = np.zeros(go.N_PARAMS)
    menv = np.zeros(go.N_PARAMS)     # the first moment of |g|: bounds m's rounding under cancellation
    lr = 2e-3
    for k in range(1, 11):
        g = (rng.normal(size=go.N_PARAMS) * 10.0 ** rng.uniform(-6, 2, go.N_PARAMS)).astype(np.float32)
        g[rng.random(go.N_PARAMS) < 0.05] = 0
        scale = np.float32(1.0 / (64 * k))
        guide.adam_step(g, float(scale), lr)
        pr, m, v = go.adam(pr, m, v, g.astype(np.float64) * float(scale), k, lr=lr)
        menv = 0.9 * menv + 0.1 * np.abs(g.astype(np.float64) * float(scale))
        st = guide.state()
        got = guide.params().astype(np.float64)
        assert st["step"] == k
        assert np.all(np.abs(st["m"] - m) <= 2e-6 * menv), k
        np.testing.assert_allclose(st["v"], v, rtol=2e-5, atol=1e-18)
        # each update is at most lr in size; float32 rounding of p and of
        # the update
        assert np.max(np.abs(got - pr)) < 2e-6, (k, np.max(np.abs(got - pr)))
        for l in range(go.LAYERS):
            kk = go.layer_in(l)
            w = go.bf16(got[go.w_offset(l):go.w_offset(l) + kk * 64].astype(np.float32)).reshape(kk, 64)
            off = 0 if l == 0 else 96 * 64 + (l - 1) * 64 * 64
            assert np.array_equal(st["w_kn"][off:off + kk * 64].reshape(kk, 64), w), l
            assert np.array_equal(st["w_nk"][off:off + kk * 64].reshape(64, kk), w.T), l
    assert np.max(np.abs(pr - p)) > 5 * lr   # the weights moved


def test_training_batch_is_one_adam_step_of_the_oracle_gradient():
    """train(samples, batch = all of them) from a reset Adam state, all
    weights 1 (below the weight cap): one Adam step with the mean gradient,
    whose first update is -lr g / (|g| + eps), i.e. lr times the sign of
    the oracle's gradient wherever that gradient is not near zero; the next
    forward pass uses the new weights."""
    rng = np.random.default_rng(16)
    g = lambda_amd.Guide(seed=2)
    p = random_params(rng).astype(np.float32)
    g.set_params(p)
    s = random_samples(rng, 1024).astype(np.float32)
    s[:, 9] = 1.0
    lr = 1e-3
    g.train(s, batch=len(s), lr=lr)
    got = g.params().astype(np.float64)
    st = g.state()
    gref, _ = go.gradient(p.astype(np.float64), s.astype(np.float64), bf=True)
    want, _, _ = go.adam(p.astype(np.float64), 0, 0, gref / len(s), 1, lr=lr)
    assert st["step"] == 1
    d = np.abs(got - want)
    big = np.abs(gref / len(s)) > 1e-3 * np.abs(gref / len(s)).max()
    print("params within 1e-6 of the oracle's Adam step: %.4f (large gradients %.4f)" % (
        np.mean(d < 1e-6), np.mean(d[big] < 1e-6)), flush=True)
    assert np.max(np.abs(got - p)) <= lr * (1 + 1e-5)
    assert np.mean(d[big] < 1e-6) > 0.995
    f = rng.random((500, 6)).astype(np.float32)
    out = g.forward(f).astype(np.float64)
    ref, _ = go.forward(got, f.astype(np.float64), bf=True)
    err = np.linalg.norm(out - ref[:, :go.N_OUT]) / np.linalg.norm(ref[:, :go.N_OUT])
    g.close()
    assert err < 5e-3


# 1000 rows: 16 blocks, one row group of the reduction; 6437 rows: 101
# blocks, a full group of 64 rows and a ragged one of 37, ragged last chunk
@pytest.mark.parametrize("n", [1000, 64 * 100 + 37])
def test_gradient_matches_oracle(guide, n):
    rng = np.random.default_rng(11)
    p = random_params(rng).astype(np.float32)
    guide.set_params(p)
    s = random_samples(rng, n).astype(np.float32)
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


def test_gradient_over_several_chunks_per_block(guide):
    """Batches above 512 x 64 rows run several 64-row chunks per block with
    the gradient held in registers across them: the gradient of such a
    batch (ragged tail included) equals the sum of the gradients of parts
    that take one chunk per block."""
    rng = np.random.default_rng(13)
    guide.set_params(random_params(rng).astype(np.float32))
    n = 2 * 512 * 64 + 37
    s = random_samples(rng, n).astype(np.float32)
    g, loss = guide.gradient(s)
    cuts = (0, 30000, 60000, n)
    parts = [guide.gradient(s[a:b]) for a, b in zip(cuts[:-1], cuts[1:])]
    gsum = sum(p[0].astype(np.float64) for p in parts)
    lsum = sum(p[1] for p in parts)
    assert abs(loss - lsum) <= 1e-4 * abs(lsum)
    np.testing.assert_allclose(g, gsum, rtol=1e-4, atol=1e-4 * np.abs(gsum).max())


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
