"""Guided rendering on the room lit through a door gap: the guided
estimator is unbiased (same image mean as the plain tracer), it collects
training samples and trains, and the trained guide lowers the error at equal
samples per pixel."""
import numpy as np
import pytest

import lambda_amd
from lambda_amd import scenes as gen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def room(tmp_path_factory):
    d = tmp_path_factory.mktemp("room")
    lrd = gen.write_guiding_room(str(d), width=128, height=72, depth=16, samples=1)
    s = lambda_amd.Scene.load(lrd)
    dev = lambda_amd.DeviceScene(s)
    yield dev
    dev.close()


@pytest.fixture(scope="module")
def reference(room):
    p = lambda_amd.Progressive(room, None, seed=0x5EED0F2EF)  # streams independent of the tested runs
    for _ in range(16):
        p.step(256)
    img = p.image()
    p.close()
    return img


def render(room, guide, steps, spp, train=True):
    p = lambda_amd.Progressive(room, guide, path_slots=1 << 18)
    stats = [p.step(spp, train=train) for _ in range(steps)]
    img = p.image()
    p.close()
    return img, stats


def fmean(img, *others):
    return float(img[lambda_amd.finite_pixels(img, *others)].mean())


def test_reference_is_mostly_finite(reference):
    """The plain tracer keeps the reference's 0/0 quirk in rare pixels."""
    bad = (~lambda_amd.finite_pixels(reference)).sum()
    print("non-finite reference pixels: %d of %d" % (bad, reference.shape[0] * reference.shape[1]), flush=True)
    assert bad < reference.shape[0] * reference.shape[1] // 100
    assert fmean(reference) > 0


def step_means(room, guide, steps, spp, train):
    """Image mean of each step's own samples (from the running images) and
    the last step's statistics."""
    p = lambda_amd.Progressive(room, guide, path_slots=1 << 18)
    prev, means, st = None, [], None
    for k in range(1, steps + 1):
        st = p.step(spp, train=train)
        img = p.image().astype(np.float64)
        cur = img if prev is None else k * img - (k - 1) * prev
        means.append(float(cur[lambda_amd.finite_pixels(cur)].mean()))
        prev = img
    p.close()
    return np.array(means), st


def test_guided_estimator_is_unbiased(room):
    """Untrained (near-uniform) and trained guides against the plain tracer:
    per-step image means agree within 4 standard errors."""
    plain, _ = step_means(room, None, 16, 64, False)
    g = lambda_amd.Guide(seed=5)
    untrained, st = step_means(room, g, 16, 64, False)
    assert st["train_samples"] > 10000
    trained, _ = step_means(room, g, 16, 64, True)
    g.close()
    for name, m in (("untrained", untrained), ("trained", trained)):
        se = np.sqrt(m.var(ddof=1) / len(m) + plain.var(ddof=1) / len(plain))
        z = (m.mean() - plain.mean()) / se
        print("%s guide mean %.5g, plain %.5g, z = %.2f" % (name, m.mean(), plain.mean(), z), flush=True)
        assert abs(z) < 4


def test_trained_guide_beats_plain_at_equal_samples(room, reference):
    g = lambda_amd.Guide(seed=6)
    guided, stats = render(room, g, 8, 64)
    g.close()
    plain, _ = render(room, None, 8, 64)
    eg = lambda_amd.rel_mse(guided, reference)
    ep = lambda_amd.rel_mse(plain, reference)
    print("512 spp relMSE: plain %.4g guided %.4g; losses %s" % (
        ep, eg, " ".join("%.3g" % s["loss"] for s in stats)), flush=True)
    gm, rm = fmean(guided, reference), fmean(reference, guided)
    print("guided mean %.5g plain mean %.5g reference %.5g" % (gm, fmean(plain, reference), rm), flush=True)
    assert np.all(np.isfinite(guided))
    assert all(np.isfinite(s["loss"]) for s in stats)
    assert abs(gm - rm) / rm < 0.03
    assert eg < ep
