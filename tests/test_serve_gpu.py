"""Batched serving: many jobs on several scenes in persistent launches.
Every job's image is bit-identical to a single render of the same settings
(reference streams), whatever the batch split; arrivals delay the jobs."""
import numpy as np
import pytest

import lambda_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def devices(scenes):
    out = [lambda_amd.DeviceScene(scenes[n]) for n in ("box", "film", "mesh", "room")]
    yield out
    for d in out:
        d.close()


JOBS = [
    dict(scene=0, width=40, height=32, samples=4, sqrt_samples=2, depth=6, arrival_ms=0.0),
    dict(scene=1, width=33, height=27, samples=9, sqrt_samples=3, depth=12, arrival_ms=5.0),
    dict(scene=2, width=48, height=20, samples=1, sqrt_samples=1, depth=10, arrival_ms=2.0),
    dict(scene=3, width=25, height=25, samples=16, sqrt_samples=4, depth=16, arrival_ms=30.0,
         lambda_min=360.0, lambda_max=830.0),
    dict(scene=0, width=64, height=64, samples=4, sqrt_samples=2, depth=4, arrival_ms=1.0, seed=7),
    dict(scene=2, width=17, height=61, samples=4, sqrt_samples=2, depth=8, arrival_ms=12.0),
]


def single(devices, job):
    jd = dict(job)
    sc = jd.pop("scene")
    jd.pop("arrival_ms")
    return devices[sc].render(tail_paths=-1, cost_order=-1, **jd)


@pytest.mark.parametrize("budget", [0, 40000])
def test_served_images_match_single_renders(devices, budget):
    images, results, stats = lambda_amd.serve(devices, JOBS, hbm_budget=budget)
    print("batches %d, wall %.1f ms, latencies %s" % (
        stats["batches"], stats["wall_ms"], " ".join("%.1f" % r["latency_ms"] for r in results)), flush=True)
    if budget:
        assert stats["batches"] > 1
    else:
        assert stats["batches"] == 1
    for job, img in zip(JOBS, images):
        ref = single(devices, job)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), job
    assert stats["paths"] == sum(j["width"] * j["height"] * j["samples"] for j in JOBS)


def test_arrivals_delay_jobs(devices):
    """A job that arrives 50 ms after the start cannot finish before then."""
    jobs = [dict(scene=0, width=32, height=32, samples=1, sqrt_samples=1, depth=4, arrival_ms=0.0),
            dict(scene=0, width=32, height=32, samples=1, sqrt_samples=1, depth=4, arrival_ms=50.0)]
    _, results, _ = lambda_amd.serve(devices, jobs)
    print(results, flush=True)
    assert results[1]["finish_ms"] >= 50.0
    assert results[0]["finish_ms"] < results[1]["finish_ms"]
    assert all(r["latency_ms"] >= 0 for r in results)
