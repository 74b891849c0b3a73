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


def test_live_server_images_match_single_renders(devices):
    """Jobs submitted to the live server, with a budget small enough that
    they run in several admitted groups: every image bit-identical to a
    single render; tickets move from queued to done."""
    srv = lambda_amd.Server(devices, hbm_budget=80000, max_launches=2)
    subs = []
    for job in JOBS:
        jd = dict(job)
        sc = jd.pop("scene")
        jd.pop("arrival_ms")
        subs.append(srv.submit(sc, **jd))
    stats = srv.drain()
    results = []
    for (t, img), job in zip(subs, JOBS):
        state, res = srv.poll(t)
        assert state == lambda_amd.JOB_DONE
        results.append(res)
        assert np.array_equal(img.view(np.uint32), single(devices, job).view(np.uint32)), job
    srv.close()
    print("launches %d, wall %.1f ms, latencies %s" % (
        stats["batches"], stats["wall_ms"], " ".join("%.1f" % r["latency_ms"] for r in results)), flush=True)
    assert stats["batches"] >= 2
    assert len({r["batch"] for r in results}) == stats["batches"]
    assert all(r["latency_ms"] > 0 for r in results)
    assert stats["paths"] == sum(j["width"] * j["height"] * j["samples"] for j in JOBS)


def test_live_server_jobs_arriving_over_time(devices):
    """Submissions from another thread while earlier jobs render: each
    finishes after its submission, and the queue drains."""
    import threading
    import time
    srv = lambda_amd.Server(devices)
    out = []

    def feed():
        for k in range(8):
            out.append(srv.submit(k % 4, width=64, height=48, samples=16, sqrt_samples=4, depth=8))
            time.sleep(0.01)

    th = threading.Thread(target=feed)
    th.start()
    th.join()
    stats = srv.drain()
    res = [srv.poll(t)[1] for t, _ in out]
    srv.close()
    print("launches %d; finish ms %s" % (stats["batches"], " ".join("%.1f" % r["finish_ms"] for r in res)), flush=True)
    assert all(r["latency_ms"] > 0 for r in res)
    assert res[-1]["finish_ms"] > res[0]["finish_ms"]


def test_live_server_rejects_bad_jobs(devices):
    srv = lambda_amd.Server(devices)
    with pytest.raises(lambda_amd.LambdaError):
        srv.submit(9, width=8, height=8)
    with pytest.raises(lambda_amd.LambdaError):
        srv.submit(0, width=8, height=8, samples=5, sqrt_samples=2)
    with pytest.raises(lambda_amd.LambdaError):
        srv.poll(12345)
    assert srv.drain()["batches"] == 0
    srv.close()
