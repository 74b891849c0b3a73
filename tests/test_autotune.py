"""The schedule autotuner: its search space, its cache key and cache, and
(on the GPU) that it returns the fastest candidate and that every candidate
renders the same reference-stream image."""
import json
import types

import numpy as np
import pytest

import lambda_amd
from lambda_amd import autotune
from conftest import load_golden_scene


def test_candidates_cover_the_schedules():
    ref = autotune.candidates(4096, lambda_amd.LAMBDA_RNG_REFERENCE)
    thr = autotune.candidates(4096, lambda_amd.LAMBDA_RNG_THROUGHPUT)
    fields = {f for f, _ in lambda_amd.RenderParams._fields_}
    for c in ref + thr:
        assert set(c) <= fields
    # persistent with and without the cost order, passes with a tail, graph
    assert any(c["tail_paths"] == 0 and c["cost_order"] > 0 for c in ref)
    assert any(c["tail_paths"] == 0 and c["cost_order"] == -1 for c in ref)
    assert any(c["tail_paths"] > 0 for c in ref)
    assert any(c.get("graph_passes", 0) > 0 for c in ref)
    # the cost order applies to reference streams only
    assert all(c["cost_order"] == 0 for c in thr)
    # tile sizes and the compiled occupancy variants
    assert {c["tile"] for c in ref} == {4, 8, 16}
    assert {c["occupancy"] for c in ref} == {2, 3, 4}
    # the voting schedule is tried with independent streams only
    assert not any(c.get("vote") for c in ref)
    assert {c["occupancy"] for c in thr if c.get("vote")} == {3, 4}


def test_key_depends_on_scene_and_frame_and_cache_is_used(tmp_path):
    box = types.SimpleNamespace(scene=load_golden_scene("box"))   # no render(): a cache hit renders nothing
    mesh = types.SimpleNamespace(scene=load_golden_scene("mesh"))
    k = autotune.key(box)
    assert k == autotune.key(box)
    assert k != autotune.key(mesh)
    assert k != autotune.key(box, width=box.scene.params.width + 8)
    cache = tmp_path / "tune.json"
    stored = {"best": {"tail_paths": 0, "path_slots": 0, "cost_order": 8}, "table": [], "trial_spp": 16}
    cache.write_text(json.dumps({k: stored}))
    assert autotune.tune(box, cache=str(cache)) == stored


@pytest.mark.gpu
def test_tune_picks_the_fastest_and_candidates_agree():
    scene = load_golden_scene("mesh")
    dev = lambda_amd.DeviceScene(scene)
    out = autotune.tune(dev, trial_spp=16, repeats=1, cache=None)
    table = out["table"]
    assert len(table) == len(autotune.candidates(scene.params.width * scene.params.height, scene.params.rng_mode))
    assert out["best"] == min(table, key=lambda e: e[1])[0]
    print("autotune:", [(c, round(ms, 2)) for c, ms in table], flush=True)
    # reference streams: every schedule renders the same image bit for bit
    ref = dev.render(tail_paths=-1, cost_order=-1)
    for cand, _ in table:
        img = dev.render(**cand)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), cand
    dev.close()
