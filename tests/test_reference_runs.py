"""The whole-run fixtures against the reference library itself (CPU; needs oracle/_ref).

These pin the fixtures: the reference reproduces them, its determinism known answer holds
(test/test_determinism.c:19-24: FallingHinges sleeps at step 274 with hash 0xE86690F4), and its
results do not depend on the worker count (test/test_determinism.c:66-124).
"""

import pytest


def test_falling_hinges_known_answer(ref_scenes, world_runs):
    ok, step, hsh = ref_scenes.falling_hinges(1, 500)
    assert ok and step == 274 and hsh == 0xE86690F4
    fh = world_runs["falling_hinges"]
    assert (fh["sleep_step"], fh["hash"]) == (274, 0xE86690F4)


@pytest.mark.parametrize("name", ["hello", "pyramid20", "pyramid20_sleep", "mixed_pile_1"])
def test_reference_reproduces_fixture(ref_scenes, world_runs, name):
    fx = next(s for s in world_runs["scenes"] if s["name"] == name)
    h = ref_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    try:
        assert ref_scenes.hash(h) == fx["hash_initial"]
        limit = min(fx["steps"], 150)
        for s in range(1, limit + 1):
            ref_scenes.step(h)
            if str(s) in fx["hashes"]:
                assert ref_scenes.hash(h) == fx["hashes"][str(s)], f"step {s}"
    finally:
        ref_scenes.destroy(h)


def test_reference_worker_count_independent(ref_scenes):
    hashes = []
    for workers in (1, 4):
        h = ref_scenes.pyramid(20, 40.0, workers=workers)
        ref_scenes.step(h, 120)
        hashes.append(ref_scenes.hash(h))
        ref_scenes.destroy(h)
    assert hashes[0] == hashes[1]
