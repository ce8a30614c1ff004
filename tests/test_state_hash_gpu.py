"""b2World_GetStateHash against the reference's.

The fixture (tests/golden/state_hashes.json, written by tests/golden/gen_golden.py from the
reference library compiled from its own sources) holds the reference's deep state hash
(src/world_snapshot.c:1025-1187) after the listed steps of eight scenes: sleeping pyramids,
mixed piles, all nine joint types with and without sleep, contact events, edits between steps
(destroyed bodies and joints, reused ids) and setters (type changes, disable/enable, shape
destruction). The hash folds body transforms and awake velocities, contact and joint impulses,
every set/color/local index and the id pools, so equality pins the solver-set, graph-coloring and
id bookkeeping as well as the numbers.
"""

import pytest

from conftest import load_golden


def state_scenes():
    return load_golden("state_hashes.json")["scenes"]


SCENE_NAMES = [s["name"] for s in state_scenes()]


def by_name(name):
    return next(s for s in state_scenes() if s["name"] == name)


def test_fixture_is_well_formed():
    scenes = state_scenes()
    assert len(scenes) == 9
    for fx in scenes:
        hashes = fx["hashes"]
        assert len(hashes) == 1 + fx["steps"] // fx["every"]
        assert all(0 <= v < 2**64 for v in hashes.values())
        # the state changes while bodies move (a sleeping scene settles later): not constant
        assert len(set(hashes.values())) > 10, fx["name"]


@pytest.mark.parametrize("name", ["joint_zoo_sleep", "setters_sleep"])
def test_reference_hash_independent_of_workers(ref_scenes, name):
    """The fixture was written with 8 workers; one worker gives the same hashes (the property
    the reference documents for this hash), so the fixture is not tied to a thread schedule."""
    fx = by_name(name)
    h = ref_scenes.create(fx["kind"], fx["iparams"], fx["fparams"], workers=1)
    try:
        assert ref_scenes.state_hash(h) == fx["hashes"]["0"]
        for step in range(1, fx["steps"] + 1):
            ref_scenes.step(h)
            if str(step) in fx["hashes"]:
                assert ref_scenes.state_hash(h) == fx["hashes"][str(step)], f"{name} step {step}"
    finally:
        ref_scenes.destroy(h)


@pytest.mark.gpu
@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("name", SCENE_NAMES)
def test_state_hash_matches_reference(gpu_scenes, name, path):
    import box2d_amd

    fx = by_name(name)
    print(f"[gpu] state hash {name} path {path}", flush=True)
    h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    box2d_amd.set_solver_path(gpu_scenes.world_id(h), path)
    try:
        assert gpu_scenes.state_hash(h) == fx["hashes"]["0"], f"{name} before the first step"
        for step in range(1, fx["steps"] + 1):
            gpu_scenes.step(h)
            want = fx["hashes"].get(str(step))
            if want is not None:
                got = gpu_scenes.state_hash(h)
                assert got == want, f"{name} path {path} step {step}: {got:#018x}, reference {want:#018x}"
    finally:
        gpu_scenes.destroy(h)
