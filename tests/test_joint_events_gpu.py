"""Joint events (b2World_GetJointEvents) against the reference's.

The fixture (tests/golden/joint_events.json, written by tests/golden/gen_golden.py from the
reference library compiled from its own sources) steps the joint zoo (all nine joint types) with
event thresholds: the joints in ascending id carry user data k + 1 and thresholds that report
every awake step, a force of 40 N, a torque of 2 N m, or never, and before step 151 every joint
moves one place in that cycle through b2Joint_SetForceThreshold / SetTorqueThreshold. After each
step the fixture holds the event count, an FNV-1a hash of the records (joint index1, generation,
user data) and the transform hash, and the full records every 50 steps. With sleep on, sleeping
joints stop reporting (reference src/solver.c:1648-1700, src/joint.c:993-1072, 1663-1675).
"""

import numpy as np
import pytest

from conftest import load_golden


def event_scenes():
    return load_golden("joint_events.json")["scenes"]


def fnv1a32(words):
    h = 0x811C9DC5
    for w in np.asarray(words, dtype=np.uint32).reshape(-1).tolist():
        for b in range(4):
            h ^= (w >> (8 * b)) & 0xFF
            h = (h * 0x01000193) & 0xFFFFFFFF
    return h


def test_fixture_covers_thresholds():
    for fx in event_scenes():
        counts = [p[0] for p in fx["per_step"]]
        assert len(counts) == fx["steps"]
        # some joints report and some do not, and the threshold change alters the set
        assert 0 < min(counts) and max(counts) < 64, fx["name"]
        before = fx["records"]["150"]
        after = fx["records"]["200"]
        assert before != after
        # records come in ascending joint index with the user data the scene assigned
        rec = np.asarray(after, dtype=np.int64).reshape(-1, 3)
        assert np.all(np.diff(rec[:, 0]) > 0)
        assert np.all(rec[:, 2] >= 1)
        assert fnv1a32(after) == fx["per_step"][199][1]


def test_reference_events_independent_of_workers(ref_scenes):
    """The fixture used 8 workers; the reference gathers per-worker bit sets, so one worker must
    give the same events."""
    fx = event_scenes()[1]
    h = ref_scenes.create(fx["kind"], fx["iparams"], fx["fparams"], workers=1)
    try:
        for step in range(1, fx["steps"] + 1):
            ref_scenes.step(h)
            ev = ref_scenes.joint_events(h)
            assert [len(ev), fnv1a32(ev), ref_scenes.hash(h)] == fx["per_step"][step - 1], f"step {step}"
    finally:
        ref_scenes.destroy(h)


@pytest.mark.gpu
@pytest.mark.parametrize("path", [1, 2, 3])
@pytest.mark.parametrize("name", ["joint_zoo_events", "joint_zoo_events_sleep"])
def test_joint_events_match_reference(gpu_scenes, name, path):
    import box2d_amd

    fx = next(s for s in event_scenes() if s["name"] == name)
    print(f"[gpu] joint events {name} path {path}", flush=True)
    h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    box2d_amd.set_solver_path(gpu_scenes.world_id(h), path)
    try:
        for step in range(1, fx["steps"] + 1):
            gpu_scenes.step(h)
            ev = gpu_scenes.joint_events(h)
            got = [len(ev), fnv1a32(ev), gpu_scenes.hash(h)]
            want = fx["per_step"][step - 1]
            assert got == want, f"{name} path {path} step {step}: [count, records, transforms] {got}, reference {want}"
            if str(step) in fx["records"]:
                assert ev.reshape(-1).tolist() == fx["records"][str(step)]
    finally:
        gpu_scenes.destroy(h)
