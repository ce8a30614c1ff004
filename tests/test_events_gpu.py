"""Contact begin, end and hit events (b2World_GetContactEvents) against the reference's.

The events scene throws circles, boxes and capsules onto a ground box with contact events on for
most shapes and hit events for half of them (some with restitution). The fixture
(tests/golden/world_runs.json "events", written by tests/golden/gen_golden.py from the reference
library) holds, per step, the number of begin, end and hit events, an FNV-1a hash of the
flattened event records (shape and contact ids and generations; hit point, normal and approach
speed as float bits) and the transform hash; full records every 25 steps. Parity is exact.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fnv1a32(words):
    h = 0x811C9DC5
    for w in np.asarray(words, dtype=np.uint32).reshape(-1).tolist():
        for b in range(4):
            h ^= (w >> (8 * b)) & 0xFF
            h = (h * 0x01000193) & 0xFFFFFFFF
    return h


def check_step(fx, step, ev, transform_hash):
    want = fx["per_step"][step - 1]
    kinds = ev[:, 0].tolist() if len(ev) else []
    got = [kinds.count(0), kinds.count(1), kinds.count(2), fnv1a32(ev), transform_hash]
    if got != want:
        rec = fx["records"].get(str(step))
        detail = ""
        if rec is not None:
            detail = f"; reference records {np.array(rec, dtype=np.uint32).reshape(-1, 10).tolist()} got {ev.tolist()}"
        raise AssertionError(f"{fx['name']} step {step}: got {got}, reference {want}{detail}")
    rec = fx["records"].get(str(step))
    if rec is not None:
        assert np.array_equal(ev.reshape(-1), np.array(rec, dtype=np.uint32))


@pytest.mark.parametrize("path", [1, 2, 3])
@pytest.mark.parametrize("name", ["events_60", "events_150"])
def test_contact_events_match_reference(gpu_scenes, world_runs, name, path):
    import box2d_amd

    fx = next(e for e in world_runs["events"] if e["name"] == name)
    print(f"[gpu] events scene {name} path {path}", flush=True)
    h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    box2d_amd.set_solver_path(gpu_scenes.world_id(h), path)
    try:
        assert gpu_scenes.body_count(h) == fx["body_count"]
        for step in range(1, fx["steps"] + 1):
            gpu_scenes.step(h)
            check_step(fx, step, gpu_scenes.contact_events(h), gpu_scenes.hash(h))
    finally:
        gpu_scenes.destroy(h)


def test_contact_events_in_a_batch(gpu_scenes, world_runs):
    """Worlds stepped as one batch report the same events as the reference, world by world."""
    import box2d_amd

    fx = next(e for e in world_runs["events"] if e["name"] == "events_60")
    handles = [gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"]) for _ in range(4)]
    ids = (box2d_amd.WorldId * len(handles))(*[gpu_scenes.world_id(h) for h in handles])
    try:
        print(f"[gpu] batch of {len(handles)} event worlds", flush=True)
        for step in range(1, fx["steps"] + 1):
            box2d_amd.step_worlds(ids)
            for h in handles:
                check_step(fx, step, gpu_scenes.contact_events(h), gpu_scenes.hash(h))
    finally:
        for h in handles:
            gpu_scenes.destroy(h)
