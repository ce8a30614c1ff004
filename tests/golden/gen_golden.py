"""Write the whole-run golden fixtures from the reference library.

Run in a container that holds the reference sources, after `make -C oracle ref` has compiled
them into oracle/_ref/libb2ref_scenes.so (the reference's own src/*.c plus the public-API scene
driver box2d_amd/scenes/scenes.c). Every fixture is data only: scene parameters, the step
protocol, and the reference's outputs (djb2 transform hashes as in shared/determinism.c:110-125,
world counters, and final transform bits for the small scenes).

    python tests/golden/gen_golden.py [--events-and-hashes-only]
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from box2d_amd.scenes_api import (SceneLib, SCENE_HELLO, SCENE_PYRAMID, SCENE_MIXED_PILE, SCENE_FALLING_HINGES,  # noqa: E402
                                  SCENE_RAIN, SCENE_JOINT_ZOO, SCENE_EVENTS, SCENE_EDITS, SCENE_SETTERS,
                                  SCENE_PRESOLVE, SCENE_SENSORS, SCENE_BULLETS)

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libb2ref_scenes.so")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).reshape(-1).tolist()


def trace(lib, h, steps, every):
    """Hash after each step listed by `every` (1-based step numbers)."""
    out = {}
    for s in range(1, steps + 1):
        lib.step(h)
        if s in every:
            out[str(s)] = lib.hash(h)
    return out


def scene_fixture(lib, name, kind, iparams, fparams, steps, every, keep_transforms):
    # the reference's results do not depend on the worker count; 8 workers keep generation quick
    h = lib.create(kind, iparams, fparams, workers=8)
    fx = {
        "name": name,
        "kind": kind,
        "iparams": list(iparams),
        "fparams": list(fparams),
        "dt": 1.0 / 60.0,
        "substeps": 4,
        "steps": steps,
        "body_count": lib.body_count(h),
        "hash_initial": lib.hash(h),
    }
    fx["hashes"] = trace(lib, h, steps, every)
    fx["counters"] = lib.counters(h)
    fx["awake_body_count"] = lib.awake_body_count(h)
    if keep_transforms:
        fx["transforms_bits"] = bits(lib.transforms(h))
        fx["velocities_bits"] = bits(lib.velocities(h))
    lib.destroy(h)
    return fx


def fnv1a32(words):
    h = 0x811C9DC5
    for w in np.asarray(words, dtype=np.uint32).reshape(-1).tolist():
        for b in range(4):
            h ^= (w >> (8 * b)) & 0xFF
            h = (h * 0x01000193) & 0xFFFFFFFF
    return h


def events_fixture(lib, name, count, seed, steps):
    """Per-step contact events of the events scene: counts of begin, end and hit events and an
    FNV-1a hash of the flattened records (b2s_contact_events), the transform hash, and the full
    records every 25 steps."""
    h = lib.create(SCENE_EVENTS, (count, seed), (), workers=8)
    fx = {"name": name, "kind": SCENE_EVENTS, "iparams": [count, seed], "fparams": [], "dt": 1.0 / 60.0,
          "substeps": 4, "steps": steps, "body_count": lib.body_count(h), "per_step": [], "records": {}}
    for step in range(1, steps + 1):
        lib.step(h)
        ev = lib.contact_events(h)
        kinds = ev[:, 0].tolist() if len(ev) else []
        fx["per_step"].append([kinds.count(0), kinds.count(1), kinds.count(2), fnv1a32(ev), lib.hash(h)])
        if step % 25 == 0:
            fx["records"][str(step)] = ev.reshape(-1).tolist()
    lib.destroy(h)
    return fx


def edits_fixture(lib, name, sleep, steps, kind=SCENE_EDITS, extra=()):
    """Per step of a scripted scene (edits, setters, pre-solve, sensors, bullets): the transform
    hash of the listed bodies, the probe hash and the listed body count."""
    iparams = (sleep,) + tuple(extra)
    h = lib.create(kind, iparams, ())
    per_step = []
    for _ in range(steps):
        lib.step(h)
        per_step.append([lib.hash(h), lib.probe_hash(h), lib.body_count(h)])
    lib.destroy(h)
    return {"name": name, "kind": kind, "iparams": list(iparams), "fparams": [], "steps": steps, "per_step": per_step}


# (name, kind, iparams, fparams, steps, every): scenes whose b2World_GetStateHash is recorded
STATE_HASH_SCENES = [
    ("pyramid20_sleep", SCENE_PYRAMID, (20, 1), (40.0,), 500, 1),
    ("mixed_pile_1", SCENE_MIXED_PILE, (64, 1, 1), (), 300, 1),
    ("joint_zoo", SCENE_JOINT_ZOO, (0,), (), 400, 1),
    ("joint_zoo_sleep", SCENE_JOINT_ZOO, (1,), (), 400, 1),
    ("events_60", SCENE_EVENTS, (60, 12345), (), 300, 1),
    ("edits_sleep", SCENE_EDITS, (1,), (), 240, 1),
    ("setters_sleep", SCENE_SETTERS, (1,), (), 200, 1),
    ("rain_small", SCENE_RAIN, (2000, 40, 0, 7), (48.0,), 400, 10),
    ("bullets_sleep", SCENE_BULLETS, (1, 0), (), 200, 1),
]


def state_hash_fixtures(lib):
    """The reference's deep state hash (b2World_GetStateHash) after every `every`-th step."""
    out = []
    for name, kind, iparams, fparams, steps, every in STATE_HASH_SCENES:
        h = lib.create(kind, iparams, fparams, workers=8)
        hashes = {"0": lib.state_hash(h)}
        for step in range(1, steps + 1):
            lib.step(h)
            if step % every == 0:
                hashes[str(step)] = lib.state_hash(h)
        lib.destroy(h)
        out.append({"name": name, "kind": kind, "iparams": list(iparams), "fparams": list(fparams),
                    "steps": steps, "every": every, "hashes": hashes})
    return out


def joint_event_fixture(lib, name, sleep, steps):
    """Per step of the joint zoo with event thresholds: the joint event count, an FNV-1a hash of
    the records (b2s_joint_events) and the transform hash; full records every 50 steps."""
    h = lib.create(SCENE_JOINT_ZOO, (sleep, 1), (), workers=8)
    fx = {"name": name, "kind": SCENE_JOINT_ZOO, "iparams": [sleep, 1], "fparams": [], "steps": steps,
          "per_step": [], "records": {}}
    for step in range(1, steps + 1):
        lib.step(h)
        ev = lib.joint_events(h)
        fx["per_step"].append([len(ev), fnv1a32(ev), lib.hash(h)])
        if step % 50 == 0:
            fx["records"][str(step)] = ev.reshape(-1).tolist()
    lib.destroy(h)
    return fx


def write_joint_events(lib):
    path = os.path.join(HERE, "joint_events.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/gen_golden.py",
                   "reference": "erincatto/box2d v3.2 b2World_GetJointEvents (src/solver.c, src/joint.c)",
                   "scenes": [joint_event_fixture(lib, "joint_zoo_events", 0, 300),
                              joint_event_fixture(lib, "joint_zoo_events_sleep", 1, 300)]},
                  f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


def write_presolve(lib):
    """The pre-solve scene (one-way platform) and the sensor scene, each with sleep off and on: per
    step the transform hash, the probe hash (the pre-solve callback's tally, the platform's flag
    and contact count; the sensor events and overlaps) and the body count."""
    path = os.path.join(HERE, "presolve.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/gen_golden.py",
                   "reference": "erincatto/box2d v3.2 pre-solve callback (src/contact.c:573-583), sensors (src/sensor.c), bullets and the pre-continuous callback (src/solver.c:181-550, 1917-1989)",
                   "scenes": [edits_fixture(lib, "presolve", 0, 200, SCENE_PRESOLVE),
                              edits_fixture(lib, "presolve_sleep", 1, 200, SCENE_PRESOLVE),
                              edits_fixture(lib, "sensors", 0, 200, SCENE_SENSORS),
                              edits_fixture(lib, "sensors_sleep", 1, 200, SCENE_SENSORS),
                              edits_fixture(lib, "bullets", 0, 200, SCENE_BULLETS, (0,)),
                              edits_fixture(lib, "bullets_sleep", 1, 200, SCENE_BULLETS, (0,)),
                              edits_fixture(lib, "bullets_precontinuous", 0, 200, SCENE_BULLETS, (1,))]},
                  f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


def write_state_hashes(lib):
    path = os.path.join(HERE, "state_hashes.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/gen_golden.py",
                   "reference": "erincatto/box2d v3.2 b2World_GetStateHash (src/world_snapshot.c)",
                   "scenes": state_hash_fixtures(lib)}, f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


def main():
    if not os.path.exists(REF_LIB):
        sys.exit(f"{REF_LIB} missing: run `make -C oracle ref` first")
    lib = SceneLib(REF_LIB)
    write_state_hashes(lib)
    write_joint_events(lib)
    write_presolve(lib)
    if "--events-and-hashes-only" in sys.argv:
        return
    fixtures = []
    fixtures.append(scene_fixture(lib, "hello", SCENE_HELLO, (), (), 120, set(range(1, 121)), True))
    fixtures.append(scene_fixture(lib, "pyramid20", SCENE_PYRAMID, (20, 0), (40.0,), 500,
                                  set(range(1, 501)), True))
    fixtures.append(scene_fixture(lib, "pyramid20_sleep", SCENE_PYRAMID, (20, 1), (40.0,), 500,
                                  set(range(1, 501)), True))
    fixtures.append(scene_fixture(lib, "pyramid100", SCENE_PYRAMID, (100, 0), (100.0,), 500,
                                  {1, 2, 5, 10, 25, 50, 100, 200, 300, 400, 500}, False))
    for seed in (1, 2, 3):
        fixtures.append(scene_fixture(lib, f"mixed_pile_{seed}", SCENE_MIXED_PILE, (64, 1, seed), (), 300,
                                      set(range(1, 301)), True))
    fixtures.append(scene_fixture(lib, "rain_small", SCENE_RAIN, (2000, 40, 0, 7), (48.0,), 400, set(range(1, 401)), True))
    fixtures.append(scene_fixture(lib, "rain", SCENE_RAIN, (20000, 100, 0, 12345), (120.0,), 1000,
                                  set(range(10, 1001, 10)) | {1, 2, 5}, False))
    fixtures.append(scene_fixture(lib, "falling_hinges_trace", SCENE_FALLING_HINGES, (), (), 300, set(range(1, 301)),
                                  True))
    fixtures.append(scene_fixture(lib, "joint_zoo", SCENE_JOINT_ZOO, (0,), (), 400, set(range(1, 401)), True))
    fixtures.append(scene_fixture(lib, "joint_zoo_sleep", SCENE_JOINT_ZOO, (1,), (), 400, set(range(1, 401)), True))
    events = [events_fixture(lib, "events_60", 60, 12345, 300), events_fixture(lib, "events_150", 150, 7, 240)]
    edits = [edits_fixture(lib, "edits", 0, 240), edits_fixture(lib, "edits_sleep", 1, 240),
             edits_fixture(lib, "setters", 0, 200, SCENE_SETTERS), edits_fixture(lib, "setters_sleep", 1, 200, SCENE_SETTERS)]
    ok, step, hsh = lib.falling_hinges(1, 500)
    hinges = {"name": "falling_hinges", "slept": ok, "sleep_step": step, "hash": hsh}
    out = {"generator": "tests/golden/gen_golden.py", "reference": "erincatto/box2d v3.2 (src/*.c, -O2 -ffp-contract=off)",
           "scenes": fixtures, "falling_hinges": hinges, "events": events, "edits": edits}
    path = os.path.join(HERE, "world_runs.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes); falling hinges {ok} {step} {hsh:#x}")


if __name__ == "__main__":
    main()
