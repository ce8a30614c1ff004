"""GPU world step parity: b2World_Step through the C API against the reference's whole runs.

Each scene is built through the public Box2D API by the scene driver (box2d_amd/scenes/scenes.c)
and stepped at 1/60 s with 4 substeps. The fixture (tests/golden/world_runs.json, written by
tests/golden/gen_golden.py from the reference library) holds the djb2 hash of all body
transforms after the listed steps (shared/determinism.c:110-125), the world counters and, for
small scenes, the final transforms and velocities as float bits. Parity is bit-exact.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

def scene(world_runs, name):
    return next(s for s in world_runs["scenes"] if s["name"] == name)


def run_and_compare(gpu_scenes, fx, path=0, pair_query=None):
    """Step the scene and compare with the fixture. pair_query forces the broadphase pair query
    (1 host tree, 2 GPU chunk scan, 3 GPU LBVH walk, 4 GPU chunk scan at the update only);
    returns how many steps ran each path (0 none, 1 host, 2 GPU at the update, 3 GPU speculative)."""
    import box2d_amd

    lib = box2d_amd.library()
    h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    wid = gpu_scenes.world_id(h)
    box2d_amd.set_solver_path(wid, path)
    if pair_query is not None:
        lib.b2amd_World_SetPairQuery(wid, pair_query)
    used = {0: 0, 1: 0, 2: 0, 3: 0}
    try:
        assert gpu_scenes.body_count(h) == fx["body_count"]
        assert gpu_scenes.hash(h) == fx["hash_initial"]
        for s in range(1, fx["steps"] + 1):
            gpu_scenes.step(h)
            used[lib.b2amd_World_GetLastPairQuery(wid)] += 1
            want = fx["hashes"].get(str(s))
            if want is not None:
                got = gpu_scenes.hash(h)
                assert got == want, f"{fx['name']}: transform hash differs first at step {s}: {got:#x} != {want:#x}"
        assert gpu_scenes.counters(h) == fx["counters"]
        assert gpu_scenes.awake_body_count(h) == fx["awake_body_count"]
        if "transforms_bits" in fx:
            got = gpu_scenes.transforms(h).view(np.uint32).reshape(-1)
            assert np.array_equal(got, np.array(fx["transforms_bits"], dtype=np.uint32))
            got = gpu_scenes.velocities(h).view(np.uint32).reshape(-1)
            assert np.array_equal(got, np.array(fx["velocities_bits"], dtype=np.uint32))
    finally:
        gpu_scenes.destroy(h)
    return used


@pytest.mark.parametrize("mode", [2, 4, 3], ids=["chunk_scan_speculative", "chunk_scan_at_update", "lbvh"])
@pytest.mark.parametrize("name", ["hello", "pyramid20_sleep", "mixed_pile_1", "mixed_pile_2", "mixed_pile_3",
                                  "falling_hinges_trace", "rain_small"])
def test_gpu_pair_query_matches_reference(gpu_scenes, world_runs, name, mode):
    """Broadphase pairs from the GPU query on every step: the chunk-box scan run speculatively
    behind the previous solve where it still applies (mode 2), the chunk-box scan at the pair
    update (mode 4) or the LBVH walk (mode 3), ordered in the reference trees' visit order. The
    whole run stays bit-identical to the reference."""
    print(f"[gpu] scene {name} (GPU pair query, mode {mode})", flush=True)
    used = run_and_compare(gpu_scenes, scene(world_runs, name), pair_query=mode)
    assert used[2] + used[3] > 0 and used[1] == 0
    if mode == 2 and used[2] + used[3] > 1:
        # after the first step's query, the speculative query serves the steps that move proxies
        assert used[3] > 0, used
    if mode != 2:
        assert used[3] == 0, used


@pytest.mark.parametrize("name", ["hello", "pyramid20", "pyramid100", "pyramid20_sleep"])
def test_scene_matches_reference(gpu_scenes, world_runs, name):
    """Default path: the fused single-workgroup solver for worlds that fit its LDS budget."""
    print(f"[gpu] scene {name} (auto path)", flush=True)
    run_and_compare(gpu_scenes, scene(world_runs, name))


@pytest.mark.parametrize("name", ["pyramid100"])
def test_scene_matches_reference_fused(gpu_scenes, world_runs, name):
    """Fused single-workgroup solver forced on a world above the auto threshold."""
    print(f"[gpu] scene {name} (fused)", flush=True)
    run_and_compare(gpu_scenes, scene(world_runs, name), path=2)


@pytest.mark.parametrize("name", ["hello", "pyramid20", "pyramid100", "pyramid20_sleep"])
def test_scene_matches_reference_color_kernels(gpu_scenes, world_runs, name):
    """Per-color kernel path (the path of worlds too large for the fused solver)."""
    print(f"[gpu] scene {name} (color kernels)", flush=True)
    run_and_compare(gpu_scenes, scene(world_runs, name), path=1)


@pytest.mark.parametrize("name", ["hello", "pyramid20", "pyramid100", "pyramid20_sleep", "falling_hinges_trace"])
def test_scene_matches_reference_xcd(gpu_scenes, world_runs, name):
    """XCD solver forced: the substep loop in one launch on one XCD's compute units, colors
    separated by barriers among its workgroups."""
    print(f"[gpu] scene {name} (xcd solver)", flush=True)
    run_and_compare(gpu_scenes, scene(world_runs, name), path=3)


@pytest.mark.parametrize("path", [0, 1, 3])
@pytest.mark.parametrize("name", ["mixed_pile_1", "mixed_pile_2", "mixed_pile_3", "falling_hinges_trace"])
def test_continuous_scenes_match_reference(gpu_scenes, world_runs, name, path):
    """Circles, capsules, hulls and rounded boxes dropped from height (fast bodies take the
    continuous collision kernel), and the hinged-box determinism scene with revolute joints."""
    print(f"[gpu] scene {name} path {path}", flush=True)
    run_and_compare(gpu_scenes, scene(world_runs, name), path=path)


@pytest.mark.parametrize("path", [1, 2, 3])
@pytest.mark.parametrize("name", ["joint_zoo", "joint_zoo_sleep"])
def test_joint_types_match_reference(gpu_scenes, world_runs, name, path):
    """Weld, prismatic, wheel, distance, motor, filter and revolute joints with their springs,
    limits and motors, among contacts: every step's transform hash equals the reference's, on the
    color kernels (per-body warm start once the graph is stable) and on the fused solver."""
    print(f"[gpu] scene {name} path {path}", flush=True)
    run_and_compare(gpu_scenes, scene(world_runs, name), path=path)


@pytest.mark.parametrize("path", [0, 1, 2, 3])
def test_rain_small_matches_reference(gpu_scenes, world_runs, path):
    """2000 mixed shapes raining into a static U-shaped chain container (one-sided chain segments
    with ghost vertices, GJK simplex caches, continuous collision against the chain)."""
    print(f"[gpu] rain_small path {path}", flush=True)
    run_and_compare(gpu_scenes, scene(world_runs, "rain_small"), path=path)


def test_rain_matches_reference(gpu_scenes, world_runs):
    """The 20,000-body mixed-shape rain, 1000 steps, hash-checked every 10 steps."""
    print("[gpu] rain 20k", flush=True)
    used = run_and_compare(gpu_scenes, scene(world_runs, "rain"))
    print(f"[gpu] rain pair query steps: host {used[1]} gpu {used[2]} speculative {used[3]} none {used[0]}", flush=True)


@pytest.mark.parametrize("path", [0, 1])
def test_falling_hinges_determinism(gpu_scenes, world_runs, path):
    """The reference determinism protocol (test/test_determinism.c:19-24, shared/determinism.c):
    step at 1/60 s with 4 substeps until no body moves; the scene sleeps at step 274 with
    transform hash 0xE86690F4."""
    import box2d_amd

    print(f"[gpu] falling hinges protocol path {path}", flush=True)
    h = gpu_scenes.create(1)
    box2d_amd.set_solver_path(gpu_scenes.world_id(h), path)
    try:
        sleep_step = -1
        for step in range(500):
            gpu_scenes.step(h)
            if gpu_scenes.move_count(h) == 0:
                sleep_step = step
                break
        fh = world_runs["falling_hinges"]
        assert sleep_step == fh["sleep_step"] == 274
        assert gpu_scenes.hash(h) == fh["hash"] == 0xE86690F4
    finally:
        gpu_scenes.destroy(h)


def test_interleaved_worlds_are_independent(gpu_scenes, world_runs):
    """Two worlds stepped alternately give the same trace as one world alone."""
    print("[gpu] interleaved worlds", flush=True)
    fx = scene(world_runs, "pyramid20")
    h1 = gpu_scenes.pyramid(20, 40.0)
    h2 = gpu_scenes.pyramid(20, 40.0)
    try:
        for s in range(1, 101):
            gpu_scenes.step(h1)
            gpu_scenes.step(h2)
            assert gpu_scenes.hash(h1) == fx["hashes"][str(s)]
            assert gpu_scenes.hash(h2) == fx["hashes"][str(s)]
    finally:
        gpu_scenes.destroy(h1)
        gpu_scenes.destroy(h2)


def test_world_recreate(gpu_scenes, world_runs):
    """Destroying and re-creating worlds reuses slots and device memory without stale state."""
    print("[gpu] recreate", flush=True)
    fx = scene(world_runs, "hello")
    for _ in range(3):
        h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
        gpu_scenes.step(h, 60)
        assert gpu_scenes.hash(h) == fx["hashes"]["60"]
        gpu_scenes.destroy(h)


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("name", ["edits", "edits_sleep"])
def test_edits_between_steps_match_reference(gpu_scenes, world_runs, name, path):
    """Body, shape and joint calls made between steps (transforms, forces and impulses, geometry
    and mass, hinge motor/spring/limit/frame changes, joint and body destruction, new bodies, a
    kinematic target) give the reference's transforms and getter results after every step."""
    import box2d_amd

    fx = next(e for e in world_runs["edits"] if e["name"] == name)
    print(f"[gpu] edits scene {name} path {path}", flush=True)
    h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    box2d_amd.set_solver_path(gpu_scenes.world_id(h), path)
    try:
        for step in range(1, fx["steps"] + 1):
            gpu_scenes.step(h)
            got = [gpu_scenes.hash(h), gpu_scenes.probe_hash(h), gpu_scenes.body_count(h)]
            assert got == fx["per_step"][step - 1], f"{name} step {step}: [transforms, getters, bodies] {got}, reference {fx['per_step'][step - 1]}"
    finally:
        gpu_scenes.destroy(h)


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("name", ["setters", "setters_sleep"])
def test_setters_and_callbacks_match_reference(gpu_scenes, world_runs, name, path):
    """Setters and user callbacks between steps: damping, gravity scale, shape materials,
    density and filters, body types (with joints attached), disable/enable, shape destruction,
    sleep settings, awake changes, the friction/restitution callbacks (set, swapped, cleared)
    evaluated on the host, the custom pair filter and the restitution and hit-event thresholds.
    Transforms and getter results (contact data included) equal the reference's after every
    step."""
    import box2d_amd

    fx = next(e for e in world_runs["edits"] if e["name"] == name)
    print(f"[gpu] setters scene {name} path {path}", flush=True)
    h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    box2d_amd.set_solver_path(gpu_scenes.world_id(h), path)
    try:
        for step in range(1, fx["steps"] + 1):
            gpu_scenes.step(h)
            got = [gpu_scenes.hash(h), gpu_scenes.probe_hash(h), gpu_scenes.body_count(h)]
            assert got == fx["per_step"][step - 1], f"{name} step {step}: [transforms, getters, bodies] {got}, reference {fx['per_step'][step - 1]}"
    finally:
        gpu_scenes.destroy(h)


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("name", ["presolve", "presolve_sleep"])
def test_presolve_callback_matches_reference(gpu_scenes, name, path):
    """Pre-solve callback on a one-way platform: the narrowphase of contacts with pre-solve
    events runs on the host around the callback, which drops points of contacts approached from
    below and second points of some pairs. Transforms, the callback's call count and argument
    tally, the platform's flag and its contact count equal the reference's after every step,
    through flag changes, relaunches and the callback cleared and restored
    (tests/golden/presolve.json)."""
    import box2d_amd
    from conftest import load_golden

    fx = next(s for s in load_golden("presolve.json")["scenes"] if s["name"] == name)
    print(f"[gpu] presolve scene {name} path {path}", flush=True)
    h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    box2d_amd.set_solver_path(gpu_scenes.world_id(h), path)
    try:
        for step in range(1, fx["steps"] + 1):
            gpu_scenes.step(h)
            got = [gpu_scenes.hash(h), gpu_scenes.probe_hash(h), gpu_scenes.body_count(h)]
            assert got == fx["per_step"][step - 1], f"{name} step {step}: [transforms, probe, bodies] {got}, reference {fx['per_step'][step - 1]}"
    finally:
        gpu_scenes.destroy(h)


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("name", ["sensors", "sensors_sleep"])
def test_sensor_events_match_reference(gpu_scenes, name, path):
    """Sensors: a static zone with a mask, a thin static strip crossed by projectiles within one
    step (continuous-collision sensor hits), a kinematic sensor sweeping and a sensor carried by
    a dynamic body, over visitors with and without sensor events, filtered by category and by the
    custom filter; sensor events toggled, a visitor and a sensor destroyed, the carrier disabled
    and enabled, a filter changed. Transforms, every sensor event and the sensors' overlap lists
    equal the reference's after each of 200 steps (tests/golden/presolve.json)."""
    import box2d_amd
    from conftest import load_golden

    fx = next(s for s in load_golden("presolve.json")["scenes"] if s["name"] == name)
    print(f"[gpu] sensor scene {name} path {path}", flush=True)
    h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    box2d_amd.set_solver_path(gpu_scenes.world_id(h), path)
    try:
        for step in range(1, fx["steps"] + 1):
            gpu_scenes.step(h)
            got = [gpu_scenes.hash(h), gpu_scenes.probe_hash(h), gpu_scenes.body_count(h)]
            assert got == fx["per_step"][step - 1], f"{name} step {step}: [transforms, sensors, bodies] {got}, reference {fx['per_step'][step - 1]}"
    finally:
        gpu_scenes.destroy(h)


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("name", ["bullets", "bullets_sleep", "bullets_precontinuous"])
def test_bullets_match_reference(gpu_scenes, name, path):
    """Bullets: five bullets and two fast non-bullet bodies fired at 200-260 m/s through a box
    stack, a thin dynamic plank, a rotating kinematic paddle and a dynamic body carrying a sensor
    ring. Bullets sweep against the moving bodies after the refit and stop at them; the other fast
    bodies sweep against static shapes only and pass. A fast body becomes a bullet and a bullet
    stops being one mid-run. With the pre-continuous callback every fast body sweeps on the host
    and the callback drops hits on the upper half of a ghost wall. Transforms, sensor events, the
    ring's overlaps, the bullet flags and the callback's call count and argument sum equal the
    reference's after each of 200 steps (tests/golden/presolve.json)."""
    import box2d_amd
    from conftest import load_golden

    fx = next(s for s in load_golden("presolve.json")["scenes"] if s["name"] == name)
    print(f"[gpu] bullet scene {name} path {path}", flush=True)
    h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    box2d_amd.set_solver_path(gpu_scenes.world_id(h), path)
    try:
        for step in range(1, fx["steps"] + 1):
            gpu_scenes.step(h)
            got = [gpu_scenes.hash(h), gpu_scenes.probe_hash(h), gpu_scenes.body_count(h)]
            assert got == fx["per_step"][step - 1], f"{name} step {step}: [transforms, probe, bodies] {got}, reference {fx['per_step'][step - 1]}"
    finally:
        gpu_scenes.destroy(h)
