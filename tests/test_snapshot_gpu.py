"""World snapshots: b2World_Snapshot, b2World_Restore and b2CreateWorldFromSnapshot.

The image layout is this library's own, so the reference cannot read it; what the reference pins
is the state. A world is stepped k steps, snapshotted, stepped m more steps, restored in place and
stepped the same m steps again, and a second world is created from the image and stepped m steps.
After every one of those steps b2World_GetStateHash (the reference's deep hash over transforms,
velocities, contact and joint impulses, every set/color/local index and the id pools) must equal
the reference's at the same step (tests/golden/state_hashes.json). A corrupted image is rejected
and leaves the world as it was.
"""

import ctypes

import pytest

from box2d_amd.scenes_api import WorldId
from conftest import load_golden

CASES = [("pyramid20_sleep", 30, 60), ("joint_zoo_sleep", 150, 80), ("mixed_pile_1", 90, 60), ("events_60", 100, 50)]


def fixture(name):
    return next(s for s in load_golden("state_hashes.json")["scenes"] if s["name"] == name)


def bind(lib):
    lib.b2World_Snapshot.restype = ctypes.c_int
    lib.b2World_Snapshot.argtypes = [WorldId, ctypes.c_void_p, ctypes.c_int]
    lib.b2World_Restore.restype = ctypes.c_bool
    lib.b2World_Restore.argtypes = [WorldId, ctypes.c_void_p, ctypes.c_int]
    lib.b2CreateWorldFromSnapshot.restype = WorldId
    lib.b2CreateWorldFromSnapshot.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.b2World_GetStateHash.restype = ctypes.c_uint64
    lib.b2World_GetStateHash.argtypes = [WorldId]
    lib.b2World_Step.restype = None
    lib.b2World_Step.argtypes = [WorldId, ctypes.c_float, ctypes.c_int]
    lib.b2DestroyWorld.argtypes = [WorldId]
    return lib


def test_cases_have_reference_hashes():
    for name, k, m in CASES:
        fx = fixture(name)
        assert fx["every"] == 1 and k + m <= fx["steps"]
        # the window changes the state (the pyramid falls asleep at step 55 within its window): a
        # restore that did nothing would not match
        assert len({fx["hashes"][str(s)] for s in range(k, k + m + 1)}) > 1


@pytest.mark.gpu
@pytest.mark.parametrize("name,k,m", CASES)
def test_snapshot_restore_and_create_match_reference(gpu_scenes, name, k, m):
    import box2d_amd

    lib = bind(box2d_amd.library())
    fx = fixture(name)
    want = fx["hashes"]
    print(f"[gpu] snapshot {name} at step {k}, {m} steps", flush=True)
    h = gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"])
    wid = gpu_scenes.world_id(h)
    try:
        gpu_scenes.step(h, k)
        assert lib.b2World_GetStateHash(wid) == want[str(k)]

        size = lib.b2World_Snapshot(wid, None, 0)
        assert size > 64
        small = (ctypes.c_uint8 * 16)()
        assert lib.b2World_Snapshot(wid, small, 16) == size and bytes(small) == bytes(16)
        image = (ctypes.c_uint8 * size)()
        assert lib.b2World_Snapshot(wid, image, size) == size
        assert lib.b2World_GetStateHash(wid) == want[str(k)]

        def step_and_check(world, label):
            for s in range(k + 1, k + m + 1):
                lib.b2World_Step(world, ctypes.c_float(1.0 / 60.0), 4)
                got = lib.b2World_GetStateHash(world)
                assert got == want[str(s)], f"{name} {label} step {s}: {got:#018x}, reference {want[str(s)]:#018x}"

        step_and_check(wid, "original")

        # a corrupted image is rejected and the world is left as it was
        bad = (ctypes.c_uint8 * size).from_buffer_copy(image)
        bad[size // 2] ^= 0x5A
        assert lib.b2World_Restore(wid, bad, size) is False
        assert lib.b2World_GetStateHash(wid) == want[str(k + m)]

        assert lib.b2World_Restore(wid, image, size) is True
        assert lib.b2World_GetStateHash(wid) == want[str(k)]
        step_and_check(wid, "restored")

        other = lib.b2CreateWorldFromSnapshot(image, size, 1)
        assert other.index1 != 0
        try:
            assert lib.b2World_GetStateHash(other) == want[str(k)]
            step_and_check(other, "created")
        finally:
            lib.b2DestroyWorld(other)
    finally:
        gpu_scenes.destroy(h)
