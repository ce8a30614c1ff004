"""Batched stepping (b2amd_StepWorlds) against the same worlds stepped one by one.

Uneven worlds (box pyramids, random box/circle piles, hanging capsule chains on revolute joints,
8-300 bodies, log-uniform) are stepped together in one batch: one launch per GPU stage over all
worlds, the substep loop as a persistent launch over a world queue. Every world's transform hash
must equal the hash of the same world stepped alone through b2World_Step on the GPU and, where the
reference library is present, through the reference.
"""

import ctypes
import os

import pytest

from conftest import REF_SCENES

pytestmark = pytest.mark.gpu

CHECK = {1, 2, 10, 30, 60, 90}


def test_batch_matches_single_worlds(gpu_scenes):
    import box2d_amd
    from box2d_amd.scenes_api import SceneLib

    specs = box2d_amd.batch_world_specs(48, lo=8, hi=300, seed=777)
    batch = [gpu_scenes.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    single = [gpu_scenes.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    ref = None
    ref_handles = []
    if os.path.exists(REF_SCENES):
        ref = SceneLib(REF_SCENES)
        ref_handles = [ref.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    ids = (box2d_amd.WorldId * len(batch))(*[gpu_scenes.world_id(h) for h in batch])
    try:
        print(f"[gpu] batch of {len(batch)} worlds, bodies {[gpu_scenes.body_count(h) for h in batch]}", flush=True)
        for step in range(1, 91):
            stats = box2d_amd.step_worlds(ids)
            for h in single:
                gpu_scenes.step(h)
            for h in ref_handles:
                ref.step(h)
            if step in CHECK:
                for i, (hb, hs) in enumerate(zip(batch, single)):
                    a, b = gpu_scenes.hash(hb), gpu_scenes.hash(hs)
                    assert a == b, f"world {i} {specs[i]}: batch {a:#x} != single {b:#x} at step {step}"
                    if ref is not None:
                        c = ref.hash(ref_handles[i])
                        assert a == c, f"world {i} {specs[i]}: batch {a:#x} != reference {c:#x} at step {step}"
        assert stats.worlds == len(batch)
        assert stats.solveThreads in (256, 512, 1024)
    finally:
        for h in batch + single:
            gpu_scenes.destroy(h)
        for h in ref_handles:
            ref.destroy(h)



def test_batch_of_joint_worlds_matches_reference(gpu_scenes, world_runs):
    """Worlds holding every joint type, stepped as one batch (persistent fused solver): each
    world's hash after every step equals the reference's."""
    import box2d_amd

    fx = next(s for s in world_runs["scenes"] if s["name"] == "joint_zoo")
    handles = [gpu_scenes.create(fx["kind"], fx["iparams"], fx["fparams"]) for _ in range(6)]
    ids = (box2d_amd.WorldId * len(handles))(*[gpu_scenes.world_id(h) for h in handles])
    try:
        print(f"[gpu] batch of {len(handles)} joint worlds", flush=True)
        for step in range(1, fx["steps"] + 1):
            box2d_amd.step_worlds(ids)
            want = fx["hashes"][str(step)]
            for i, h in enumerate(handles):
                got = gpu_scenes.hash(h)
                assert got == want, f"joint world {i}: {got:#x} != reference {want:#x} at step {step}"
    finally:
        for h in handles:
            gpu_scenes.destroy(h)
