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

CHECK = {1, 2, 3, 10, 30, 60, 90}


def _set_batch_threads(threads):
    import box2d_amd

    lib = box2d_amd.library()
    lib.b2amd_SetBatchSolverThreads.restype = ctypes.c_int
    assert lib.b2amd_SetBatchSolverThreads(threads) == 0


@pytest.mark.parametrize("threads,hi,seed", [(0, 300, 777), (64, 1024, 781), (256, 1024, 778), (512, 1024, 779),
                                             (1024, 1024, 780)],
                         ids=["tuned_8-300", "wg64_8-1024", "wg256_8-1024", "wg512_8-1024", "wg1024_8-1024"])
def test_batch_matches_single_worlds(gpu_scenes, threads, hi, seed):
    """48 uneven worlds (8 to 300 or 1024 bodies) as one batch, with the persistent solver's
    workgroup size from the offline-tuned table per world-size bucket or forced to 64, 256, 512 or
    1024 threads for every bucket."""
    import box2d_amd
    from box2d_amd.scenes_api import SceneLib

    specs = box2d_amd.batch_world_specs(48, lo=8, hi=hi, seed=seed)
    batch = [gpu_scenes.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    single = [gpu_scenes.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    ref = None
    ref_handles = []
    if os.path.exists(REF_SCENES):
        ref = SceneLib(REF_SCENES)
        ref_handles = [ref.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    ids = (box2d_amd.WorldId * len(batch))(*[gpu_scenes.world_id(h) for h in batch])
    _set_batch_threads(threads)
    try:
        print(f"[gpu] batch of {len(batch)} worlds (threads {threads}), bodies {[gpu_scenes.body_count(h) for h in batch]}",
              flush=True)
        for step in range(1, 91):
            stats = box2d_amd.step_worlds(ids)
            if threads != 0:
                assert stats.solveThreads == threads
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
        assert stats.solveThreads in (64, 128, 256, 512, 1024)
        assert sum(stats.bucketWorlds) == len(batch)
        if threads == 0:
            # one launch: the workgroup size of the bucket with the most worlds, each bucket's warm form
            table = box2d_amd.batch_tuning()
            most = max(range(4), key=lambda k: stats.bucketWorlds[k])
            for k in range(4):
                if stats.bucketWorlds[k]:
                    assert stats.bucketThreads[k] == table[most]["table_threads"]
                    assert stats.bucketWarmByBody[k] == table[k]["table_warm_by_body"]
    finally:
        _set_batch_threads(0)
        for h in batch + single:
            gpu_scenes.destroy(h)
        for h in ref_handles:
            ref.destroy(h)


@pytest.mark.parametrize("threads,warm", [(64, 1), (128, 0), (1024, 1)], ids=["wg64_warm_by_body", "wg128_by_color",
                                                                           "wg1024_warm_by_body"])
def test_batch_bucket_overrides_match_single_worlds(gpu_scenes, threads, warm):
    """Every bucket's configuration overridden (b2amd_SetBatchTuning) before the worlds are built:
    32 uneven worlds (8 to 2000 bodies, so all four buckets) stepped 40 times as a batch end as if
    stepped alone, with the warm start body by body or color by color."""
    import box2d_amd

    lib = box2d_amd.library()
    specs = box2d_amd.batch_world_specs(32, lo=8, hi=2000, seed=900 + threads + warm)
    for k in range(4):
        assert lib.b2amd_SetBatchTuning(k, threads, warm) == 0
    batch = [gpu_scenes.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    single = [gpu_scenes.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    ids = (box2d_amd.WorldId * len(batch))(*[gpu_scenes.world_id(h) for h in batch])
    try:
        print(f"[gpu] bucket overrides threads {threads} warm {warm}", flush=True)
        for step in range(1, 41):
            stats = box2d_amd.step_worlds(ids)
            for h in single:
                gpu_scenes.step(h)
            assert all(stats.bucketThreads[k] == threads for k in range(4) if stats.bucketWorlds[k])
            assert all(stats.bucketWarmByBody[k] == warm for k in range(4) if stats.bucketWorlds[k])
            if step in CHECK or step == 40:
                for i, (hb, hs) in enumerate(zip(batch, single)):
                    a, b = gpu_scenes.hash(hb), gpu_scenes.hash(hs)
                    assert a == b, f"world {i} {specs[i]}: batch {a:#x} != single {b:#x} at step {step}"
        assert sum(1 for k in range(4) if stats.bucketWorlds[k]) >= 3
    finally:
        for k in range(4):
            lib.b2amd_SetBatchTuning(k, 0, 0)
        for h in batch + single:
            gpu_scenes.destroy(h)


def test_batch_with_worlds_beyond_the_persistent_solver(gpu_scenes, world_runs):
    """Two 100-row pyramids (5050 bodies each) and a 105-row pyramid (5565 awake bodies, more than
    the persistent solver's 5400 LDS slots) batched with small worlds: the large world's substep
    loop runs on its own after the batch's, and every world still ends as if stepped alone (the
    100-row pyramid also as the reference's)."""
    import box2d_amd

    fx = next(s for s in world_runs["scenes"] if s["name"] == "pyramid100")
    specs = box2d_amd.batch_world_specs(6, lo=8, hi=300, seed=901)
    batch = [gpu_scenes.pyramid(100, fx["fparams"][0], sleep=False), gpu_scenes.pyramid(105, 40.0, sleep=False)]
    batch += [gpu_scenes.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    single = [gpu_scenes.pyramid(105, 40.0, sleep=False)] + [gpu_scenes.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    ids = (box2d_amd.WorldId * len(batch))(*[gpu_scenes.world_id(h) for h in batch])
    try:
        assert gpu_scenes.body_count(batch[1]) > 5400
        print(f"[gpu] batch with oversized worlds, bodies {[gpu_scenes.body_count(h) for h in batch]}", flush=True)
        for step in range(1, 41):
            stats = box2d_amd.step_worlds(ids)
            for h in single:
                gpu_scenes.step(h)
            want = fx["hashes"].get(str(step))
            if want is not None:
                assert gpu_scenes.hash(batch[0]) == want, f"100-row pyramid differs from the reference at step {step}"
            if step in CHECK or step == 40:
                for i, (hb, hs) in enumerate(zip(batch[1:], single)):
                    a, b = gpu_scenes.hash(hb), gpu_scenes.hash(hs)
                    assert a == b, f"world {i + 1}: batch {a:#x} != single {b:#x} at step {step}"
        assert stats.worlds == len(batch)
    finally:
        for h in batch + single:
            gpu_scenes.destroy(h)



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
