"""The reference's own benchmark scenes, stepped by box2d_amd, against the reference library.

oracle/_ref/libb2amd_bench.so is the reference's shared/benchmarks.c (with shared/human.c and
shared/utils.c) compiled unchanged against box2d_amd's public headers and linked with its
library, driven as the reference's benchmark app drives it (benchmark/main.c:318-351). The
fixture tests/golden/bench_scenes.json (tests/golden/gen_bench_scenes.py) holds what the
reference library gives for the same scenes: per step, the move-event count, the awake body
count and a hash of every moved body's id and transform bits. Parity is exact at every step of
every scene: compounds, joint grid, junkyard (kinematic pusher driven by
b2Body_SetTargetTransform), large pyramid, many pyramids, rain (ragdolls created and destroyed
while stepping), smash, spinner (chain ground, motorized revolute joint), tumbler and washer
(kinematic containers).
"""

import os
import time

import numpy as np
import pytest

from oracle.bench_scenes import AMD_BENCH, BenchScenes

pytestmark = pytest.mark.gpu

SCENES = ["compounds", "joint_grid", "junkyard", "large_pyramid", "many_pyramids", "rain", "smash", "spinner",
          "tumbler", "washer"]


@pytest.fixture(scope="module")
def amd_bench():
    if not os.path.exists(AMD_BENCH):
        pytest.skip("oracle/_ref/libb2amd_bench.so not built (needs the reference's shared/ sources)")
    import box2d_amd

    box2d_amd.library()
    return BenchScenes(AMD_BENCH)


@pytest.mark.parametrize("name", SCENES)
def test_reference_benchmark_scene(amd_bench, bench_scenes_golden, name):
    fx = bench_scenes_golden[name]
    print(f"[gpu] reference benchmark scene {name}, {fx['steps']} steps", flush=True)
    amd_bench.create(name)
    start = time.perf_counter()
    try:
        for step in range(1, fx["steps"] + 1):
            got = list(amd_bench.step())
            if step == 1:
                first = np.array(fx["first_moves"], dtype=np.uint32).reshape(-1, 7)
                moves = amd_bench.moves()[: len(first)]
                assert np.array_equal(moves, first), f"{name} step 1 move records differ"
            want = fx["per_step"][step - 1]
            assert got == want, f"{name} step {step}: [moves, awake, hash] {got}, reference {want}"
        ms = 1e3 * (time.perf_counter() - start) / fx["steps"]
        print(f"[gpu] {name}: {fx['steps']} steps match, {ms:.3f} ms/step with the per-step hash", flush=True)
    finally:
        amd_bench.destroy()
