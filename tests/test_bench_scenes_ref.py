"""CPU checks of the benchmark-scene fixture and its driver.

The fixture tests/golden/bench_scenes.json was written from the reference library by
tests/golden/gen_bench_scenes.py. Where oracle/_ref/libb2ref_bench.so is built (it needs the
reference sources), the first steps of every scene are run again and must give the fixture's
numbers, so the fixture and the driver agree.
"""

import os

import pytest

from oracle.bench_scenes import REF_BENCH, BenchScenes

SCENE_STEPS = {"compounds": 500, "joint_grid": 500, "junkyard": 800, "large_pyramid": 500, "many_pyramids": 200,
               "rain": 1000, "smash": 300, "spinner": 500, "tumbler": 750, "washer": 500}


def test_fixture_covers_the_reference_benchmark_table(bench_scenes_golden):
    # step counts of the reference benchmark app (benchmark/main.c:150-159)
    assert {n: s["steps"] for n, s in bench_scenes_golden.items()} == SCENE_STEPS
    for name, s in bench_scenes_golden.items():
        assert len(s["per_step"]) == s["steps"]
        # a step reports a move event for every body that was awake in it, so never fewer
        # than the bodies still awake after it
        assert all(len(r) == 3 and r[0] >= r[1] for r in s["per_step"])
        # every scene has awake bodies moving at the start
        assert s["per_step"][0][0] > 0


@pytest.mark.skipif(not os.path.exists(REF_BENCH), reason="oracle/_ref/libb2ref_bench.so not built")
@pytest.mark.parametrize("name", sorted(SCENE_STEPS))
def test_reference_library_reproduces_fixture(bench_scenes_golden, name):
    lib = BenchScenes(REF_BENCH)
    lib.create(name)
    try:
        for step in range(1, 9):
            assert list(lib.step()) == bench_scenes_golden[name]["per_step"][step - 1]
    finally:
        lib.destroy()
