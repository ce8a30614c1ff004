import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
REF_SCENES = os.path.join(ROOT, "oracle", "_ref", "libb2ref_scenes.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950) and the built HIP library")


def load_golden(name):
    with open(os.path.join(GOLDEN_DIR, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def world_runs():
    return load_golden("world_runs.json")


@pytest.fixture(scope="session")
def bench_scenes_golden():
    return {s["name"]: s for s in load_golden("bench_scenes.json")["scenes"]}


@pytest.fixture(scope="session")
def gpu_scenes():
    import box2d_amd

    return box2d_amd.scenes()


@pytest.fixture(scope="session")
def ref_scenes():
    """The reference library compiled from its own sources (present only where they were)."""
    if not os.path.exists(REF_SCENES):
        pytest.skip("oracle/_ref/libb2ref_scenes.so not built (needs the reference sources)")
    from box2d_amd.scenes_api import SceneLib

    return SceneLib(REF_SCENES)
