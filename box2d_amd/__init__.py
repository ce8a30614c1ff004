"""box2d_amd: Box2D v3's world step on AMD Instinct MI355X (gfx950).

The native library ``lib/libbox2d_amd.so`` exports the Box2D C API (see include/box2d/box2d.h);
``b2World_Step`` runs its arithmetic in hand-written HIP kernels. This package only locates the
built libraries and wraps the scene driver for tests and benchmarks.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIBRARY_PATH = os.path.join(LIB_DIR, "libbox2d_amd.so")
SCENES_PATH = os.path.join(LIB_DIR, "libb2amd_scenes.so")

_lib = None


def library():
    """Load libbox2d_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise RuntimeError(
                f"{LIBRARY_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        _lib = ctypes.CDLL(LIBRARY_PATH)
    return _lib


from .scenes_api import (SceneLib, SCENE_PYRAMID, SCENE_FALLING_HINGES, SCENE_HELLO, SCENE_MIXED_PILE,  # noqa: E402
                         SCENE_RAIN, SCENE_BATCH_WORLD, SCENE_PUSH, WorldId)


class KernelStats(ctypes.Structure):
    """b2amd_KernelStats (include/box2d_amd.h): one kernel class's launches, time and bytes."""
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_int), ("totalMs", ctypes.c_double),
                ("algorithmicBytes", ctypes.c_double)]


class BatchStats(ctypes.Structure):
    _fields_ = [("solveKernelMs", ctypes.c_float), ("solveThreads", ctypes.c_int), ("solveGrid", ctypes.c_int),
                ("worlds", ctypes.c_int), ("phaseMs", ctypes.c_float * 5), ("solveAlgorithmicBytes", ctypes.c_double),
                ("bucketWorlds", ctypes.c_int * 4), ("bucketThreads", ctypes.c_int * 4),
                ("bucketWarmByBody", ctypes.c_int * 4), ("bucketWorkgroups", ctypes.c_int * 4),
                ("bucketMs", ctypes.c_float * 4)]


def batch_tuning():
    """The batched solver's configuration per world-size bucket: a list of dicts with the bucket's
    largest awake body count, the workgroup size and warm-start form in force and the offline
    table's."""
    lib = library()
    out = []
    for k in range(4):
        v = [ctypes.c_int() for _ in range(5)]
        if lib.b2amd_GetBatchTuning(k, *[ctypes.byref(x) for x in v]) != 0:
            break
        out.append({"bucket": k, "max_bodies": v[0].value, "threads": v[1].value, "warm_by_body": v[2].value,
                    "table_threads": v[3].value, "table_warm_by_body": v[4].value})
    return out


def step_worlds(world_ids, count=1, dt=1.0 / 60.0, substeps=4):
    """Step the worlds (a ctypes array of WorldId) together count times with b2amd_StepWorlds.
    Returns the BatchStats of the last step."""
    lib = library()
    stats = BatchStats()
    for _ in range(count):
        lib.b2amd_StepWorlds(world_ids, len(world_ids), ctypes.c_float(dt), substeps, ctypes.byref(stats))
    return stats


def batch_world_specs(count, lo=8, hi=1024, seed=12345):
    """Sizes log-uniform in [lo, hi], a world kind per world (0 pyramid, 1 pile, 2 chains) and a
    seed per world, all from one seeded generator."""
    import numpy as np

    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)
    kinds = rng.integers(0, 3, count)
    return [(int(s), int(k), int(seed + 1 + i)) for i, (s, k) in enumerate(zip(sizes, kinds))]


SOLVER_AUTO = 0
SOLVER_COLOR_KERNELS = 1
SOLVER_FUSED = 2
SOLVER_XCD = 3


def set_solver_path(world_id, path):
    """Select the solver path of a world (SOLVER_AUTO, SOLVER_COLOR_KERNELS, SOLVER_FUSED or
    SOLVER_XCD)."""
    library().b2amd_World_SetSolverPath(world_id, ctypes.c_int(path))


def scenes():
    """Scene driver bound to the GPU library."""
    library()
    return SceneLib(SCENES_PATH)
