"""The C-ABI library loads and exports every function the public headers declare (no GPU calls)."""

import ctypes
import os
import re
import subprocess

import pytest

from conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")
HEADERS = [os.path.join(INCLUDE, "box2d", "box2d.h"), os.path.join(INCLUDE, "box2d", "collision.h"),
           os.path.join(INCLUDE, "box2d", "math_functions.h"), os.path.join(INCLUDE, "box2d", "types.h"),
           os.path.join(INCLUDE, "box2d", "base.h"), os.path.join(INCLUDE, "box2d_amd.h")]

_DECL = re.compile(r"B2_API\s+[^;{]*?\b([a-z]\w*)\s*\(", re.S)


def declared_symbols():
    names = set()
    for path in HEADERS:
        with open(path) as f:
            text = f.read()
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"(?m)^\s*#.*$", "", text)
        for m in _DECL.finditer(text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_step_api():
    names = declared_symbols()
    for required in ("b2CreateWorld", "b2DestroyWorld", "b2World_Step", "b2CreateBody", "b2CreatePolygonShape",
                     "b2CreateRevoluteJoint", "b2Body_GetTransform", "b2World_GetBodyEvents", "b2CollidePolygons",
                     "b2amd_World_GetKernelStats", "b2amd_CollideBatch"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import box2d_amd

    lib = box2d_amd.library()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert missing == [], f"declared but not exported: {missing}"


def test_scene_driver_loads():
    import box2d_amd

    s = box2d_amd.scenes()
    for name in ("b2s_create", "b2s_step", "b2s_hash", "b2s_world_id", "b2s_counters", "b2s_falling_hinges"):
        assert hasattr(s.lib, name)


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_headers_compile(tmp_path, lang):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('#include "box2d/box2d.h"\n#include "box2d/math_functions.h"\n#include "box2d_amd.h"\n'
                   "int main(void){ b2WorldDef d = b2DefaultWorldDef(); (void)d; return 0; }\n")
    cc = "gcc" if lang == "c" else "g++"
    std = "-std=c17" if lang == "c" else "-std=c++17"
    subprocess.run([cc, std, "-fsyntax-only", "-Wall", "-I", INCLUDE, str(src)], check=True)


def test_layouts_match_reference_sizes():
    """Structs the caller fills are passed by pointer: their sizes must match the reference ABI."""
    import box2d_amd

    lib = box2d_amd.library()

    class WorldId(ctypes.Structure):
        _fields_ = [("index1", ctypes.c_uint16), ("generation", ctypes.c_uint16)]

    class WorldDef(ctypes.Structure):
        _fields_ = [("raw", ctypes.c_ubyte * 108), ("internalValue", ctypes.c_int)]

    lib.b2DefaultWorldDef.restype = WorldDef
    lib.b2DefaultWorldDef.argtypes = []
    d = lib.b2DefaultWorldDef()
    # internalValue (the def cookie) is the last int of the 112-byte b2WorldDef
    cookie = d.internalValue
    assert cookie == 1152023
    assert ctypes.sizeof(WorldId) == 4


def test_batch_tuning_table_is_well_formed():
    """The offline-tuned table of the batched solver (host calls only): four buckets with rising
    body limits, the last one reaching the persistent solver's 5400-body capacity, each with a valid
    workgroup size and warm-start form; overrides are validated and restored."""
    import box2d_amd

    lib = box2d_amd.library()
    table = box2d_amd.batch_tuning()
    assert [t["bucket"] for t in table] == [0, 1, 2, 3]
    limits = [t["max_bodies"] for t in table]
    assert limits == sorted(limits) and limits[-1] >= 5400
    for t in table:
        assert t["table_threads"] in (64, 128, 256, 512, 1024)
        assert t["table_warm_by_body"] in (0, 1)
        assert (t["threads"], t["warm_by_body"]) == (t["table_threads"], t["table_warm_by_body"])
    assert lib.b2amd_SetBatchTuning(4, 256, 0) == -1
    assert lib.b2amd_SetBatchTuning(0, 96, 0) == -1
    assert lib.b2amd_SetBatchTuning(0, 256, 2) == -1
    assert lib.b2amd_SetBatchTuning(1, 128, 1) == 0
    assert box2d_amd.batch_tuning()[1]["threads"] == 128 and box2d_amd.batch_tuning()[1]["warm_by_body"] == 1
    assert lib.b2amd_SetBatchTuning(1, 0, 0) == 0
    assert box2d_amd.batch_tuning()[1]["threads"] == table[1]["table_threads"]
