"""Time of impact and GJK distance against the reference fixture, on the host and on the GPU.

tests/golden/toi_distance.npz (tests/golden/gen_toi.py) holds 4000 time-of-impact queries (half
general, half continuous-collision-like: a static chain segment far from the sweep's frame origin
against a falling convex shape) and 4000 distance queries with the reference's outputs. The
comparison is bit-exact over every output field (the distance's debug simplex count excepted).
"""

import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR
from b2types import DistanceInput, DistanceOutput, SimplexCache, TOIInput, TOIOutput, bind_distance_functions


def load():
    d = np.load(os.path.join(GOLDEN_DIR, "toi_distance.npz"))
    n = d["toi_inputs"].shape[0]
    toi_in = (TOIInput * n).from_buffer_copy(d["toi_inputs"].tobytes())
    dist_in = (DistanceInput * n).from_buffer_copy(d["dist_inputs"].tobytes())
    return n, toi_in, d["toi_outputs"], dist_in, d["dist_outputs"], d["dist_caches"]


def distance_words(raw):
    """pointA, pointB, normal, distance, iterations (drop the simplex count)."""
    return np.frombuffer(raw, dtype=np.uint32).reshape(-1, 9)[:, :8]


def check_toi(got_bytes, want):
    got = np.frombuffer(got_bytes, dtype=np.uint32).reshape(want.shape[0], -1)
    want = want.view(np.uint32).reshape(want.shape[0], -1)
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} TOI outputs differ, first {int(bad[0])}: {got[bad[0]].tolist()} vs {want[bad[0]].tolist()}"


def test_host_toi_and_distance_match_fixture():
    import box2d_amd

    lib = bind_distance_functions(box2d_amd.library())
    n, toi_in, toi_out, dist_in, dist_out, dist_caches = load()
    out = (TOIOutput * n)()
    for i in range(n):
        out[i] = lib.b2TimeOfImpact(ctypes.byref(toi_in[i]))
    check_toi(bytes(out), toi_out)

    douts = (DistanceOutput * n)()
    caches = (SimplexCache * n)()
    for i in range(n):
        douts[i] = lib.b2ShapeDistance(ctypes.byref(dist_in[i]), ctypes.byref(caches[i]), None, 0)
    assert np.array_equal(distance_words(bytes(douts)), distance_words(dist_out.tobytes()))
    assert bytes(caches) == dist_caches.tobytes()


@pytest.mark.gpu
def test_gpu_toi_and_distance_match_fixture():
    import box2d_amd

    lib = box2d_amd.library()
    n, toi_in, toi_out, dist_in, dist_out, dist_caches = load()
    print("[gpu] toi batch", flush=True)
    out = (TOIOutput * n)()
    assert lib.b2amd_TimeOfImpactBatch(toi_in, n, out) == 0
    check_toi(bytes(out), toi_out)

    print("[gpu] distance batch", flush=True)
    douts = (DistanceOutput * n)()
    caches = (SimplexCache * n)()
    assert lib.b2amd_ShapeDistanceBatch(dist_in, caches, n, douts) == 0
    assert np.array_equal(distance_words(bytes(douts)), distance_words(dist_out.tobytes()))
    assert bytes(caches) == dist_caches.tobytes()
