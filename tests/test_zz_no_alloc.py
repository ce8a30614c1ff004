"""b2amd_CollideBatchNoAlloc (the narrowphase batch staged in a fixed device array instead of
allocated buffers) gives the same manifolds as b2amd_CollideBatch, bit for bit, on the
reference's manifold fixture (tests/golden/manifolds.npz, written from the reference's
b2Collide* functions): two calls in a row, the second with fewer pairs, so the second reads
the staging array that the first left behind. Counts beyond the staging capacity are refused
on the host before any GPU work."""

import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from b2types import LocalManifold, Shape, Transform, manifold_array_words  # noqa: E402

CAPACITY = 4096


def _pairs(idx):
    d = np.load(os.path.join(ROOT, "tests", "golden", "manifolds.npz"))
    n = len(idx)
    a = (Shape * n).from_buffer_copy(d["shapes_a"][idx].tobytes())
    b = (Shape * n).from_buffer_copy(d["shapes_b"][idx].tobytes())
    x = (Transform * n).from_buffer_copy(d["transforms"][idx].tobytes())
    return a, b, x, d["manifolds"][idx]


def _lib():
    import box2d_amd

    lib = box2d_amd.library()
    lib.b2amd_CollideBatchNoAlloc.restype = ctypes.c_int
    lib.b2amd_CollideBatch.restype = ctypes.c_int
    return lib


def test_no_alloc_refuses_counts_beyond_capacity():
    lib = _lib()
    n = CAPACITY + 1
    a, b, x = (Shape * n)(), (Shape * n)(), (Transform * n)()
    out = (LocalManifold * n)()
    assert lib.b2amd_CollideBatchNoAlloc(a, b, x, n, out) == -2
    assert lib.b2amd_CollideBatchNoAlloc(a, b, x, 0, out) == 0


@pytest.mark.gpu
def test_no_alloc_matches_collide_batch_twice():
    lib = _lib()
    total = np.load(os.path.join(ROOT, "tests", "golden", "manifolds.npz"))["manifolds"].shape[0]
    rng = np.random.default_rng(7)
    first = rng.choice(total, CAPACITY, replace=False)
    second = rng.choice(total, 1000, replace=False)
    for label, idx in (("first call, 4096 pairs", first), ("second call, 1000 pairs", second)):
        print(f"b2amd_CollideBatchNoAlloc: {label}", flush=True)
        a, b, x, want = _pairs(idx)
        n = len(idx)
        got = (LocalManifold * n)()
        ref = (LocalManifold * n)()
        assert lib.b2amd_CollideBatchNoAlloc(a, b, x, n, got) == 0
        assert lib.b2amd_CollideBatch(a, b, x, n, ref) == 0
        got_w = manifold_array_words(got)
        assert np.array_equal(got_w, manifold_array_words(ref)), label
        assert np.array_equal(got_w, want), label
