"""Narrowphase parity: the manifold functions against the reference fixture.

The fixture (tests/golden/manifolds.npz, written by tests/golden/gen_manifolds.py from the
reference's src/manifold.c) holds 600 random pairs of each primary pair type. The comparison is
bit-exact over every word of b2LocalManifold: normal, points, separations, feature ids, count.
"""

import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN_DIR, REF_SCENES
from b2types import (PAIR_TYPES, LocalManifold, Shape, Transform, bind_manifold_functions, collide_host,
                     manifold_array_words, manifold_words)


def load_fixture():
    d = np.load(os.path.join(GOLDEN_DIR, "manifolds.npz"))
    n = d["manifolds"].shape[0]
    a = (Shape * n).from_buffer_copy(d["shapes_a"].tobytes())
    b = (Shape * n).from_buffer_copy(d["shapes_b"].tobytes())
    x = (Transform * n).from_buffer_copy(d["transforms"].tobytes())
    return a, b, x, d["manifolds"]


def first_mismatch(got, want, shapes_a, shapes_b):
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    if len(bad) == 0:
        return None
    i = int(bad[0])
    return (f"{len(bad)} of {len(want)} manifolds differ; first at pair {i} "
            f"(types {shapes_a[i].type},{shapes_b[i].type}): got {got[i].tolist()} want {want[i].tolist()}")


def test_fixture_covers_every_pair_type():
    a, b, _, m = load_fixture()
    types = {(a[i].type, b[i].type) for i in range(len(a))}
    assert types == set(PAIR_TYPES)
    counts = np.bincount(m[:, 10].astype(np.int64), minlength=3)
    assert counts[1] > 200 and counts[2] > 200


def test_host_manifolds_match_reference_fixture():
    """libbox2d_amd's host b2Collide* (the source the collide kernel compiles per lane)."""
    import box2d_amd

    lib = bind_manifold_functions(box2d_amd.library())
    a, b, x, want = load_fixture()
    got = np.stack([manifold_words(collide_host(lib, a[i], b[i], x[i])) for i in range(len(a))])
    msg = first_mismatch(got, want, a, b)
    assert msg is None, msg


def test_reference_library_reproduces_fixture():
    """The fixture is what the reference computes (guards against a stale fixture)."""
    if not os.path.exists(REF_SCENES):
        pytest.skip("reference library not built here")
    lib = bind_manifold_functions(ctypes.CDLL(REF_SCENES))
    a, b, x, want = load_fixture()
    got = np.stack([manifold_words(collide_host(lib, a[i], b[i], x[i])) for i in range(0, len(a), 7)])
    assert np.array_equal(got, want[::7])


@pytest.mark.gpu
def test_gpu_collide_batch_matches_reference_fixture():
    import box2d_amd

    lib = box2d_amd.library()
    a, b, x, want = load_fixture()
    n = len(a)
    out = (LocalManifold * n)()
    lib.b2amd_CollideBatch.restype = ctypes.c_int
    rc = lib.b2amd_CollideBatch(a, b, x, n, out)
    assert rc == 0
    got = manifold_array_words(out)
    msg = first_mismatch(got, want, a, b)
    assert msg is None, msg
