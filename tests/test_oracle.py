"""Pin the CPU restatement (oracle/b2_oracle.c) before it is trusted as a checker.

Each function is checked against what does not depend on it: the reference fixture (outputs of
the reference's own functions), the reference library itself when it was built here, the
known answers of the reference's unit tests, and mathematics (libm, float64 brute force).
"""

import ctypes
import math
import os

import numpy as np
import pytest

from conftest import REF_SCENES, ROOT
from b2types import (PAIR_TYPES, LocalManifold, Polygon, SegmentDistanceResult, Shape, Transform, Vec2, Rot,
                     bind_manifold_functions, collide_host, manifold_array_words, manifold_words)
from test_manifolds import load_fixture

ORACLE = os.path.join(ROOT, "oracle", "_build", "libb2oracle.so")


class Soft(ctypes.Structure):
    _fields_ = [("biasRate", ctypes.c_float), ("massScale", ctypes.c_float), ("impulseScale", ctypes.c_float)]


@pytest.fixture(scope="module")
def oracle():
    if not os.path.exists(ORACLE):
        pytest.skip("oracle not built (make -C oracle oracle)")
    lib = ctypes.CDLL(ORACLE)
    lib.o_atan2.restype = ctypes.c_float
    lib.o_atan2.argtypes = [ctypes.c_float, ctypes.c_float]
    lib.o_make_soft.restype = Soft
    lib.o_make_soft.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
    lib.o_segment_distance.restype = SegmentDistanceResult
    lib.o_segment_distance.argtypes = [Vec2, Vec2, Vec2, Vec2]
    lib.o_collide_polygons.restype = LocalManifold
    lib.o_collide_polygons.argtypes = [ctypes.POINTER(Polygon), ctypes.POINTER(Polygon), Transform]
    return lib


def fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def test_oracle_manifolds_match_reference_fixture(oracle):
    a, b, x, want = load_fixture()
    n = len(a)
    out = (LocalManifold * n)()
    assert oracle.b2o_collide_batch(a, b, x, n, out) == 0
    got = manifold_array_words(out)
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} differ, first {int(bad[0])}: {got[bad[0]].tolist()} vs {want[bad[0]].tolist()}"


def test_oracle_manifold_invariants(oracle):
    """Unit normals, at most 2 points, feature ids inside the shapes' vertex ranges."""
    a, b, x, _ = load_fixture()
    n = len(a)
    out = (LocalManifold * n)()
    oracle.b2o_collide_batch(a, b, x, n, out)
    for i in range(n):
        m = out[i]
        assert 0 <= m.pointCount <= 2
        if m.pointCount == 0:
            continue
        assert abs(math.hypot(m.normal.x, m.normal.y) - 1.0) < 1e-5
        for k in range(m.pointCount):
            pid = m.points[k].id
            ia, ib = pid >> 8, pid & 0xFF
            ca = a[i].polygon.count if a[i].type == 3 else 2
            cb = b[i].polygon.count if b[i].type == 3 else 2
            assert ia < ca and ib < cb
        # the closest point of a manifold lies within the speculative margin
        assert min(m.points[k].separation for k in range(m.pointCount)) <= 4 * 0.005 + 1e-6


def test_box_overlap_known_answer(oracle):
    """test/test_collision.c:202-216: two unit boxes offset by 0.9 overlap by 0.1 with 2 points."""
    box = Polygon()
    verts = [(-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)]
    norms = [(0.0, -1.0), (1.0, 0.0), (0.0, 1.0), (-1.0, 0.0)]
    for i in range(4):
        box.vertices[i] = Vec2(*verts[i])
        box.normals[i] = Vec2(*norms[i])
    box.count = 4
    box.radius = 0.0
    xf = Transform(Vec2(0.9, 0.0), Rot(1.0, 0.0))
    m = oracle.o_collide_polygons(ctypes.byref(box), ctypes.byref(box), xf)
    assert m.pointCount == 2
    for k in range(2):
        assert abs(m.points[k].separation + 0.1) < 1e-6
    assert (m.normal.x, m.normal.y) == (1.0, 0.0)


def test_segment_distance_known_answers(oracle):
    """test/test_distance.c:12-43 and a float64 brute-force minimum."""
    r = oracle.o_segment_distance(Vec2(-1, -1), Vec2(-1, 1), Vec2(2, 0), Vec2(1, 0))
    assert abs(r.distanceSquared - 4.0) < 1e-6
    assert (r.fraction1, r.fraction2) == (0.5, 1.0)
    rng = np.random.default_rng(5)
    for _ in range(200):
        p1, q1, p2, q2 = (rng.uniform(-2, 2, 2) for _ in range(4))
        r = oracle.o_segment_distance(Vec2(*p1), Vec2(*q1), Vec2(*p2), Vec2(*q2))
        t = np.linspace(0, 1, 401)
        A = p1[None, :] + t[:, None] * (q1 - p1)[None, :]
        B = p2[None, :] + t[:, None] * (q2 - p2)[None, :]
        d2 = ((A[:, None, :] - B[None, :, :]) ** 2).sum(-1).min()
        assert r.distanceSquared <= d2 + 1e-6
        assert r.distanceSquared >= d2 - 0.05 * math.sqrt(d2) - 1e-4


def test_atan2_against_libm_and_reference(oracle):
    rng = np.random.default_rng(7)
    y = rng.uniform(-10, 10, 20000).astype(np.float32)
    x = rng.uniform(-10, 10, 20000).astype(np.float32)
    out = np.zeros_like(x)
    oracle.b2o_atan2_batch(fptr(y), fptr(x), fptr(out), len(x))
    # the reference's polynomial has max error about 1e-4 rad (src/math_functions.c:94)
    assert np.max(np.abs(out - np.arctan2(y, x))) < 2e-4
    assert oracle.o_atan2(0.0, 0.0) == 0.0
    if os.path.exists(REF_SCENES):
        ref = ctypes.CDLL(REF_SCENES)
        ref.b2Atan2.restype = ctypes.c_float
        ref.b2Atan2.argtypes = [ctypes.c_float, ctypes.c_float]
        for i in range(0, len(x), 37):
            assert ref.b2Atan2(float(y[i]), float(x[i])) == out[i]


def test_cos_sin_against_libm_and_reference(oracle):
    a = np.linspace(-20, 20, 20001).astype(np.float32)
    c = np.zeros_like(a)
    s = np.zeros_like(a)
    oracle.b2o_cos_sin_batch(fptr(a), fptr(c), fptr(s), len(a))
    assert np.max(np.abs(c - np.cos(a.astype(np.float64)))) < 2e-3
    assert np.max(np.abs(s - np.sin(a.astype(np.float64)))) < 2e-3
    assert np.max(np.abs(c * c + s * s - 1.0)) < 1e-6
    if os.path.exists(REF_SCENES):
        ref = ctypes.CDLL(REF_SCENES)

        class CosSin(ctypes.Structure):
            _fields_ = [("cosine", ctypes.c_float), ("sine", ctypes.c_float)]

        ref.b2ComputeCosSin.restype = CosSin
        ref.b2ComputeCosSin.argtypes = [ctypes.c_float]
        for i in range(0, len(a), 41):
            cs = ref.b2ComputeCosSin(float(a[i]))
            assert (cs.cosine, cs.sine) == (c[i], s[i])


def test_make_soft_closed_form(oracle):
    """Against float64 closed forms: bias = w/(2z+hw), impulse = 1/(1+hw(2z+hw)), mass + impulse = 1."""
    for hz, z, h in [(30.0, 10.0, 1 / 240), (60.0, 0.0, 1 / 240), (5.0, 0.7, 1 / 60)]:
        s = oracle.o_make_soft(hz, z, h)
        w = 2 * math.pi * hz
        assert abs(s.biasRate - w / (2 * z + h * w)) < 1e-4 * s.biasRate
        assert abs(s.impulseScale - 1 / (1 + h * w * (2 * z + h * w))) < 1e-6
        assert abs(s.massScale + s.impulseScale - 1.0) < 1e-6
    s = oracle.o_make_soft(0.0, 1.0, 0.01)
    assert (s.biasRate, s.massScale, s.impulseScale) == (0.0, 0.0, 0.0)
