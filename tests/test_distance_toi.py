"""GJK distance and time of impact against the reference library (CPU).

libbox2d_amd's host b2ShapeDistance / b2TimeOfImpact compile the same source the continuous
collision kernel runs. Random convex proxies and sweeps are evaluated by both libraries and the
outputs compared bit for bit (reference src/distance.c:424-611, 1148-1441). Known answers from
test/test_distance.c:45-120 pin the reference build itself.
"""

import ctypes
import math

import numpy as np
import pytest

from b2types import (DistanceInput, Rot, ShapeProxy, SimplexCache, Sweep, TOIInput, Transform, Vec2,
                     bind_distance_functions)


def random_proxy(rng):
    kind = rng.integers(0, 3)
    p = ShapeProxy()
    if kind == 0:
        p.count = 1
        p.points[0] = Vec2(*rng.uniform(-0.3, 0.3, 2))
        p.radius = rng.uniform(0.05, 0.5)
    elif kind == 1:
        p.count = 2
        c = rng.uniform(-0.2, 0.2, 2)
        d = rng.uniform(-0.6, 0.6, 2)
        p.points[0] = Vec2(*(c - d))
        p.points[1] = Vec2(*(c + d))
        p.radius = rng.uniform(0.0, 0.3)
    else:
        n = int(rng.integers(3, 9))
        angles = np.sort(rng.uniform(0, 2 * math.pi, n))
        r = rng.uniform(0.3, 0.9)
        for i in range(n):
            p.points[i] = Vec2(r * math.cos(angles[i]), r * math.sin(angles[i]))
        p.count = n
        p.radius = rng.uniform(0.0, 0.1) if rng.uniform() < 0.5 else 0.0
    return p


def rot(a):
    return Rot(math.cos(a), math.sin(a))


def bits(*vals):
    return np.array(vals, dtype=np.float32).view(np.uint32).tolist()


@pytest.fixture(scope="module")
def libs(ref_scenes):
    import box2d_amd

    mine = bind_distance_functions(box2d_amd.library())
    ref = bind_distance_functions(ctypes.CDLL(ref_scenes.path))
    return mine, ref


def test_shape_distance_matches_reference(libs):
    mine, ref = libs
    rng = np.random.default_rng(11)
    for trial in range(3000):
        inp = DistanceInput()
        inp.proxyA = random_proxy(rng)
        inp.proxyB = random_proxy(rng)
        inp.transform = Transform(Vec2(*rng.uniform(-2.0, 2.0, 2)), rot(rng.uniform(-math.pi, math.pi)))
        inp.useRadii = bool(rng.integers(0, 2))
        c1, c2 = SimplexCache(), SimplexCache()
        a = mine.b2ShapeDistance(ctypes.byref(inp), ctypes.byref(c1), None, 0)
        b = ref.b2ShapeDistance(ctypes.byref(inp), ctypes.byref(c2), None, 0)
        got = bits(a.pointA.x, a.pointA.y, a.pointB.x, a.pointB.y, a.normal.x, a.normal.y, a.distance) + [a.iterations]
        want = bits(b.pointA.x, b.pointA.y, b.pointB.x, b.pointB.y, b.normal.x, b.normal.y, b.distance) + [b.iterations]
        assert got == want, f"trial {trial}"
        assert bytes(c1) == bytes(c2), f"trial {trial} cache"


def test_time_of_impact_matches_reference(libs):
    mine, ref = libs
    rng = np.random.default_rng(12)
    hits = 0
    for trial in range(3000):
        inp = TOIInput()
        inp.proxyA = random_proxy(rng)
        inp.proxyB = random_proxy(rng)
        a1 = rng.uniform(-math.pi, math.pi)
        inp.sweepA = Sweep(Vec2(0, 0), Vec2(0, 0), Vec2(*rng.uniform(-0.2, 0.2, 2)), rot(a1), rot(a1 + rng.uniform(-0.3, 0.3)))
        start = rng.uniform(-3.0, 3.0, 2)
        end = -start * rng.uniform(0.0, 1.5) + rng.uniform(-0.5, 0.5, 2)
        b1 = rng.uniform(-math.pi, math.pi)
        inp.sweepB = Sweep(Vec2(*rng.uniform(-0.1, 0.1, 2)), Vec2(*start), Vec2(*end), rot(b1), rot(b1 + rng.uniform(-1.0, 1.0)))
        inp.maxFraction = 1.0 if trial % 3 else float(rng.uniform(0.2, 1.0))
        x = mine.b2TimeOfImpact(ctypes.byref(inp))
        y = ref.b2TimeOfImpact(ctypes.byref(inp))
        got = [x.state] + bits(x.point.x, x.point.y, x.normal.x, x.normal.y, x.fraction)
        want = [y.state] + bits(y.point.x, y.point.y, y.normal.x, y.normal.y, y.fraction)
        assert got == want, f"trial {trial}: {got} vs {want}"
        hits += y.state == 3
    assert hits > 300


def test_sweep_transform_matches_reference(libs):
    mine, ref = libs
    rng = np.random.default_rng(13)
    for _ in range(500):
        sw = Sweep(Vec2(*rng.uniform(-1, 1, 2)), Vec2(*rng.uniform(-5, 5, 2)), Vec2(*rng.uniform(-5, 5, 2)),
                   rot(rng.uniform(-3, 3)), rot(rng.uniform(-3, 3)))
        t = float(rng.uniform(0, 1))
        a = mine.b2GetSweepTransform(ctypes.byref(sw), t)
        b = ref.b2GetSweepTransform(ctypes.byref(sw), t)
        assert bytes(a) == bytes(b)


def test_distance_known_answer(libs):
    """test/test_distance.c:45-73: two unit-separated squares are 1 apart along x."""
    _, ref = libs
    for lib in libs:
        sq = ShapeProxy()
        pts = [(-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)]
        for i, p in enumerate(pts):
            sq.points[i] = Vec2(*p)
        sq.count = 4
        sq.radius = 0.0
        inp = DistanceInput(sq, sq, Transform(Vec2(2.0, 0.0), Rot(1.0, 0.0)), False)
        cache = SimplexCache()
        out = lib.b2ShapeDistance(ctypes.byref(inp), ctypes.byref(cache), None, 0)
        assert abs(out.distance - 1.0) < 1e-6
        assert abs(out.normal.x - 1.0) < 1e-6
