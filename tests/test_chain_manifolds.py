"""One-sided chain-segment manifolds against the reference library (CPU).

Random chain segments with convex, flat and concave ghost neighbors collide with circles,
capsules and polygons posed around them, on both sides and near the segment ends. The capsule and
polygon forms carry a GJK simplex cache across calls, as a persisting contact does: each case is
evaluated three times with a slowly moving pose, feeding the cache forward, and every manifold and
cache state is compared bit for bit with the reference (src/manifold.c:1042-1641).
"""

import ctypes
import math

import numpy as np

from b2types import (Capsule, ChainSegment, Circle, Polygon, Rot, Segment, SimplexCache, Transform, Vec2,
                     bind_chain_functions, manifold_words)


def make_polygon(lib, rng):
    lib.b2MakeRoundedBox.restype = Polygon
    lib.b2MakeRoundedBox.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
    return lib.b2MakeRoundedBox(float(rng.uniform(0.1, 0.6)), float(rng.uniform(0.1, 0.6)),
                                float(rng.choice([0.0, 0.05, 0.1])))


def random_segment(rng):
    p1 = rng.uniform(-1, 1, 2)
    d = rng.uniform(-2, 2, 2)
    p2 = p1 + d
    e = d / max(np.linalg.norm(d), 1e-3)
    # ghost vertices bend the neighbors convex, flat or concave
    def ghost(anchor, sign):
        a = rng.uniform(-1.2, 1.2)
        rot = np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
        return anchor + sign * rot @ e * rng.uniform(0.3, 1.5)
    s = ChainSegment()
    s.ghost1 = Vec2(*ghost(p1, -1.0))
    s.segment = Segment(Vec2(*p1), Vec2(*p2))
    s.ghost2 = Vec2(*ghost(p2, 1.0))
    return s, p1, p2


def test_chain_manifolds_match_reference(ref_scenes):
    import box2d_amd

    mine = bind_chain_functions(box2d_amd.library())
    ref = bind_chain_functions(ctypes.CDLL(ref_scenes.path))
    rng = np.random.default_rng(21)
    nonempty = 0
    for trial in range(2500):
        seg, p1, p2 = random_segment(rng)
        kind = trial % 3
        t = rng.uniform(-0.3, 1.3)
        base = p1 + t * (p2 - p1) + rng.uniform(-0.8, 0.8, 2)
        angle = rng.uniform(-math.pi, math.pi)
        c1, c2 = SimplexCache(), SimplexCache()
        for k in range(3):
            xf = Transform(Vec2(*(base + k * rng.uniform(-0.05, 0.05, 2))), Rot(math.cos(angle + 0.05 * k),
                                                                               math.sin(angle + 0.05 * k)))
            if kind == 0:
                circle = Circle(Vec2(*rng.uniform(-0.1, 0.1, 2)), float(rng.uniform(0.05, 0.6)))
                a = mine.b2CollideChainSegmentAndCircle(ctypes.byref(seg), ctypes.byref(circle), xf)
                b = ref.b2CollideChainSegmentAndCircle(ctypes.byref(seg), ctypes.byref(circle), xf)
            elif kind == 1:
                h = rng.uniform(0.05, 0.6)
                capsule = Capsule(Vec2(-h, 0.0), Vec2(h, 0.0), float(rng.uniform(0.02, 0.3)))
                a = mine.b2CollideChainSegmentAndCapsule(ctypes.byref(seg), ctypes.byref(capsule), xf, ctypes.byref(c1))
                b = ref.b2CollideChainSegmentAndCapsule(ctypes.byref(seg), ctypes.byref(capsule), xf, ctypes.byref(c2))
            else:
                poly = make_polygon(ref, rng)
                a = mine.b2CollideChainSegmentAndPolygon(ctypes.byref(seg), ctypes.byref(poly), xf, ctypes.byref(c1))
                b = ref.b2CollideChainSegmentAndPolygon(ctypes.byref(seg), ctypes.byref(poly), xf, ctypes.byref(c2))
            wa, wb = manifold_words(a), manifold_words(b)
            assert np.array_equal(wa, wb), f"trial {trial} step {k} kind {kind}: {wa.tolist()} vs {wb.tolist()}"
            assert bytes(c1) == bytes(c2), f"trial {trial} step {k}: cache differs"
            nonempty += b.pointCount > 0
    assert nonempty > 1000
