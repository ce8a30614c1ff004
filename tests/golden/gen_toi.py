"""Write the time-of-impact and GJK distance fixtures from the reference library.

Inputs mix general random convex proxies and sweeps with continuous-collision-like cases: a
static chain segment far from the sweep origin (the reference's base frame) against a fast,
falling rounded polygon, capsule or circle. Outputs come from the reference's b2TimeOfImpact and
b2ShapeDistance (src/distance.c) in oracle/_ref/libb2ref_scenes.so. Data only.

    python tests/golden/gen_toi.py
"""

import ctypes
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from b2types import (DistanceInput, Rot, ShapeProxy, SimplexCache, Sweep, TOIInput, Transform, Vec2,  # noqa: E402
                     bind_distance_functions)

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libb2ref_scenes.so")


def rot(a):
    return Rot(math.cos(a), math.sin(a))


def random_proxy(rng, radius=None):
    kind = int(rng.integers(0, 3))
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
        r = rng.uniform(0.2, 0.6)
        for i in range(n):
            p.points[i] = Vec2(r * math.cos(angles[i]), r * math.sin(angles[i]))
        p.count = n
        p.radius = 0.05 if rng.uniform() < 0.5 else 0.0
    if radius is not None:
        p.radius = radius
    return p


def toi_case(rng, k):
    inp = TOIInput()
    if k % 2 == 0:
        # general
        inp.proxyA = random_proxy(rng)
        inp.proxyB = random_proxy(rng)
        a1 = rng.uniform(-math.pi, math.pi)
        inp.sweepA = Sweep(Vec2(0, 0), Vec2(0, 0), Vec2(*rng.uniform(-0.2, 0.2, 2)), rot(a1), rot(a1 + rng.uniform(-0.3, 0.3)))
        start = rng.uniform(-3.0, 3.0, 2)
        end = -start * rng.uniform(0.0, 1.5) + rng.uniform(-0.5, 0.5, 2)
        b1 = rng.uniform(-math.pi, math.pi)
        inp.sweepB = Sweep(Vec2(*rng.uniform(-0.1, 0.1, 2)), Vec2(*start), Vec2(*end), rot(b1), rot(b1 + rng.uniform(-1.0, 1.0)))
        inp.maxFraction = 1.0 if k % 3 else float(rng.uniform(0.2, 1.0))
    else:
        # static chain segment in world coordinates, the fast body's start as the frame origin
        base = np.array([rng.uniform(-30, 30), rng.uniform(0.2, 2.0)], dtype=np.float32)
        x0 = float(np.floor(base[0] / 2.0) * 2.0 + 2.0 * int(rng.integers(-1, 2)))
        seg = ShapeProxy()
        seg.points[0] = Vec2(x0 + 2.0, 0.0)
        seg.points[1] = Vec2(x0, 0.0)
        seg.count = 2
        seg.radius = 0.0
        inp.proxyA = seg
        c = Vec2(float(np.float32(0.0) - base[0]), float(np.float32(0.0) - base[1]))
        inp.sweepA = Sweep(Vec2(0, 0), c, c, Rot(1.0, 0.0), Rot(1.0, 0.0))
        inp.proxyB = random_proxy(rng)
        drop = rng.uniform(0.2, 1.5)
        a = rng.uniform(-math.pi, math.pi)
        w = rng.uniform(-0.2, 0.2)
        inp.sweepB = Sweep(Vec2(*rng.uniform(-0.05, 0.05, 2)), Vec2(0, 0), Vec2(float(rng.uniform(-0.1, 0.1)), -drop),
                           rot(a), rot(a + w))
        inp.maxFraction = 1.0
    return inp


def main():
    lib = bind_distance_functions(ctypes.CDLL(REF_LIB))
    rng = np.random.default_rng(31)
    n = 4000
    toi_in = (TOIInput * n)()
    toi_out = []
    for k in range(n):
        toi_in[k] = toi_case(rng, k)
        o = lib.b2TimeOfImpact(ctypes.byref(toi_in[k]))
        toi_out.append(bytes(o))
    dist_in = (DistanceInput * n)()
    caches_in = (SimplexCache * n)()
    dist_out, caches_out = [], []
    for k in range(n):
        d = DistanceInput()
        d.proxyA = random_proxy(rng)
        d.proxyB = random_proxy(rng)
        d.transform = Transform(Vec2(*rng.uniform(-2.0, 2.0, 2)), rot(rng.uniform(-math.pi, math.pi)))
        d.useRadii = bool(k & 1)
        dist_in[k] = d
        cache = SimplexCache()
        o = lib.b2ShapeDistance(ctypes.byref(d), ctypes.byref(cache), None, 0)
        dist_out.append(bytes(o))
        caches_out.append(bytes(cache))
    path = os.path.join(HERE, "toi_distance.npz")
    np.savez_compressed(path,
                        toi_inputs=np.frombuffer(bytes(toi_in), dtype=np.uint8).reshape(n, -1),
                        toi_outputs=np.frombuffer(b"".join(toi_out), dtype=np.uint8).reshape(n, -1),
                        dist_inputs=np.frombuffer(bytes(dist_in), dtype=np.uint8).reshape(n, -1),
                        dist_outputs=np.frombuffer(b"".join(dist_out), dtype=np.uint8).reshape(n, -1),
                        dist_caches=np.frombuffer(b"".join(caches_out), dtype=np.uint8).reshape(n, -1))
    states = np.frombuffer(b"".join(toi_out), dtype=np.int32).reshape(n, -1)[:, 0]
    print(f"wrote {path} ({os.path.getsize(path)} bytes); toi states {np.bincount(states, minlength=5).tolist()}")


if __name__ == "__main__":
    main()
