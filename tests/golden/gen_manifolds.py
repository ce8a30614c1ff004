"""Write the narrowphase golden fixture from the reference's manifold functions.

Shapes are built with the reference's own constructors (b2MakeBox, b2MakeRoundedBox,
b2ComputeHull + b2MakePolygon, include/box2d/collision.h) and collided with its b2Collide*
functions (src/manifold.c) from oracle/_ref/libb2ref_scenes.so, which `make -C oracle ref`
compiles from the reference sources. The fixture holds the raw b2amd_Shape records, the
relative transforms and the reference manifolds as uint32 words: data only.

    python tests/golden/gen_manifolds.py
"""

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from b2types import (CAPSULE, CIRCLE, PAIR_TYPES, POLYGON, SEGMENT, Hull, LocalManifold, Polygon, Rot,  # noqa: E402
                     Shape, Transform, Vec2, bind_manifold_functions, collide_host, manifold_words)

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libb2ref_scenes.so")
PER_TYPE = 600
SEED = 12345


def make_shape(lib, rng, kind):
    s = Shape()
    s.type = kind
    if kind == CIRCLE:
        s.circle.center = Vec2(*rng.uniform(-0.3, 0.3, 2))
        s.circle.radius = rng.uniform(0.05, 0.6)
    elif kind == CAPSULE:
        h = rng.uniform(0.0, 0.7)
        a = rng.uniform(-np.pi, np.pi)
        c = rng.uniform(-0.2, 0.2, 2)
        d = np.array([np.cos(a), np.sin(a)]) * h
        s.capsule.center1 = Vec2(*(c - d))
        s.capsule.center2 = Vec2(*(c + d))
        s.capsule.radius = rng.uniform(0.02, 0.3)
    elif kind == SEGMENT:
        p = rng.uniform(-1.0, 1.0, 2)
        q = p + rng.uniform(-1.5, 1.5, 2)
        s.segment.point1 = Vec2(*p)
        s.segment.point2 = Vec2(*q)
    else:
        form = rng.integers(0, 4)
        if form == 0:
            poly = lib.b2MakeBox(ctypes.c_float(rng.uniform(0.1, 0.8)), ctypes.c_float(rng.uniform(0.1, 0.8)))
        elif form == 1:
            poly = lib.b2MakeRoundedBox(ctypes.c_float(rng.uniform(0.1, 0.6)), ctypes.c_float(rng.uniform(0.1, 0.6)),
                                        ctypes.c_float(rng.uniform(0.01, 0.2)))
        else:
            while True:
                n = int(rng.integers(3, 9))
                pts = (Vec2 * 8)()
                for i in range(n):
                    a = rng.uniform(-np.pi, np.pi)
                    r = rng.uniform(0.2, 0.9)
                    pts[i] = Vec2(r * np.cos(a), r * np.sin(a))
                hull = lib.b2ComputeHull(pts, n)
                if hull.count >= 3:
                    break
            radius = 0.0 if form == 2 else rng.uniform(0.01, 0.15)
            poly = lib.b2MakePolygon(ctypes.byref(hull), ctypes.c_float(radius))
        s.polygon = poly
    return s


def main():
    if not os.path.exists(REF_LIB):
        sys.exit(f"{REF_LIB} missing: run `make -C oracle ref` first")
    lib = bind_manifold_functions(ctypes.CDLL(REF_LIB))
    lib.b2MakeBox.restype = Polygon
    lib.b2MakeBox.argtypes = [ctypes.c_float, ctypes.c_float]
    lib.b2MakeRoundedBox.restype = Polygon
    lib.b2MakeRoundedBox.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
    lib.b2ComputeHull.restype = Hull
    lib.b2ComputeHull.argtypes = [ctypes.POINTER(Vec2), ctypes.c_int]
    lib.b2MakePolygon.restype = Polygon
    lib.b2MakePolygon.argtypes = [ctypes.POINTER(Hull), ctypes.c_float]

    rng = np.random.default_rng(SEED)
    shapes_a, shapes_b, xfs, outs = [], [], [], []
    for ta, tb in PAIR_TYPES:
        for k in range(PER_TYPE):
            a = make_shape(lib, rng, ta)
            b = make_shape(lib, rng, tb)
            # relative pose: mostly overlapping or near contact, some exact alignments
            reach = 1.0 if k % 4 else 0.4
            p = rng.uniform(-reach, reach, 2)
            angle = rng.uniform(-np.pi, np.pi) if k % 5 else (np.pi / 2) * int(rng.integers(0, 4))
            xf = Transform(Vec2(*p), Rot(np.cos(angle), np.sin(angle)))
            m = collide_host(lib, a, b, xf)
            shapes_a.append(bytes(a))
            shapes_b.append(bytes(b))
            xfs.append(bytes(xf))
            outs.append(manifold_words(m))

    path = os.path.join(HERE, "manifolds.npz")
    np.savez_compressed(path,
                        shapes_a=np.frombuffer(b"".join(shapes_a), dtype=np.uint8).reshape(len(shapes_a), -1),
                        shapes_b=np.frombuffer(b"".join(shapes_b), dtype=np.uint8).reshape(len(shapes_b), -1),
                        transforms=np.frombuffer(b"".join(xfs), dtype=np.uint8).reshape(len(xfs), -1),
                        manifolds=np.stack(outs))
    counts = np.stack(outs)[:, 10]
    print(f"wrote {path} ({os.path.getsize(path)} bytes): {len(outs)} pairs, point counts "
          f"{np.bincount(counts.astype(np.int64), minlength=3).tolist()}")


if __name__ == "__main__":
    main()
