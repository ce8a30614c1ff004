"""Write the benchmark-scene fixture from the reference library.

The scenes are the reference's own benchmark scenes (shared/benchmarks.c, run as
benchmark/main.c runs them), compiled with the driver oracle/bench_scenes.c into
oracle/_ref/libb2ref_bench.so by `make -C oracle bench`. For every step of every scene the
fixture holds the move-event count, the awake body count and the FNV-1a hash of the move events
sorted by body index (index, generation, transform bits, fell-asleep flag), plus the first 64 sorted move
records of the first step. Data only.

    python tests/golden/gen_bench_scenes.py
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.bench_scenes import REF_BENCH, BenchScenes  # noqa: E402


def main():
    if not os.path.exists(REF_BENCH):
        sys.exit(f"{REF_BENCH} missing: run `make -C oracle bench` first")
    lib = BenchScenes(REF_BENCH)
    scenes = []
    for name in lib.names:
        steps = lib.steps(name)
        lib.create(name)
        per_step = []
        first = None
        for s in range(steps):
            per_step.append(list(lib.step()))
            if s == 0:
                first = lib.moves()[:64].reshape(-1).tolist()
        lib.destroy()
        scenes.append({"name": name, "steps": steps, "per_step": per_step, "first_moves": first})
        print(f"{name}: {steps} steps, last {per_step[-1]}", flush=True)
    out = {"generator": "tests/golden/gen_bench_scenes.py",
           "reference": "erincatto/box2d v3.2 (src/*.c, shared/benchmarks.c, -O2 -ffp-contract=off)",
           "protocol": "default world def, continuous on, scene step function then b2World_Step(1/60, 4)",
           "scenes": scenes}
    path = os.path.join(HERE, "bench_scenes.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
