"""Timing of the reference's own ten benchmark scenes: box2d_amd on the GPU against the reference
library on the host's cores, with the final state checked against the reference's.

The scene builders are the reference's shared/benchmarks.c (with shared/human.c and
shared/utils.c), compiled unchanged against box2d_amd's headers into
oracle/_ref/libb2amd_bench.so and against the reference library into oracle/_ref/libb2ref_bench.so
(`make -C oracle bench`). Each scene runs as the reference's benchmark app runs it
(benchmark/main.c:318-351): one untimed step, then steps-1 timed steps of 1/60 s with 4 substeps.
The GPU run's last step is reduced to the move-event hash of tests/golden/bench_scenes.json and
compared with the reference's. Prints one JSON line.

    python tests/bench_reference_scenes.py [--threads T] [--no-cpu] [--scenes a,b,...]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle.bench_scenes import AMD_BENCH, REF_BENCH, BenchScenes  # noqa: E402

# The reference's published figures: AMD Ryzen 9 7950X, SSE2 build, total ms over steps-1 timed
# steps at 1 and 8 threads (benchmark/amd7950x_sse2/<scene>.csv:2,9).
PUBLISHED_MS = {"compounds": (1398.62, 368.938), "joint_grid": (2388.9, 530.946), "junkyard": (3229.01, 757.608),
                "large_pyramid": (1134.89, 242.273), "many_pyramids": (1888.15, 301.354), "rain": (7207.07, 1589.56),
                "smash": (1267.67, 376.557), "spinner": (4026.1, 1012.35), "tumbler": (1426.18, 424.385),
                "washer": (4085.76, 901.712)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0, help="reference worker threads (0: min(16, host cores))")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--scenes", default="")
    args = ap.parse_args()
    threads = args.threads or min(16, os.cpu_count() or 1)

    import box2d_amd

    lib = box2d_amd.library()
    with open(os.path.join(ROOT, "tests", "golden", "bench_scenes.json")) as f:
        golden = {x["name"]: x for x in json.load(f)["scenes"]}
    amd = BenchScenes(AMD_BENCH)
    ref = None if args.no_cpu or not os.path.exists(REF_BENCH) else BenchScenes(REF_BENCH)
    names = args.scenes.split(",") if args.scenes else amd.names
    out = {}
    for name in names:
        n = amd.steps(name)
        amd.create(name)
        amd.run(1)
        lib.b2amd_Synchronize()
        t0 = time.perf_counter()
        amd.run(n - 2)
        last = list(amd.step())
        lib.b2amd_Synchronize()
        gpu_ms = (time.perf_counter() - t0) * 1e3 / (n - 1)
        amd.destroy()
        pub1, pub8 = PUBLISHED_MS[name]
        rec = {"steps": n, "gpu_ms_per_step": round(gpu_ms, 4),
               "state_match": bool(last == golden[name]["per_step"][n - 1]),
               "published_7950x_1t_ms_per_step": round(pub1 / (n - 1), 4),
               "published_7950x_8t_ms_per_step": round(pub8 / (n - 1), 4)}
        if ref is not None:
            ref.create(name, workers=threads)
            ref.run(1)
            t0 = time.perf_counter()
            ref.run(n - 1)
            cpu_ms = (time.perf_counter() - t0) * 1e3 / (n - 1)
            ref.destroy()
            rec.update({"cpu_ms_per_step": round(cpu_ms, 4), "cpu_threads": threads,
                        "speedup_vs_cpu": round(cpu_ms / gpu_ms, 3)})
        out[name] = rec
        print(f"[scenes] {name}: {rec}", file=sys.stderr, flush=True)
    print(json.dumps({"reference_scenes": out, "protocol": "benchmark/main.c: 1 untimed step, steps-1 timed",
                      "cpu_kind": "reference library (oracle/_ref, -O2 -ffp-contract=off, built-in scheduler)"}))


if __name__ == "__main__":
    main()
