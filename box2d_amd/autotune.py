"""Offline tuner of the batched persistent solver (b2amd_StepWorlds) on the GPU it runs on.

For each world-size bucket (awake bodies 1-64, 65-256, 257-1024, 1025-5400) it builds a batch of
worlds whose sizes fall in the bucket (box pyramids, box/circle piles and hanging capsule chains,
sizes log-uniform, seeded), and for every candidate configuration -- workgroup size 64 to 1024
threads, warm start body by body or color by color -- it steps a fresh copy of the batch and takes
the median GPU time of the persistent launch over the timed steps. The fastest configuration per
bucket is written as the table of csrc/batch_tuning.h (rebuild the library to use it), and every
timing to a markdown report. At run time each world takes its bucket's warm-start form, and a
batch's single launch takes the workgroup size of the bucket that holds the most of its worlds. The configuration never changes a result, only the
time; the batch tests check that for every workgroup size and both warm-start forms.

    python -m box2d_amd.autotune --header gpurun_out/batch_tuning.h --report gpurun_out/batch_autotune.md
"""

import argparse
import datetime
import statistics
import sys
import time

import box2d_amd

THREADS = (64, 128, 256, 512, 1024)
WARM = (0, 1)
# (bucket, smallest and largest world of its batch, worlds in the batch)
BUCKETS = ((0, 8, 64, 2048), (1, 65, 256, 1024), (2, 257, 1024, 256), (3, 1025, 5000, 48))
DT = 1.0 / 60.0
SUBSTEPS = 4


def measure(scenes, lib, bucket, lo, hi, worlds, threads, warm, warmup, steps, seed):
    """Median GPU milliseconds of the bucket's persistent launch over `steps` batch steps of a fresh
    batch (after `warmup` steps), with the bucket set to (threads, warm)."""
    assert lib.b2amd_SetBatchTuning(bucket, threads, warm) == 0
    # worlds whose body count falls in the bucket (a pyramid rounds its target down to whole rows,
    # chain worlds stop at 160 bodies), the same ones for every configuration
    handles = []
    round_seed = seed
    while len(handles) < worlds:
        for spec in box2d_amd.batch_world_specs(worlds, lo=lo, hi=hi, seed=round_seed):
            h = scenes.create(box2d_amd.SCENE_BATCH_WORLD, spec)
            if lo <= scenes.body_count(h) <= hi and len(handles) < worlds:
                handles.append(h)
            else:
                scenes.destroy(h)
        round_seed += 1000
    try:
        ids = (box2d_amd.WorldId * len(handles))(*[scenes.world_id(h) for h in handles])
        box2d_amd.step_worlds(ids, warmup, DT, SUBSTEPS)
        times = []
        for _ in range(steps):
            st = box2d_amd.step_worlds(ids, 1, DT, SUBSTEPS)
            if st.bucketWorlds[bucket] != len(handles) or st.bucketThreads[bucket] != threads or \
                    st.bucketWarmByBody[bucket] != warm:
                raise RuntimeError(f"bucket {bucket}: the batch ran {st.bucketWorlds[bucket]} worlds with "
                                   f"({st.bucketThreads[bucket]}, {st.bucketWarmByBody[bucket]}), asked for "
                                   f"{len(handles)} with ({threads}, {warm})")
            times.append(st.bucketMs[bucket])
        bodies = sum(scenes.body_count(h) for h in handles)
        return statistics.median(times), bodies
    finally:
        for h in handles:
            scenes.destroy(h)
        lib.b2amd_SetBatchTuning(bucket, 0, 0)


def write_header(path, table, source):
    rows = ", ".join(f"{{ {m}, {t}, {w} }}" for m, t, w in table)
    text = f"""// Configuration of the batched persistent solver per world-size bucket, chosen offline on an
// MI355X by box2d_amd/autotune.py (which rewrites this file): for each bucket, the workgroup size
// and warm-start form with the lowest median GPU time of the substep-loop launch over a batch of
// worlds of that size. A batch's launch takes the size of the bucket with the most worlds; each
// world its bucket's warm-start form. The results do not depend on these values, only the time.
#pragma once

// {{ max awake bodies, workgroup threads, warm start body by body }}
#define B2AMD_BATCH_TUNING_TABLE                                                                                         \\
	{{ {rows} }}
#define B2AMD_BATCH_TUNING_SOURCE "{source}"
"""
    with open(path, "w") as f:
        f.write(text)


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--header", default="batch_tuning.h", help="where to write the tuned table header")
    ap.add_argument("--report", default="batch_autotune.md", help="where to write the timings")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--seed", type=int, default=4242)
    args = ap.parse_args()

    lib = box2d_amd.library()
    scenes = box2d_amd.scenes()
    lib.b2amd_SetBatchSolverThreads(0)
    limits = [t["max_bodies"] for t in box2d_amd.batch_tuning()]

    t0 = time.perf_counter()
    results = {}
    table = []
    lines = ["# Batched solver autotune", "",
             f"`python -m box2d_amd.autotune --header {args.header} --report {args.report}` "
             f"({args.warmup} warm-up and {args.steps} timed batch steps per configuration, fresh worlds each, "
             f"{SUBSTEPS} substeps at {DT:.5f} s). Cells: median GPU ms of the bucket's persistent launch "
             "(HIP events around the dispatch); the chosen configuration is in bold.", ""]
    for bucket, lo, hi, worlds in BUCKETS:
        cell = {}
        bodies = 0
        for threads in THREADS:
            for warm in WARM:
                ms, bodies = measure(scenes, lib, bucket, lo, hi, worlds, threads, warm, args.warmup, args.steps,
                                     args.seed + bucket)
                cell[(threads, warm)] = ms
                print(f"bucket {bucket} ({lo}-{hi} bodies, {worlds} worlds): threads {threads} warm {warm}: "
                      f"{ms:.3f} ms", flush=True)
        best = min(cell, key=cell.get)
        results[bucket] = cell
        table.append((limits[bucket], best[0], best[1]))
        lines += [f"## Bucket {bucket}: worlds of {lo}-{hi} awake bodies ({worlds} worlds, {bodies} bodies)", "",
                  "| workgroup threads | warm start color by color | warm start body by body |", "|---|---|---|"]
        for threads in THREADS:
            row = []
            for warm in WARM:
                v = f"{cell[(threads, warm)]:.3f}"
                row.append(f"**{v}**" if (threads, warm) == best else v)
            lines.append(f"| {threads} | {row[0]} | {row[1]} |")
        lines.append("")
    source = f"box2d_amd/autotune.py, {datetime.date.today().isoformat()}"
    write_header(args.header, table, source)
    lines += ["## Table", "", "| bucket | max awake bodies | threads | warm start body by body |", "|---|---|---|---|"]
    lines += [f"| {k} | {m} | {t} | {w} |" for k, (m, t, w) in enumerate(table)]
    lines += ["", f"Tuning took {time.perf_counter() - t0:.0f} s."]
    with open(args.report, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.header} and {args.report}: {table}")


if __name__ == "__main__":
    sys.exit(main())
