#!/usr/bin/env python3
"""Benchmark of the Box2D world step on one AMD Instinct MI355X.

Workload: the reference's large_pyramid benchmark scene (shared/benchmarks.c:88-131): a
100-row triangular pyramid of 5050 dynamic 0.5 m boxes (density 1, friction 0.6) on a static
100 x 1 m (half-extent) ground box, gravity (0, -10), 1/60 s steps with 4 substeps, sleep off.
Bodies are created through the public Box2D C API by the scene driver and stepped with
b2World_Step; all body, shape and contact state stays resident in GPU memory between steps.

Protocol (benchmark/main.c:307-377 style): W untimed warm-up steps, then K timed steps bracketed
by device synchronization; one JSON line on stdout. value = body-substeps per second
(bodies x substeps / seconds per step). After the timed region a second pass of the same scene
records HIP events around every kernel launch on the world's stream to price the dominant kernel
against HBM bandwidth (algorithmic bytes per launch / average launch duration). The CPU baseline
is the reference library compiled from its own sources (oracle/_ref) stepping the same scene on
the host's cores, timed on a bounded sample of steps.
"""

import argparse
import csv
import ctypes
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SUBSTEPS = 4
DT = 1.0 / 60.0
# Published reference figure for this scene and metric: large_pyramid, 8 threads, SSE2, AMD Ryzen 9
# 7950X: 242.273 ms over 499 steps (benchmark/amd7950x_sse2/large_pyramid.csv:9) -> 41.6 M body-substeps/s.
PUBLISHED_BODY_SUBSTEPS = 5050 * SUBSTEPS / (242.273e-3 / 499)
METRIC = "body-substeps/s & ms/step: batched+adjoint vs plain GPU, hash-matched; CPU ref"


def golden_hash(rows, ground, steps):
    path = os.path.join(ROOT, "tests", "golden", "world_runs.json")
    try:
        with open(path) as f:
            runs = json.load(f)
    except OSError:
        return None
    for s in runs["scenes"]:
        if s["kind"] == 0 and s["iparams"][:2] == [rows, 0] and abs(s["fparams"][0] - ground) < 1e-6:
            return s["hashes"].get(str(steps))
    return None


# kernel class (b2amd_KernelStats name) -> kernel symbol in rocprofv3 output
KERNEL_SYMBOLS = {"relax": "ColorKernel<2>", "solve": "ColorKernel<1>", "color_tail": "ColorTailKernel<2>",
                  "integrate_velocities_warm_start": "IntegrateWarmStartKernel"}
COPY_MB = 256


def pmc_per_dispatch(counter, child_args, symbol, tmp):
    """Average of one rocprofv3 counter (KiB per dispatch) over the dispatches of `symbol` in a
    child run of this script; None when the profiler is missing or the run fails."""
    out = os.path.join(tmp, counter + "_" + str(abs(hash(tuple(child_args)))))
    cmd = ["timeout", "-s", "KILL", "90", "rocprofv3", "--pmc", counter, "-f", "csv", "-d", out, "-o", "c", "--",
           sys.executable, os.path.abspath(__file__)] + child_args
    try:
        subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=tmp, timeout=120, check=True)
    except (OSError, subprocess.SubprocessError):
        return None
    values = []
    for dirpath, _, files in os.walk(out):
        for name in files:
            if name.endswith("counter_collection.csv"):
                with open(os.path.join(dirpath, name)) as f:
                    for row in csv.DictReader(f):
                        if symbol in row.get("Kernel_Name", "") and row.get("Counter_Name") == counter:
                            values.append(float(row["Counter_Value"]))
    return sum(values) / len(values) if values else None


def live_traffic(kernel, rows, ground, warmup, steps):
    """HBM bytes per launch of the dominant kernel, measured now: FETCH_SIZE and WRITE_SIZE of the
    kernel over the same window of the same scene, each scaled by what the counter reads for a
    plain device copy of COPY_MB MiB (the counters' units differ from bytes by a fixed factor).
    Returns (bytes or None, details)."""
    symbol = KERNEL_SYMBOLS.get(kernel)
    if symbol is None or shutil.which("rocprofv3") is None:
        return None, {"reason": "no kernel symbol or no rocprofv3"}
    probe = ["--traffic-probe", "--rows", str(rows), "--ground", str(ground), "--warmup", str(warmup), "--steps",
             str(steps)]
    copy = ["--copy-calibration", str(COPY_MB)]
    copy_kib = COPY_MB * 1024.0
    tmp = tempfile.mkdtemp(prefix="b2amd_pmc_")
    try:
        fetch = pmc_per_dispatch("FETCH_SIZE", probe, symbol, tmp)
        write = pmc_per_dispatch("WRITE_SIZE", probe, symbol, tmp)
        copy_fetch = pmc_per_dispatch("FETCH_SIZE", copy, "copyBuffer", tmp)
        copy_write = pmc_per_dispatch("WRITE_SIZE", copy, "copyBuffer", tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    details = {"kernel_symbol": symbol, "fetch_kib": fetch, "write_kib": write, "copy_mib": COPY_MB,
               "copy_fetch_kib": copy_fetch, "copy_write_kib": copy_write}
    if None in (fetch, write, copy_fetch, copy_write) or copy_fetch <= 0 or copy_write <= 0:
        details["reason"] = "a counter run failed"
        return None, details
    details["fetch_scale"] = copy_kib / copy_fetch
    details["write_scale"] = copy_kib / copy_write
    traffic = 1024.0 * (fetch * details["fetch_scale"] + write * details["write_scale"])
    return traffic, details


def kernel_stats(lib, world_id):
    from box2d_amd import KernelStats

    arr = (KernelStats * 32)()
    lib.b2amd_World_GetKernelStats.restype = ctypes.c_int
    n = lib.b2amd_World_GetKernelStats(world_id, arr, 32)
    out = []
    for i in range(n):
        k = arr[i]
        if k.launches > 0:
            out.append({"name": k.name.decode(), "launches": k.launches, "total_ms": k.totalMs,
                        "bytes": k.algorithmicBytes})
    return out


def host_cores():
    """CPUs this process may run on: its affinity set, capped by a cgroup v2 CPU quota if any."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return n


def cpu_baseline(rows, ground, warmup, steps, max_threads, budget_s=30.0):
    """Reference library (compiled from the reference sources) on this host's cores, stepping the
    same scene over the same step window as the GPU's timed region: warmup untimed steps, then
    steps timed. Thread counts 1..max_threads with the reference's built-in scheduler, the
    minimum of 4 runs per count (benchmark/main.c:307-377; 2 runs when 4 would exceed about
    budget_s of CPU time at roughly 2 ms per step). The fastest count is the baseline; the
    published window (steps 2..500) is timed at that count as well unless it is the window."""
    path = os.path.join(ROOT, "oracle", "_ref", "libb2ref_scenes.so")
    if not os.path.exists(path):
        return None
    from box2d_amd.scenes_api import SceneLib

    repeats = 4 if max_threads * (warmup + steps) * 4 * 2e-3 <= budget_s else 2
    long_window = not (warmup == 1 and steps == 499)
    ref = SceneLib(path)
    sweep = []
    bodies = None
    for t in range(1, max_threads + 1):
        best = None
        for _ in range(repeats):
            h = ref.pyramid(rows, ground, sleep=False, workers=t)
            bodies = ref.body_count(h)
            ref.step(h, warmup)
            run = ref.step_timed(h, steps) / steps
            ref.destroy(h)
            best = run if best is None else min(best, run)
        sweep.append({"threads": t, "ms_per_step": best, "body_substeps_per_s": bodies * SUBSTEPS / (best * 1e-3)})
    top = min(sweep, key=lambda r: r["ms_per_step"])
    out = {"value": top["body_substeps_per_s"], "unit": "body-substeps/s", "cores": top["threads"],
           "kind": "reference", "ms_per_step": top["ms_per_step"],
           "sample": f"steps {warmup + 1}..{warmup + steps} of the same {rows}-row pyramid (the GPU's timed window), "
                     f"workerCount=1..{max_threads} with the reference's built-in scheduler, minimum of {repeats} "
                     f"runs per count; value at the fastest count",
           "host_cores_available": max_threads, "thread_sweep": sweep}
    if long_window:
        best = None
        for _ in range(2):
            h = ref.pyramid(rows, ground, sleep=False, workers=top["threads"])
            ref.step(h)
            run = ref.step_timed(h, 499) / 499
            ref.destroy(h)
            best = run if best is None else min(best, run)
        out["steps_2_500_ms_per_step"] = best
    return out


def batched_workload(lib, scenes, worlds, steps, sample_stride=8, warmup=4):
    """Uneven batched worlds: body counts log-uniform in [8, 1024], each a box pyramid, a random
    box/circle pile or hanging capsule chains (seeded), stepped together with b2amd_StepWorlds.
    Then every sample_stride-th world is built again and stepped alone through b2World_Step, world
    after world within each step (the plain per-world GPU loop): its steps after the warm-up are
    timed, and each world's final transform hash must equal the batch's copy of that world."""
    import box2d_amd

    specs = box2d_amd.batch_world_specs(worlds, lo=8, hi=1024, seed=12345)
    handles = [scenes.create(box2d_amd.SCENE_BATCH_WORLD, spec) for spec in specs]
    bodies = sum(scenes.body_count(h) for h in handles)
    ids = (box2d_amd.WorldId * worlds)(*[scenes.world_id(h) for h in handles])
    box2d_amd.step_worlds(ids, warmup, DT, SUBSTEPS)  # warm-up
    lib.b2amd_Synchronize()
    solve_ms = 0.0
    solve_bytes = 0.0
    phase_ms = [0.0] * 5
    t0 = time.perf_counter()
    for _ in range(steps):
        st = box2d_amd.step_worlds(ids, 1, DT, SUBSTEPS)
        solve_ms += st.solveKernelMs
        solve_bytes += st.solveAlgorithmicBytes
        for k in range(5):
            phase_ms[k] += st.phaseMs[k]
    lib.b2amd_Synchronize()
    batched_s = (time.perf_counter() - t0) / steps
    sample = list(range(0, worlds, sample_stride))
    batch_hashes = [scenes.hash(handles[i]) for i in sample]
    for h in handles:
        scenes.destroy(h)

    loop = [scenes.create(box2d_amd.SCENE_BATCH_WORLD, specs[i]) for i in sample]
    loop_bodies = sum(scenes.body_count(h) for h in loop)
    for _ in range(warmup):
        for h in loop:
            scenes.step(h, 1, DT, SUBSTEPS)
    lib.b2amd_Synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for h in loop:
            scenes.step(h, 1, DT, SUBSTEPS)
    lib.b2amd_Synchronize()
    loop_s = (time.perf_counter() - t0) / steps
    loop_hashes = [scenes.hash(h) for h in loop]
    for h in loop:
        scenes.destroy(h)
    mismatched = [sample[k] for k in range(len(sample)) if batch_hashes[k] != loop_hashes[k]]
    loop_rate = loop_bodies * SUBSTEPS / loop_s
    batch_rate = bodies * SUBSTEPS / batched_s
    names = ("host_pairs_uploads", "gpu_collide", "host_touch_coloring", "gpu_solve", "host_refit_sleep")
    return {"worlds": worlds, "bodies": bodies, "steps": steps, "warmup": warmup, "ms_per_step": batched_s * 1e3,
            "body_substeps_per_s": batch_rate, "solve_kernel_ms_per_step": solve_ms / steps,
            "solve_workgroup_threads": st.solveThreads, "solve_workgroups": st.solveGrid,
            "solve_buckets": [{"bucket": k, "worlds": st.bucketWorlds[k], "threads": st.bucketThreads[k],
                               "warm_by_body": st.bucketWarmByBody[k], "workgroups": st.bucketWorkgroups[k],
                               "gpu_ms": st.bucketMs[k]} for k in range(4) if st.bucketWorlds[k]],
            "solver_tuning_table": box2d_amd.batch_tuning(),
            "phase_ms_per_step": {names[k]: phase_ms[k] / steps for k in range(5)},
            "hash_match": not mismatched, "hash_sample_worlds": len(sample), "hash_mismatched_worlds": mismatched[:16],
            "plain_loop_worlds": len(sample), "plain_loop_bodies": loop_bodies,
            "plain_loop_ms_per_step": loop_s * 1e3, "plain_loop_body_substeps_per_s": loop_rate,
            "speedup_vs_plain_loop": batch_rate / loop_rate,
            "solve_kernel_roofline": {"bound": "hbm", "achieved": solve_bytes / (solve_ms * 1e-3) / 1e9,
                                      "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                      "frac": solve_bytes / (solve_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                      "algorithmic_bytes_per_launch": solve_bytes / steps,
                                      "avg_launch_ms": solve_ms / steps}}


def rain_workload(lib, scenes, steps, cpu_steps, cpu_threads, no_cpu):
    """The 20,000-body mixed-shape rain into the U-shaped chain container (20k bodies, 1000 steps):
    whole-run GPU time and the final hash against the reference's; the reference library on the
    host's cores over the first cpu_steps steps, with the GPU's time over the same steps beside it."""
    want = None
    path = os.path.join(ROOT, "tests", "golden", "world_runs.json")
    try:
        with open(path) as f:
            runs = json.load(f)
        fx = next(x for x in runs["scenes"] if x["name"] == "rain")
        want = fx["hashes"].get(str(steps))
    except (OSError, StopIteration):
        pass
    iparams, fparams = (20000, 100, 0, 12345), (120.0,)
    h = scenes.create(4, iparams, fparams)
    wid = scenes.world_id(h)
    bodies = scenes.body_count(h)
    used = {0: 0, 1: 0, 2: 0, 3: 0}
    first_ms = None
    lib.b2amd_Synchronize()
    t0 = time.perf_counter()
    for k in range(1, steps + 1):
        scenes.step(h, 1, DT, SUBSTEPS)
        used[lib.b2amd_World_GetLastPairQuery(wid)] += 1
        if k == cpu_steps:
            lib.b2amd_Synchronize()
            first_ms = (time.perf_counter() - t0) * 1e3 / cpu_steps
    lib.b2amd_Synchronize()
    total = time.perf_counter() - t0
    got = scenes.hash(h)
    scenes.destroy(h)
    out = {"workload": "rain: 20,000 mixed shapes (circles, capsules, hulls, rounded boxes) into a U chain container",
           "bodies": bodies, "steps": steps, "gpu_ms_per_step": total * 1e3 / steps,
           "gpu_body_substeps_per_s": bodies * SUBSTEPS * steps / total, "hash": f"{got:#010x}",
           "hash_match": None if want is None else bool(got == want),
           "pair_query_steps": {"gpu_speculative": used[3], "gpu_at_update": used[2], "host_tree": used[1], "none": used[0]},
           "gpu_ms_per_step_first_steps": first_ms, "first_steps": cpu_steps}
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libb2ref_scenes.so")
    if not no_cpu and os.path.exists(ref_path):
        from box2d_amd.scenes_api import SceneLib

        ref = SceneLib(ref_path)
        hr = ref.create(4, iparams, fparams, workers=cpu_threads)
        ms = ref.step_timed(hr, cpu_steps, DT, SUBSTEPS) / cpu_steps
        ref.destroy(hr)
        out["cpu_baseline"] = {"value": ms, "unit": "ms/step", "cores": cpu_threads, "kind": "reference",
                               "sample": f"steps 1..{cpu_steps} of the same rain, workerCount={cpu_threads}"}
        out["speedup_vs_cpu_first_steps"] = ms / first_ms if first_ms else None
    return out


def push_workload(lib, worlds, steps, loop_worlds):
    """Differentiable push (256 worlds of a 10-row box pyramid plus a pusher box, 60 steps): the
    batched forward with every step recorded on a tape, the GPU adjoint over the tape, and the
    same worlds stepped without recording and one by one through b2World_Step."""
    import numpy as np

    import box2d_amd
    from box2d_amd.adjoint import BODY_STATE, PushWorlds

    rng = np.random.default_rng(4)
    params = np.stack([rng.uniform(2, 6, worlds), rng.uniform(-0.5, 0.5, worlds), rng.uniform(-1, 1, worlds),
                       rng.uniform(0.3, 0.9, worlds)], 1)
    goals = np.stack([rng.uniform(-1, 1, worlds), rng.uniform(8.5, 9.5, worlds)], 1)

    # plain batched forward (no tape)
    pw = PushWorlds(params)
    bodies = sum(len(b) for b in pw.body_ids)
    ids = (box2d_amd.WorldId * worlds)(*pw.ids)
    lib.b2amd_Synchronize()
    t0 = time.perf_counter()
    box2d_amd.step_worlds(ids, steps, DT, SUBSTEPS)
    lib.b2amd_Synchronize()
    plain_s = time.perf_counter() - t0
    pw.destroy()

    # recorded forward + backward
    pw = PushWorlds(params)
    t0 = time.perf_counter()
    tape = pw.record(steps)
    lib.b2amd_Synchronize()
    fwd_s = time.perf_counter() - t0
    final = tape.final_states()
    fg = np.zeros((worlds, tape.body_stride), dtype=BODY_STATE)
    for w in range(worlds):
        top = int(pw.body_ids[w][-2])
        fg[w]["center"][top] = 2.0 * (final[w]["center"][top] - goals[w])
    t0 = time.perf_counter()
    ig, fr, stats = tape.backward(fg)
    bwd_s = time.perf_counter() - t0
    pw.destroy()

    # plain per-world loop on a subset
    pw = PushWorlds(params[:loop_worlds])
    t0 = time.perf_counter()
    for _ in range(steps):
        for h in pw.handles:
            pw.scenes.step(h, 1, DT, SUBSTEPS)
    lib.b2amd_Synchronize()
    loop_s = (time.perf_counter() - t0) * worlds / loop_worlds
    pw.destroy()

    work = bodies * SUBSTEPS * steps
    return {"workload": f"push: {worlds} worlds x (10-row pyramid + pusher box), {steps} steps, loss |top - goal|^2",
            "worlds": worlds, "bodies": bodies, "steps": steps,
            "forward_recorded_ms_per_step": fwd_s * 1e3 / steps, "forward_plain_ms_per_step": plain_s * 1e3 / steps,
            "backward_ms_per_step": bwd_s * 1e3 / steps, "backward_kernel_ms": stats.backwardMs,
            "fwd_bwd_body_substeps_per_s": work / (fwd_s + bwd_s), "plain_batched_body_substeps_per_s": work / plain_s,
            "plain_loop_body_substeps_per_s": work / loop_s, "plain_loop_sample_worlds": loop_worlds,
            "replay_mismatches": stats.replayMismatches, "unsupported": stats.unsupported}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=499)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=100)
    ap.add_argument("--ground", type=float, default=100.0, help="ground box half width (m)")
    ap.add_argument("--profile-steps", type=int, default=60, help="steps of the per-kernel timing pass")
    ap.add_argument("--cpu-seconds", type=float, default=30.0, help="bound on the CPU baseline thread sweep")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="largest reference worker count of the sweep (0: min(32, CPUs in this process's affinity/quota))")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--batched-worlds", type=int, default=4096, help="worlds of the batched workload (0: skip)")
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--rain-steps", type=int, default=1000, help="steps of the rain workload (0: skip)")
    ap.add_argument("--rain-cpu-steps", type=int, default=1000, help="CPU-baseline sample of the rain (steps from the start)")
    ap.add_argument("--push-worlds", type=int, default=256, help="worlds of the differentiable push (0: skip)")
    ap.add_argument("--push-steps", type=int, default=60)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB megabytes 10 times (scales rocprofv3 FETCH_SIZE and "
                         "WRITE_SIZE to bytes) and exit")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only step the pyramid warmup + steps times (the child run the counter passes profile)")
    ap.add_argument("--no-traffic", action="store_true", help="skip the live rocprofv3 counter passes")
    args = ap.parse_args()

    if args.copy_calibration > 0:
        # a device-to-device copy through the HIP runtime (one copy kernel dispatch each)
        hip = ctypes.CDLL("libamdhip64.so")
        n = args.copy_calibration * (1 << 20)
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(n)):
            sys.exit("hipMalloc failed")
        hip.hipMemset(src, 1, ctypes.c_size_t(n))
        for _ in range(10):
            hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        hip.hipFree(src)
        hip.hipFree(dst)
        print(json.dumps({"copy_calibration_bytes_per_dispatch": n, "dispatches": 10}))
        return

    if args.traffic_probe:
        import box2d_amd

        scenes = box2d_amd.scenes()
        h = scenes.pyramid(args.rows, args.ground, sleep=False)
        scenes.step(h, args.warmup + args.steps, DT, SUBSTEPS)
        box2d_amd.library().b2amd_Synchronize()
        scenes.destroy(h)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported. A single Box2D world "
              "does not shard across GPUs; independent worlds would be split by world range.", file=sys.stderr)
        sys.exit(2)

    import box2d_amd
    from box2d_amd import WorldId

    lib = box2d_amd.library()
    scenes = box2d_amd.scenes()
    scenes.lib.b2s_world_id.restype = ctypes.c_uint32

    h = scenes.pyramid(args.rows, args.ground, sleep=False)
    bodies = scenes.body_count(h)
    packed = scenes.lib.b2s_world_id(h)
    wid = WorldId(packed & 0xFFFF, packed >> 16)

    scenes.step(h, args.warmup, DT, SUBSTEPS)
    lib.b2amd_Synchronize()
    t0 = time.perf_counter()
    # the steps run back to back in the scene driver, which reads each step's b2Profile off the clock
    _, profile = scenes.step_profiled(h, args.steps, DT, SUBSTEPS)
    lib.b2amd_Synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = bodies * SUBSTEPS * args.steps / elapsed
    # host-side phases of the timed steps (b2Profile, ms/step): pairs = broadphase pair update
    # (host tree query or GPU LBVH round trip); collide = collide launch to the end of touch
    # processing (GPU narrowphase + readback + prepareConstraints, the host touch/coloring work);
    # solverSetup = uploads and solver lists; solve = solve launch to the end of the step's host
    # work (constraints = the GPU solve window incl. the host work overlapped with it, then
    # transforms, refit and sleepIslands on the host)
    phases = {k: profile[k] for k in ("step", "pairs", "solverSetup", "collide", "prepareConstraints", "solve",
                                      "constraints", "splitIslands", "transforms", "refit", "sleepIslands")}

    # the reference's published window (benchmark/main.c: one untimed step, then 499 timed)
    hw = scenes.pyramid(args.rows, args.ground, sleep=False)
    scenes.step(hw, 1, DT, SUBSTEPS)
    lib.b2amd_Synchronize()
    t1 = time.perf_counter()
    scenes.step(hw, 499, DT, SUBSTEPS)
    lib.b2amd_Synchronize()
    long_s = (time.perf_counter() - t1) / 499
    long_hash = scenes.hash(hw)
    scenes.destroy(hw)
    long_want = golden_hash(args.rows, args.ground, 500)
    window_2_500 = {"ms_per_step": long_s * 1e3, "body_substeps_per_s": bodies * SUBSTEPS / long_s,
                    "vs_published": (bodies * SUBSTEPS / long_s) / PUBLISHED_BODY_SUBSTEPS if args.rows == 100 else None,
                    "hash_match": None if long_want is None else bool(long_hash == long_want)}

    # the same window with the move events read after every step (b2World_GetBodyEvents), as a
    # caller that draws or syncs transforms does: the transforms come back from the GPU per step
    class BodyEvents(ctypes.Structure):
        _fields_ = [("moveEvents", ctypes.c_void_p), ("moveCount", ctypes.c_int)]

    lib.b2World_GetBodyEvents.restype = BodyEvents
    lib.b2World_GetBodyEvents.argtypes = [WorldId]
    he = scenes.pyramid(args.rows, args.ground, sleep=False)
    pe = scenes.lib.b2s_world_id(he)
    wide = WorldId(pe & 0xFFFF, pe >> 16)
    scenes.step(he, args.warmup, DT, SUBSTEPS)
    lib.b2amd_Synchronize()
    moved = 0
    t2 = time.perf_counter()
    for _ in range(args.steps):
        scenes.step(he, 1, DT, SUBSTEPS)
        moved += lib.b2World_GetBodyEvents(wide).moveCount
    lib.b2amd_Synchronize()
    events_ms = (time.perf_counter() - t2) * 1e3 / args.steps
    scenes.destroy(he)
    with_body_events = {"ms_per_step": events_ms, "move_events_per_step": moved / args.steps,
                        "note": "same window, b2World_GetBodyEvents read after every step from Python"}

    total_steps = args.warmup + args.steps
    want = golden_hash(args.rows, args.ground, total_steps)
    got = scenes.hash(h)
    hash_match = None if want is None else bool(got == want)
    counters = scenes.counters(h)

    # per-kernel timing pass on a fresh copy of the scene (same steps as the timed region)
    h2 = scenes.pyramid(args.rows, args.ground, sleep=False)
    p2 = scenes.lib.b2s_world_id(h2)
    wid2 = WorldId(p2 & 0xFFFF, p2 >> 16)
    scenes.step(h2, args.warmup, DT, SUBSTEPS)
    lib.b2amd_World_EnableKernelTiming(wid2, ctypes.c_bool(True))
    lib.b2amd_World_ResetKernelStats(wid2)
    scenes.step(h2, min(args.profile_steps, args.steps), DT, SUBSTEPS)
    stats = kernel_stats(lib, wid2)
    lib.b2amd_World_EnableKernelTiming(wid2, ctypes.c_bool(False))
    scenes.destroy(h2)

    dominant = max(stats, key=lambda k: k["total_ms"]) if stats else None
    roofline = None
    if dominant is not None:
        avg_ms = dominant["total_ms"] / dominant["launches"]
        bytes_per_launch = dominant["bytes"] / dominant["launches"]
        achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                    "kernel": dominant["name"],
                    "avg_launch_us": avg_ms * 1e3, "algorithmic_bytes_per_launch": bytes_per_launch}
    kernels = [{"name": k["name"], "launches_per_step": k["launches"] / min(args.profile_steps, args.steps),
                "avg_us": 1e3 * k["total_ms"] / k["launches"],
                "gbs": (k["bytes"] / (k["total_ms"] * 1e-3) / 1e9) if k["total_ms"] > 0 else 0.0} for k in stats]
    gpu_busy_ms = sum(k["total_ms"] for k in stats) / max(1, min(args.profile_steps, args.steps))

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(32, host_cores())
        cpu = cpu_baseline(args.rows, args.ground, args.warmup, args.steps, threads, args.cpu_seconds)

    scenes.destroy(h)

    batched = None
    if args.batched_worlds > 0:
        batched = batched_workload(lib, scenes, args.batched_worlds, args.batched_steps)

    lib.b2amd_World_GetLastPairQuery.restype = ctypes.c_int
    rain = None
    if args.rain_steps > 0:
        threads = args.cpu_threads or min(32, host_cores())
        rain = rain_workload(lib, scenes, args.rain_steps, min(args.rain_cpu_steps, args.rain_steps), threads, args.no_cpu)

    push = None
    if args.push_worlds > 0:
        push = push_workload(lib, args.push_worlds, args.push_steps, min(32, args.push_worlds))

    # HBM traffic of the dominant kernel: rocprofv3 counter passes on child runs of the same
    # window, after every timed workload so the profiler perturbs none of them
    if roofline is not None and not args.no_traffic:
        traffic, details = live_traffic(roofline["kernel"], args.rows, args.ground, args.warmup, args.steps)
        roofline["traffic"] = traffic
        roofline["traffic_counters"] = details
        if traffic is not None:
            roofline["traffic_over_algorithmic"] = traffic / roofline["algorithmic_bytes_per_launch"]

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / PUBLISHED_BODY_SUBSTEPS if (args.rows == 100) else None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"large_pyramid: {args.rows}-row box pyramid ({bodies} bodies) on a "
                               f"{args.ground:g}x1 m ground, plain single-world GPU step",
                   "bodies": bodies, "substeps": SUBSTEPS, "dt": DT, "sleep": False,
                   "contacts": counters["contacts"], "colors": [c for c in counters["colors"] if c],
                   "baseline": "reference large_pyramid 8T SSE2 on Ryzen 9 7950X, 41.6M body-substeps/s"},
        "hash_match": hash_match,
        "hash": f"{got:#010x}",
        "roofline": roofline,
        "host_phases_ms_per_step": phases,
        "gpu_busy_ms_per_step": gpu_busy_ms,
        "spin_wait": bool(lib.b2amd_SpinWaitActive()),
        "window_steps_2_500": window_2_500,
        "with_body_events": with_body_events,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "batched": batched,
        "rain": rain,
        "push_adjoint": push,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
