"""ctypes wrapper of the benchmark-scene driver (oracle/bench_scenes.c). Test infrastructure.

The driver runs the reference's own benchmark scenes (shared/benchmarks.c with shared/human.c
and shared/utils.c) against the library it was linked with: oracle/_ref/libb2ref_bench.so is
the reference library, oracle/_ref/libb2amd_bench.so is box2d_amd's.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BENCH = os.path.join(HERE, "_ref", "libb2ref_bench.so")
AMD_BENCH = os.path.join(HERE, "_ref", "libb2amd_bench.so")


class BenchScenes:
    def __init__(self, path):
        self.lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        self.lib.bs_scene_name.restype = ctypes.c_char_p
        self.lib.bs_create.restype = ctypes.c_uint32
        self.lib.bs_create_workers.restype = ctypes.c_uint32
        self._out = (ctypes.c_uint32 * 3)()
        self.names = [self.lib.bs_scene_name(i).decode() for i in range(self.lib.bs_scene_count())]

    def steps(self, name):
        return self.lib.bs_scene_steps(self.names.index(name))

    def create(self, name, workers=1):
        """Builds the scene in a fresh world; returns (index1, generation) of the world id."""
        packed = self.lib.bs_create_workers(self.names.index(name), workers)
        return packed & 0xFFFF, packed >> 16

    def run(self, count):
        """count benchmark steps with no per-step reduction (for timing)."""
        self.lib.bs_run(count)

    def step(self):
        """One benchmark step; returns (move count, awake body count, sorted move-event hash)."""
        self.lib.bs_step(self._out)
        return tuple(int(x) for x in self._out)

    def moves(self, capacity=1 << 16):
        buf = np.zeros(7 * capacity, dtype=np.uint32)
        n = self.lib.bs_moves(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), buf.size)
        return buf[: 7 * n].reshape(n, 7)

    def destroy(self):
        self.lib.bs_destroy()
