"""Test-side binding of the CPU replay oracle for the adjoint (oracle/adjoint_oracle.cpp).

Packs a recorded world (every step record, the final state, the shapes and the static bodies) into
the flat arrays the oracle takes, and runs its float, float64 and reverse-mode float64 replays.
Used by the adjoint tests only.
"""

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "_build", "libb2adjoint_oracle.so")

SHAPE = np.dtype([("type", "<i4"), ("padding", "<i4"), ("vertices", "<f4", 16), ("normals", "<f4", 16),
                  ("centroid", "<f4", 2), ("radius", "<f4"), ("count", "<i4")])
assert SHAPE.itemsize == 152

_lib = None


def oracle():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} is missing: run `make -C oracle oracle`")
        lib = ctypes.CDLL(ORACLE_PATH)
        vp = ctypes.c_void_p
        lib.b2o_adjoint_replay.restype = ctypes.c_int
        lib.b2o_adjoint_replay.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int, vp,
                                           ctypes.c_int, vp, vp, ctypes.c_int, vp, vp, vp, vp, vp, vp]
        lib.b2o_adjoint_decisions.argtypes = [ctypes.c_int]
        lib.b2o_adjoint_decisions_info.argtypes = [ctypes.POINTER(ctypes.c_int)] * 4
        lib.b2o_adjoint_capture_step_grads.argtypes = [vp]
        _lib = lib
    return _lib


class WorldRecord:
    """All records of one world in the oracle's layout."""

    def __init__(self, headers, bodies, body_start, cons, cons_start, shapes, friction, statics):
        self.headers = np.ascontiguousarray(headers)
        self.bodies = np.ascontiguousarray(bodies)
        self.body_start = np.ascontiguousarray(body_start, dtype=np.int32)
        self.cons = np.ascontiguousarray(cons)
        self.cons_start = np.ascontiguousarray(cons_start, dtype=np.int32)
        self.shapes = np.ascontiguousarray(shapes)
        self.friction = np.ascontiguousarray(friction, dtype=np.float32)
        self.statics = np.ascontiguousarray(statics)
        self.steps = len(self.headers) - 1

    @classmethod
    def from_tape(cls, tape, world):
        from box2d_amd.adjoint import TAPE_BODY

        steps = tape.lib.b2amd_Tape_GetStepCount(tape.handle, world)
        headers, bodies, cons = [], [], []
        body_start, cons_start = [], []
        nb_total = nc_total = 0
        for k in range(steps + 1):
            h, b, c = tape.read_step(world, k)
            headers.append(h)
            body_start.append(nb_total)
            bodies.append(b)
            nb_total += len(b)
            if k < steps:
                cons_start.append(nc_total)
                cons.append(c)
                nc_total += len(c)
        shapes = np.zeros(tape.shape_stride, dtype=SHAPE)
        friction = np.zeros(tape.shape_stride, dtype=np.float32)
        for s in range(tape.shape_stride):
            r = tape.shape(world, s)
            if r is None:
                continue
            raw = bytes(r[0])
            shapes[s] = np.frombuffer(raw, dtype=SHAPE)[0]
            friction[s] = r[2]
        statics = np.zeros(tape.body_stride, dtype=TAPE_BODY)
        for b in range(tape.body_stride):
            st = tape.body(world, b)
            if st is not None:
                statics[b] = st
        return cls(np.array(headers), np.concatenate(bodies), body_start, np.concatenate(cons), cons_start, shapes,
                   friction, statics)

    def save(self, path, **extra):
        np.savez_compressed(path, headers=self.headers, bodies=self.bodies, body_start=self.body_start, cons=self.cons,
                            cons_start=self.cons_start, shapes=self.shapes, friction=self.friction,
                            statics=self.statics, **extra)

    @classmethod
    def load(cls, npz, prefix=""):
        g = lambda k: npz[prefix + k]  # noqa: E731
        return cls(g("headers"), g("bodies"), g("body_start"), g("cons"), g("cons_start"), g("shapes"),
                   g("friction"), g("statics"))

    def initial_state(self):
        """float64 (bodies, 7) state of record 0: center, rotation, velocity, angular velocity."""
        nb = int(self.headers[0]["bodyCount"])
        b = self.bodies[:nb]
        return np.concatenate([b["center"], b["rotation"], b["linearVelocity"], b["angularVelocity"][:, None]],
                              axis=1).astype(np.float64)

    def slot_of(self, body_id):
        nb = int(self.headers[0]["bodyCount"])
        ids = list(self.bodies[:nb]["bodyId"])
        return ids.index(body_id) + 1

    def replay(self, mode, target_slot, goal, state=None, friction=None, decisions=0):
        """mode 0 float, 1 float64, 2 float64 with reverse-mode gradients. decisions: 0 free
        branches, 1 record the run's branch pattern, 2 force the recorded pattern. Returns a dict."""
        lib = oracle()
        lib.b2o_adjoint_decisions(decisions)
        nb = int(self.headers[0]["bodyCount"])
        state = self.initial_state() if state is None else np.ascontiguousarray(state, dtype=np.float64)
        fr = self.friction.astype(np.float64) if friction is None else np.ascontiguousarray(friction, dtype=np.float64)
        goal = np.ascontiguousarray(goal, dtype=np.float64)
        loss = np.zeros(1, dtype=np.float64)
        gstate = np.zeros((nb, 7), dtype=np.float64)
        gfr = np.zeros(len(self.shapes), dtype=np.float64)
        counts = np.zeros(2, dtype=np.int32)
        sig = np.zeros(1, dtype=np.uint64)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        lib.b2o_adjoint_replay(mode, self.steps, p(self.headers), p(self.bodies), p(self.body_start), p(self.cons),
                               p(self.cons_start), p(self.shapes), p(self.friction), len(self.shapes), p(self.statics),
                               len(self.statics), p(state), p(fr), target_slot, p(goal), p(loss), p(gstate), p(gfr),
                               p(counts), p(sig))
        info = [ctypes.c_int() for _ in range(4)]
        lib.b2o_adjoint_decisions_info(*[ctypes.byref(x) for x in info])
        if decisions != 1:
            lib.b2o_adjoint_decisions(0)
        return {"loss": float(loss[0]), "grad_state": gstate, "grad_friction": gfr, "float_mismatches": int(counts[0]),
                "structure_mismatches": int(counts[1]), "signature": int(sig[0]), "decisions": info[0].value,
                "flips": info[1].value, "overrun": bool(info[2].value), "consumed": info[3].value}

    def record_pattern(self, target_slot, goal):
        """Record the branch pattern of the float replay (the recorded forward's pattern)."""
        r = self.replay(0, target_slot, goal, decisions=1)
        oracle().b2o_adjoint_decisions(0)
        return r["decisions"]

    def push_gradients(self, pusher_slot, pusher_shape, target_slot, goal, fixed=True):
        """Reverse-mode float64 gradient of the loss with respect to (vx, vy, omega, friction).
        fixed: hold the recorded forward's branch pattern (active set, clamps) fixed."""
        if fixed:
            self.record_pattern(target_slot, goal)
        r = self.replay(2, target_slot, goal, decisions=2 if fixed else 0)
        g = r["grad_state"][pusher_slot - 1]
        return np.array([g[4], g[5], g[6], r["grad_friction"][pusher_shape]]), r

    def push_fd(self, pusher_slot, pusher_shape, target_slot, goal, h=1e-4, fixed=True):
        """Central differences (float64 replay) of the loss in (vx, vy, omega, friction). With
        fixed, every run is forced onto the recorded forward's branch pattern; stable says whether
        the perturbed runs' own values took that pattern too (a contact-stable run)."""
        dec = 0
        if fixed:
            self.record_pattern(target_slot, goal)
            dec = 2
        base = self.replay(1, target_slot, goal, decisions=dec)
        grads = np.zeros(4)
        stable = (base["flips"] == 0) if fixed else True
        for k in range(4):
            vals = []
            for sgn in (1.0, -1.0):
                state = self.initial_state()
                fr = self.friction.astype(np.float64)
                if k < 3:
                    state[pusher_slot - 1, 4 + k] += sgn * h
                else:
                    fr[pusher_shape] += sgn * h
                r = self.replay(1, target_slot, goal, state, fr, decisions=dec)
                if fixed:
                    stable = stable and r["flips"] == 0 and not r["overrun"]
                else:
                    stable = stable and r["signature"] == base["signature"] and r["structure_mismatches"] == 0
                vals.append(r["loss"])
            grads[k] = (vals[0] - vals[1]) / (2.0 * h)
        return grads, stable, base
