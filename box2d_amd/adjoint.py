"""Differentiable stepping: step recording (tape) and the GPU adjoint, behind torch.autograd.

The native library records every step of the worlds on a tape (b2amd_Tape_Create): the body state
at the step's start and the step's contact constraints in solve order with the discrete decisions
the step took. b2amd_Tape_Backward recomputes each step's substeps on the GPU from its record and
runs the hand-written adjoint of the step in reverse (one workgroup per world). The active contact
set and every clamp pattern are those of the recorded forward.

``PushSimulation`` is the differentiable push workload: batched worlds, each a box pyramid with a
pusher box whose initial velocity, spin and friction are the parameters; the forward steps all
worlds together with b2amd_StepWorlds and returns the final centers of every scene body.
"""

import ctypes

import numpy as np

from . import WorldId, library, scenes

BODY_STATE = np.dtype([("center", "<f4", 2), ("rotation", "<f4", 2), ("linearVelocity", "<f4", 2),
                       ("angularVelocity", "<f4"), ("padding", "<f4")])

TAPE_STEP = np.dtype([("bodyCount", "<i4"), ("constraintCount", "<i4"), ("jointCount", "<i4"),
                      ("activeColorCount", "<i4"), ("colorStart", "<i4", 25), ("overflowStart", "<i4"),
                      ("overflowCount", "<i4"), ("subStepCount", "<i4"), ("timeStep", "<f4"), ("h", "<f4"),
                      ("inv_h", "<f4"), ("contactSpeed", "<f4"), ("maxLinearSpeed", "<f4"),
                      ("maxAngularSpeed", "<f4"), ("gravity", "<f4", 2), ("enableWarmStarting", "<i4"),
                      ("inv_dt", "<f4"), ("padding", "<i4", 2)])

TAPE_BODY = np.dtype([("bodyId", "<i4"), ("flags", "<u4"), ("center", "<f4", 2), ("origin", "<f4", 2),
                      ("rotation", "<f4", 2), ("linearVelocity", "<f4", 2), ("angularVelocity", "<f4"),
                      ("invMass", "<f4"), ("invInertia", "<f4"), ("gravityScale", "<f4"), ("linearDamping", "<f4"),
                      ("angularDamping", "<f4"), ("localCenter", "<f4", 2), ("force", "<f4", 2), ("torque", "<f4"),
                      ("padding", "<f4")])

TAPE_CONSTRAINT = np.dtype([("contactId", "<i4"), ("shapeIdA", "<i4"), ("shapeIdB", "<i4"), ("bodyIdA", "<i4"),
                            ("bodyIdB", "<i4"), ("slotA", "<i4"), ("slotB", "<i4"), ("pointCount", "<i4"),
                            ("info", "<u4"), ("id", "<u2", 2), ("normal", "<f4", 2), ("anchorA", "<f4", (2, 2)),
                            ("anchorB", "<f4", (2, 2)), ("separation", "<f4", 2), ("baseSeparation", "<f4", 2),
                            ("normalImpulse", "<f4", 2), ("tangentImpulse", "<f4", 2), ("friction", "<f4"),
                            ("restitution", "<f4"), ("rollingResistance", "<f4"), ("tangentSpeed", "<f4"),
                            ("biasRate", "<f4"), ("massScale", "<f4"), ("impulseScale", "<f4"),
                            ("cachedRotationA", "<f4", 2), ("cachedRotationB", "<f4", 2), ("padding", "<f4")])

assert BODY_STATE.itemsize == 32 and TAPE_STEP.itemsize == 176
assert TAPE_BODY.itemsize == 88 and TAPE_CONSTRAINT.itemsize == 160

UNSUPPORTED_BITS = {1: "overflow constraints", 2: "joints", 4: "restitution or rolling resistance",
                    8: "non-polygon manifold", 16: "speed cap", 32: "awake set changed", 64: "motion locks",
                    128: "continuous collision"}


class AdjointStats(ctypes.Structure):
    _fields_ = [("backwardMs", ctypes.c_float), ("replayMismatches", ctypes.c_int), ("unsupported", ctypes.c_int),
                ("worlds", ctypes.c_int), ("steps", ctypes.c_int)]


class Shape(ctypes.Structure):
    """b2amd_Shape with the polygon member of the union (b2Polygon: 8 vertices, 8 normals, centroid,
    radius, count)."""
    _fields_ = [("type", ctypes.c_int), ("padding", ctypes.c_int), ("vertices", ctypes.c_float * 16),
                ("normals", ctypes.c_float * 16), ("centroid", ctypes.c_float * 2), ("radius", ctypes.c_float),
                ("count", ctypes.c_int)]


def _ptr(a, ctype=ctypes.c_void_p):
    return a.ctypes.data_as(ctype)


def _bind(lib):
    if getattr(lib, "_tape_bound", False):
        return lib
    vp = ctypes.c_void_p
    lib.b2amd_Tape_Create.restype = vp
    lib.b2amd_Tape_Create.argtypes = [ctypes.POINTER(WorldId), ctypes.c_int]
    lib.b2amd_Tape_Finish.restype = ctypes.c_int
    lib.b2amd_Tape_Finish.argtypes = [vp]
    lib.b2amd_Tape_Destroy.argtypes = [vp]
    lib.b2amd_Tape_GetStepCount.restype = ctypes.c_int
    lib.b2amd_Tape_GetStepCount.argtypes = [vp, ctypes.c_int]
    lib.b2amd_Tape_GetBodyStride.restype = ctypes.c_int
    lib.b2amd_Tape_GetBodyStride.argtypes = [vp]
    lib.b2amd_Tape_GetShapeStride.restype = ctypes.c_int
    lib.b2amd_Tape_GetShapeStride.argtypes = [vp]
    lib.b2amd_Tape_ReadStep.restype = ctypes.c_int
    lib.b2amd_Tape_ReadStep.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int, vp, ctypes.c_int]
    lib.b2amd_Tape_GetFinalStates.restype = ctypes.c_int
    lib.b2amd_Tape_GetFinalStates.argtypes = [vp, vp]
    lib.b2amd_Tape_Backward.restype = ctypes.c_int
    lib.b2amd_Tape_Backward.argtypes = [vp, vp, vp, vp, ctypes.POINTER(AdjointStats)]
    lib.b2amd_Tape_GetBody.restype = ctypes.c_int
    lib.b2amd_Tape_GetBody.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    lib.b2amd_Tape_GetShape.restype = ctypes.c_int
    lib.b2amd_Tape_GetShape.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Shape),
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)]
    lib._tape_bound = True
    return lib


class Tape:
    """Recording of the steps of a set of worlds and the reverse pass over it."""

    def __init__(self, world_ids):
        self.lib = _bind(library())
        self.count = len(world_ids)
        arr = (WorldId * self.count)(*world_ids)
        self.handle = self.lib.b2amd_Tape_Create(arr, self.count)
        if not self.handle:
            raise RuntimeError("b2amd_Tape_Create failed (a world is already recorded on another tape)")
        self.steps = None

    def finish(self):
        self.steps = self.lib.b2amd_Tape_Finish(self.handle)
        self.body_stride = self.lib.b2amd_Tape_GetBodyStride(self.handle)
        self.shape_stride = self.lib.b2amd_Tape_GetShapeStride(self.handle)
        return self.steps

    def final_states(self):
        out = np.zeros((self.count, self.body_stride), dtype=BODY_STATE)
        self.lib.b2amd_Tape_GetFinalStates(self.handle, _ptr(out))
        return out

    def backward(self, final_grad):
        """final_grad: BODY_STATE array (worlds, body_stride). Returns (initial gradient in the same
        layout, shape friction gradient (worlds, shape_stride), AdjointStats)."""
        fg = np.ascontiguousarray(final_grad, dtype=BODY_STATE)
        assert fg.shape == (self.count, self.body_stride)
        ig = np.zeros_like(fg)
        fr = np.zeros((self.count, max(self.shape_stride, 1)), dtype=np.float32)
        stats = AdjointStats()
        rc = self.lib.b2amd_Tape_Backward(self.handle, _ptr(fg), _ptr(ig), _ptr(fr), ctypes.byref(stats))
        if rc != 0:
            raise RuntimeError("b2amd_Tape_Backward failed (tape not finished)")
        return ig, fr, stats

    def read_step(self, world, step):
        """(header, bodies, constraints) of one record; step == step count gives the final state."""
        h = np.zeros(1, dtype=TAPE_STEP)
        if self.lib.b2amd_Tape_ReadStep(self.handle, world, step, _ptr(h), None, 0, None, 0) != 0:
            raise IndexError(f"no record {step} for world {world}")
        nb, nc = int(h["bodyCount"][0]), int(h["constraintCount"][0])
        bodies = np.zeros(max(nb, 1), dtype=TAPE_BODY)
        cons = np.zeros(max(nc, 1), dtype=TAPE_CONSTRAINT)
        self.lib.b2amd_Tape_ReadStep(self.handle, world, step, _ptr(h), _ptr(bodies), nb, _ptr(cons), nc)
        return h[0], bodies[:nb], cons[:nc]

    def shape(self, world, shape_id):
        """(shape, body id, friction) of a shape of a recorded world, or None."""
        s = Shape()
        body = ctypes.c_int()
        friction = ctypes.c_float()
        if self.lib.b2amd_Tape_GetShape(self.handle, world, shape_id, ctypes.byref(s), ctypes.byref(body),
                                        ctypes.byref(friction)) != 0:
            return None
        return s, body.value, friction.value

    def body(self, world, body_id):
        """Current state of any body (static ones included) in TAPE_BODY form, or None."""
        out = np.zeros(1, dtype=TAPE_BODY)
        if self.lib.b2amd_Tape_GetBody(self.handle, world, body_id, _ptr(out)) != 0:
            return None
        return out[0]

    def destroy(self):
        if self.handle:
            self.lib.b2amd_Tape_Destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def unsupported_names(bits):
    return [name for bit, name in UNSUPPORTED_BITS.items() if bits & bit]


class PushWorlds:
    """Batched differentiable push worlds built from a (worlds, 4) parameter array: pusher
    velocity (vx, vy), spin and friction."""

    def __init__(self, params, rows=10):
        from . import step_worlds  # noqa: F401  (imported for the caller's convenience)

        self.scenes = scenes()
        self.params = np.asarray(params, dtype=np.float32)
        self.handles = [self.scenes.push(float(p[0]), float(p[1]), float(p[2]), float(p[3]), rows) for p in self.params]
        self.ids = [self.scenes.world_id(h) for h in self.handles]
        self.body_ids = [self.scenes.body_ids(h) for h in self.handles]
        self.tape = None

    def record(self, steps, dt=1.0 / 60.0, substeps=4):
        from . import step_worlds

        self.tape = Tape(self.ids)
        arr = (WorldId * len(self.ids))(*self.ids)
        step_worlds(arr, steps, dt, substeps)
        self.tape.finish()
        return self.tape

    def pusher_shape_ids(self):
        out = []
        for w, bids in enumerate(self.body_ids):
            pusher = int(bids[-1])
            sid = next(s for s in range(self.tape.shape_stride)
                       if (r := self.tape.shape(w, s)) is not None and r[1] == pusher)
            out.append(sid)
        return out

    def destroy(self):
        if self.tape is not None:
            self.tape.destroy()
            self.tape = None
        for h in self.handles:
            self.scenes.destroy(h)
        self.handles = []


def _torch_function():
    import torch

    class PushSimulation(torch.autograd.Function):
        """params (worlds, 4) -> final centers (worlds, bodies, 2) of every scene body after
        `steps` steps. Backward runs the GPU adjoint of the recorded steps."""

        @staticmethod
        def forward(ctx, params, steps, rows):
            worlds = PushWorlds(params.detach().cpu().numpy(), rows)
            tape = worlds.record(steps)
            final = tape.final_states()
            nb = len(worlds.body_ids[0])
            centers = np.stack([final[w]["center"][worlds.body_ids[w]] for w in range(len(worlds.ids))])
            ctx.worlds = worlds
            ctx.nb = nb
            return torch.from_numpy(centers.astype(np.float32)).to(params.device)

        @staticmethod
        def backward(ctx, grad_centers):
            worlds = ctx.worlds
            tape = worlds.tape
            g = grad_centers.detach().cpu().numpy().astype(np.float32)
            fg = np.zeros((tape.count, tape.body_stride), dtype=BODY_STATE)
            for w in range(tape.count):
                fg[w]["center"][worlds.body_ids[w]] = g[w]
            ig, fr, stats = tape.backward(fg)
            ctx.stats = stats
            shape_ids = worlds.pusher_shape_ids()
            grad = np.zeros((tape.count, 4), dtype=np.float32)
            for w in range(tape.count):
                pusher = int(worlds.body_ids[w][-1])
                grad[w, 0:2] = ig[w]["linearVelocity"][pusher]
                grad[w, 2] = ig[w]["angularVelocity"][pusher]
                grad[w, 3] = fr[w, shape_ids[w]]
            worlds.destroy()
            if stats.unsupported:
                raise RuntimeError("adjoint: unsupported features in the recorded steps: "
                                   + ", ".join(unsupported_names(stats.unsupported)))
            return torch.from_numpy(grad).to(grad_centers.device), None, None

    return PushSimulation


_PUSH = None


def push_simulation(params, steps=60, rows=10):
    """Differentiable push: final centers (worlds, bodies, 2) as a function of params (worlds, 4)."""
    global _PUSH
    if _PUSH is None:
        _PUSH = _torch_function()
    return _PUSH.apply(params, steps, rows)
