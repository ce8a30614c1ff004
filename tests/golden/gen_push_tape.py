"""Write tests/golden/push_tape.npz: the step records of one small differentiable push world and
the GPU adjoint's gradient for it (needs the built library and a GPU).

World: a 4-row box pyramid (10 boxes) on the 40 x 1 m ground with a 1 m pusher box launched at it,
(vx, vy, omega, friction) = (4.5, 0.2, 0.5, 0.7), 40 steps of 1/60 s with 4 substeps; loss
|top box center - (0.3, 3.2)|^2. The CPU tests replay these records with the oracle (float replay
bit for bit, float64 reverse mode and central differences) without a GPU.

    python tests/golden/gen_push_tape.py [output.npz]
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

PARAMS = np.array([[4.5, 0.2, 0.5, 0.7]], dtype=np.float32)
GOAL = np.array([[0.3, 3.2]])
ROWS = 4
STEPS = 40


def main():
    from adjoint_ref import WorldRecord
    from box2d_amd.adjoint import BODY_STATE, PushWorlds

    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "push_tape.npz")
    pw = PushWorlds(PARAMS, rows=ROWS)
    tape = pw.record(STEPS)
    final = tape.final_states()
    top, pusher = int(pw.body_ids[0][-2]), int(pw.body_ids[0][-1])
    fg = np.zeros((1, tape.body_stride), dtype=BODY_STATE)
    fg[0]["center"][top] = 2.0 * (final[0]["center"][top] - GOAL[0])
    ig, fr, stats = tape.backward(fg)
    assert stats.unsupported == 0 and stats.replayMismatches == 0, (stats.unsupported, stats.replayMismatches)
    shape = pw.pusher_shape_ids()[0]
    grad = np.array([ig[0]["linearVelocity"][pusher][0], ig[0]["linearVelocity"][pusher][1],
                     ig[0]["angularVelocity"][pusher], fr[0, shape]], dtype=np.float64)
    rec = WorldRecord.from_tape(tape, 0)
    rec.save(out, params=PARAMS, goal=GOAL, top_body=np.array([top]), pusher_body=np.array([pusher]),
             pusher_shape=np.array([shape]), gpu_grad=grad,
             final_center=np.array(final[0]["center"][top], dtype=np.float32))
    print("wrote", out, "grad", grad)
    pw.destroy()


if __name__ == "__main__":
    main()
