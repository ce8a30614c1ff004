"""CPU checks of the adjoint oracle on recorded steps (tests/golden/push_tape.npz, written on a GPU
by tests/golden/gen_push_tape.py from the product's forward and backward passes).

The gradient is that of the recorded run with its branch pattern held fixed (active contact set,
speculative and soft bias branches, normal and friction clamps): the oracle records the pattern of
its float replay (the recorded forward's) and forces it on its float64 runs.

- The oracle's float replay reproduces every recorded step bit for bit: the records are what the
  GPU forward computed (itself hash-identical to the reference), so this pins the oracle's
  restatement of the step.
- A record altered in one bit is caught (the check is sensitive).
- The oracle's float64 reverse-mode gradient equals central differences (h = 1e-4) of its
  float64 replay on the same branch pattern.
- The GPU adjoint's gradient stored with the records agrees with both to 1e-3.
"""

import os

import numpy as np
import pytest

from adjoint_ref import WorldRecord

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "push_tape.npz")

pytestmark = pytest.mark.skipif(not os.path.exists(FIXTURE), reason="push_tape.npz fixture not generated")


@pytest.fixture(scope="module")
def fixture():
    npz = np.load(FIXTURE)
    rec = WorldRecord.load(npz)
    return rec, npz


def _slots(rec, npz):
    return (rec.slot_of(int(npz["pusher_body"][0])), int(npz["pusher_shape"][0]), rec.slot_of(int(npz["top_body"][0])),
            npz["goal"][0])


def test_float_replay_is_bit_exact(fixture):
    rec, npz = fixture
    ps, sh, ts, goal = _slots(rec, npz)
    r = rec.replay(0, ts, goal)
    assert r["float_mismatches"] == 0
    assert r["structure_mismatches"] == 0
    # the replay's final top-box center is the recorded one
    dx = np.float32(np.sqrt(r["loss"]))
    ref = np.linalg.norm(npz["final_center"].astype(np.float64) - goal)
    assert abs(dx - ref) < 1e-5


def test_altered_record_is_detected(fixture):
    rec, npz = fixture
    ps, sh, ts, goal = _slots(rec, npz)
    bodies = rec.bodies.copy()
    k = int(rec.body_start[5])
    v = bodies[k]["linearVelocity"].view(np.uint32)
    v[0] ^= 1  # one ulp in one recorded velocity
    bodies[k]["linearVelocity"] = v.view(np.float32)
    bad = WorldRecord(rec.headers, bodies, rec.body_start, rec.cons, rec.cons_start, rec.shapes, rec.friction,
                      rec.statics)
    assert bad.replay(0, ts, goal)["float_mismatches"] > 0


def test_reverse_mode_matches_central_differences(fixture):
    rec, npz = fixture
    ps, sh, ts, goal = _slots(rec, npz)
    g_rev, r = rec.push_gradients(ps, sh, ts, goal)
    assert r["consumed"] == r["decisions"] and not r["overrun"]
    g_fd, _, _ = rec.push_fd(ps, sh, ts, goal, h=1e-4)
    np.testing.assert_allclose(g_rev, g_fd, rtol=1e-4, atol=1e-6 * np.abs(g_fd).max())


def test_gpu_gradient_matches_fp64_adjoint_and_fd(fixture):
    rec, npz = fixture
    ps, sh, ts, goal = _slots(rec, npz)
    g_rev, _ = rec.push_gradients(ps, sh, ts, goal)
    g_fd, _, _ = rec.push_fd(ps, sh, ts, goal, h=1e-4)
    g_gpu = npz["gpu_grad"]
    assert np.linalg.norm(g_gpu - g_rev) <= 1e-3 * np.linalg.norm(g_rev)
    assert np.linalg.norm(g_gpu - g_fd) <= 1e-3 * np.linalg.norm(g_fd)


def test_free_branches_differ_from_the_recorded_pattern(fixture):
    """Without the pattern held fixed, float64 takes some near-tie branches differently from the
    float32 forward; the gradient of that other smooth piece is not the recorded run's. This is why
    the comparisons above hold the recorded pattern (it also shows the forcing is effective)."""
    rec, npz = fixture
    ps, sh, ts, goal = _slots(rec, npz)
    g_fixed, r = rec.push_gradients(ps, sh, ts, goal)
    g_free, _ = rec.push_gradients(ps, sh, ts, goal, fixed=False)
    assert r["flips"] > 0
    assert np.linalg.norm(g_free - g_fixed) > 1e-3 * np.linalg.norm(g_fixed)
