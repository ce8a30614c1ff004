"""Adjoint of recorded steps on the GPU (b2amd_Tape_*), checked against the CPU replay oracle.

Workload: differentiable push worlds, each a 10-row pyramid of 1 m boxes with a 1 m pusher box
launched at it. Parameters (vx, vy, omega, friction) ~ U([2,6], [-0.5,0.5], [-1,1], [0.3,0.9]);
loss = |top box center - goal|^2 after the recorded steps. Checks:
  - the record is a faithful checkpoint: the oracle's float replay reproduces every recorded step
    bit for bit, and the GPU backward's own recompute reports no mismatch;
  - GPU adjoint vs the oracle's float64 reverse-mode adjoint and vs central differences (h = 1e-4)
    of the float64 replay, all with the recorded run's branch pattern held fixed (active contact
    set, bias branches, normal and friction clamps), to 1e-3 relative;
  - the torch.autograd.Function returns the same gradient as the tape.
"""

import numpy as np
import pytest

from adjoint_ref import WorldRecord

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3


def _push_params(count, seed):
    rng = np.random.default_rng(seed)
    params = np.stack([rng.uniform(2, 6, count), rng.uniform(-0.5, 0.5, count), rng.uniform(-1, 1, count),
                       rng.uniform(0.3, 0.9, count)], 1)
    goals = np.stack([rng.uniform(-1, 1, count), rng.uniform(8.5, 9.5, count)], 1)
    return params, goals


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


@pytest.fixture(scope="module")
def recorded():
    from box2d_amd.adjoint import BODY_STATE, PushWorlds

    params, goals = _push_params(6, 2024)
    pw = PushWorlds(params)
    tape = pw.record(60)
    final = tape.final_states()
    fg = np.zeros((len(params), tape.body_stride), dtype=BODY_STATE)
    for w in range(len(params)):
        top = int(pw.body_ids[w][-2])
        fg[w]["center"][top] = 2.0 * (final[w]["center"][top] - goals[w])
    ig, fr, stats = tape.backward(fg)
    yield pw, tape, goals, final, ig, fr, stats
    pw.destroy()


def test_tape_records_and_backward_recompute_are_exact(recorded):
    pw, tape, goals, final, ig, fr, stats = recorded
    assert tape.steps == 60
    assert stats.unsupported == 0, stats.unsupported
    assert stats.replayMismatches == 0
    assert stats.worlds == len(goals)
    # the final records hold the worlds' states
    for w in range(len(goals)):
        assert np.all(np.isfinite(final[w]["center"][pw.body_ids[w]]))


def test_oracle_float_replay_reproduces_records(recorded):
    pw, tape, goals, *_ = recorded
    for w in range(2):
        rec = WorldRecord.from_tape(tape, w)
        top = rec.slot_of(int(pw.body_ids[w][-2]))
        r = rec.replay(0, top, goals[w])
        assert r["float_mismatches"] == 0
        assert r["structure_mismatches"] == 0


def test_gpu_adjoint_matches_fp64_adjoint_and_fd(recorded):
    pw, tape, goals, final, ig, fr, stats = recorded
    shape_ids = pw.pusher_shape_ids()
    for w in range(4):
        pusher = int(pw.body_ids[w][-1])
        top = int(pw.body_ids[w][-2])
        g_gpu = np.array([ig[w]["linearVelocity"][pusher][0], ig[w]["linearVelocity"][pusher][1],
                          ig[w]["angularVelocity"][pusher], fr[w, shape_ids[w]]], dtype=np.float64)
        rec = WorldRecord.from_tape(tape, w)
        ps, ts = rec.slot_of(pusher), rec.slot_of(top)
        g_rev, _ = rec.push_gradients(ps, shape_ids[w], ts, goals[w])
        g_fd, stable, _ = rec.push_fd(ps, shape_ids[w], ts, goals[w])
        print(f"[gpu] push world {w}: gpu {g_gpu} fp64 adjoint {g_rev} fd {g_fd} contact-stable {stable}", flush=True)
        assert _rel(g_rev, g_fd) < 1e-4, (w, g_rev, g_fd)
        assert _rel(g_gpu, g_rev) < REL_TOL, (w, g_gpu, g_rev)
        assert _rel(g_gpu, g_fd) < REL_TOL, (w, g_gpu, g_fd)


def test_gpu_adjoint_matches_fd_of_the_unforced_forward():
    """Central differences of the plain float64 forward, with no branch forced: each perturbed run
    makes its own contact, bias and clamp decisions. Eight single-box targets are pushed for 20
    steps. A world counts as contact-stable when its unforced perturbed runs keep the recorded
    run's branch signature and land where the runs forced onto the recorded pattern land (the two
    difference quotients agree to 1e-3 relative), so that no decision of the forward with a visible
    effect changed within the step h = 1e-6. In every contact-stable world the GPU gradient must
    equal the unforced differences to 1e-3 relative, and at least two of the eight worlds must be
    contact-stable (on MI355X three are: the others cross a decision the signature does not hold,
    and there the unforced differences leave the forced ones by 1e-3 to 2x)."""
    from box2d_amd.adjoint import BODY_STATE, PushWorlds

    rng = np.random.default_rng(8)
    params = np.stack([rng.uniform(2, 6, 8), rng.uniform(-0.5, 0.5, 8), rng.uniform(-1, 1, 8),
                       rng.uniform(0.3, 0.9, 8)], 1)
    goals = np.array([[0.0, 2.0]] * 8)
    pw = PushWorlds(params, rows=1)
    try:
        tape = pw.record(20)
        final = tape.final_states()
        fg = np.zeros((8, tape.body_stride), dtype=BODY_STATE)
        for w in range(8):
            top = int(pw.body_ids[w][-2])
            fg[w]["center"][top] = 2.0 * (final[w]["center"][top] - goals[w])
        ig, fr, stats = tape.backward(fg)
        assert stats.unsupported == 0
        shape_ids = pw.pusher_shape_ids()
        stable_worlds = []
        for w in range(8):
            pusher = int(pw.body_ids[w][-1])
            top = int(pw.body_ids[w][-2])
            g_gpu = np.array([ig[w]["linearVelocity"][pusher][0], ig[w]["linearVelocity"][pusher][1],
                              ig[w]["angularVelocity"][pusher], fr[w, shape_ids[w]]], dtype=np.float64)
            rec = WorldRecord.from_tape(tape, w)
            ps, ts = rec.slot_of(pusher), rec.slot_of(top)
            g_free, same_signature, _ = rec.push_fd(ps, shape_ids[w], ts, goals[w], h=1e-6, fixed=False)
            g_forced, _, _ = rec.push_fd(ps, shape_ids[w], ts, goals[w], h=1e-6, fixed=True)
            stable = same_signature and np.linalg.norm(g_free) > 0 and _rel(g_free, g_forced) < 1e-3
            print(f"[gpu] push world {w}: gpu {g_gpu} unforced fd {g_free} forced fd {g_forced} contact-stable {stable}",
                  flush=True)
            if stable:
                stable_worlds.append(w)
                assert _rel(g_gpu, g_free) < REL_TOL, (w, g_gpu, g_free)
        assert len(stable_worlds) >= 2, stable_worlds
    finally:
        pw.destroy()


def test_torch_function_gradient():
    import torch

    from box2d_amd.adjoint import push_simulation

    params_np, goals = _push_params(3, 99)
    params = torch.tensor(params_np, dtype=torch.float32, requires_grad=True)
    centers = push_simulation(params, 30)
    goal = torch.tensor(goals, dtype=torch.float32)
    loss = ((centers[:, -2, :] - goal) ** 2).sum()
    loss.backward()
    g = params.grad.numpy()
    assert np.all(np.isfinite(g))

    # same gradient through the tape API directly
    from box2d_amd.adjoint import BODY_STATE, PushWorlds

    pw = PushWorlds(params_np)
    tape = pw.record(30)
    final = tape.final_states()
    fg = np.zeros((3, tape.body_stride), dtype=BODY_STATE)
    for w in range(3):
        top = int(pw.body_ids[w][-2])
        fg[w]["center"][top] = 2.0 * (final[w]["center"][top] - goals[w])
    ig, fr, _ = tape.backward(fg)
    sids = pw.pusher_shape_ids()
    for w in range(3):
        pusher = int(pw.body_ids[w][-1])
        ref = [ig[w]["linearVelocity"][pusher][0], ig[w]["linearVelocity"][pusher][1],
               ig[w]["angularVelocity"][pusher], fr[w, sids[w]]]
        np.testing.assert_allclose(g[w], ref, rtol=1e-6, atol=1e-6)
    pw.destroy()
