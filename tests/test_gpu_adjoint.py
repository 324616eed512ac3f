"""GPU adjoint (b2g_tape_*) against the host f64 model of the same frozen run.

A batch of differentiable worlds (boxes and circles in 20 x 20 m boxes, per-body friction) is
stepped on the GPU with a tape open (body-state checkpoints every 16 steps). The recorded
structure -- contacts, manifold types and feature ids, solve order, carries and position-iteration
counts, read back with b2g_tape_export -- drives the host f64 model (tests/host/adjoint_check.cpp,
pinned against central differences in test_adjoint_host.py). The GPU's frozen replay must
reproduce its own forward run, and its gradients of the loss sum_w |c_body0(w) - goal_w|^2 with
respect to the initial velocities and the frictions must agree with the host f64 reverse pass.
"""
import numpy as np
import pytest

from adjoint_model import HostAdjoint, state6

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def _run(n_worlds, n, steps, every=16):
    import b2g
    from b2g import scenes
    batch = scenes.differentiable_worlds(n_worlds, n)
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)  # a tape needs continuous physics off
    w.add_batched(batch.body_defs, batch.body_world, batch.fixture_body, batch.shape_defs, batch.fixture_defs)
    s0 = w.body_states()
    w.tape_begin(every)
    for _ in range(steps):
        w.step(DT, 8, 3)
    w.tape_end()
    return batch, w, s0


def _cotangent(batch, final):
    cot = np.zeros((len(final), 6), np.float64)
    b0 = batch.meta["body0"]
    cot[b0, 0:2] = 2.0 * (final[b0, 0:2] - batch.meta["goal"])
    return cot


def test_gpu_tape_replay_reproduces_forward():
    batch, w, _ = _run(8, 24, 40)
    steps, slots = w.tape_info()
    assert steps == 40 and slots > 0
    fwd = state6(w.body_states())
    rep = w.tape_replay()
    dyn = batch.body_defs["type"] == 2
    err = np.abs(rep[dyn] - fwd[dyn]).max()
    assert err <= 1e-4, "frozen replay differs from the forward run by %g" % err


def test_gpu_adjoint_matches_host_f64(oracle, oracle64):
    """30 taped steps of 8 worlds of 24 bodies (checkpoints every 16): the GPU's f32 gradients
    against the host f64 reverse pass on the GPU's recorded structure, within 1e-3 of each
    world's gradient scale. Worlds whose f64 re-simulation leaves the recorded manifold features
    are counted, printed and left out (at most 2 of 8)."""
    batch, w, s0 = _run(8, 24, 30)
    final = state6(w.body_states())
    cot = _cotangent(batch, final)
    g, gf = w.tape_backward(cot.astype(np.float32))
    host = HostAdjoint(oracle, batch, w.mass())
    for k in range(w.tape_info()[0]):
        slots, imp, offs, params = w.tape_export(k)
        host.add_step(slots, imp[:, :4], params)
    hs0 = state6(s0)
    out, mism = host.forward(hs0)
    wid = batch.body_world
    # The f64 model recomputes every manifold from its own states. Where f32 and f64 runs fall on
    # different sides of a feature change (a contact near a corner), the model follows another
    # branch than the recording for that slot, and its world's gradients are of another function.
    # Such worlds are left out and counted.
    skip = set(np.unique(wid[host.mismatch_bodies()]).tolist())
    print("f64 re-simulation left the recorded branches in %d of 8 worlds: %s" % (len(skip), sorted(skip)))
    assert len(skip) <= 2, "f64 replay left the recorded branches in %d of 8 worlds" % len(skip)
    hg, hgf, _ = host.backward(hs0, _cotangent(batch, out))
    dyn = np.nonzero(batch.body_defs["type"] == 2)[0]
    # per world: initial-velocity gradients and friction gradients, relative to the world's scale
    fworld = wid[batch.fixture_body]
    for wi in range(batch.meta["worlds"]):
        if wi in skip:
            continue
        bsel = dyn[wid[dyn] == wi]
        a, b = g[bsel][:, 3:6].astype(np.float64), hg[bsel][:, 3:6]
        scale = max(1e-6, np.abs(b).max())
        assert np.abs(a - b).max() <= 1e-3 * scale, "world %d: velocity gradients differ" % wi
        fsel = np.nonzero(fworld == wi)[0]
        fa, fb = gf[fsel].astype(np.float64), hgf[fsel]
        fscale = max(1e-6, np.abs(fb).max())
        assert np.abs(fa - fb).max() <= 1e-3 * fscale + 1e-6, "world %d: friction gradients differ" % wi
