"""GPU adjoint against f64 central differences at the differentiable-rollout workload's size:
1024 worlds of 64 boxes and circles (20 x 20 m boxes, per-body friction U[0.2, 1.0]), taped with
body-state checkpoints every 16 steps, loss |c_body0 - goal|^2 per world.

On 32 sampled worlds the GPU's f32 gradients with respect to body 0's initial velocity, five
random other initial-velocity components and four random frictions are compared with central
differences (h = 1e-3) of the host f64 frozen model on the GPU's recorded structure; they agree
within 1e-3 of the world's gradient scale at every smooth parameter. Not compared, and counted:
worlds whose f64 re-simulation leaves the recorded manifold features (its derivative is then of
another function) and parameters where the central differences at h and h / 10 disagree (a clamp
switching inside the step). The horizon is 32 steps: by 60 steps the f64 re-simulation leaves the
recorded features in every sampled world, so central differences no longer differentiate the
recorded run.
"""
import numpy as np
import pytest

from adjoint_model import fd_check_world, state6, tape_structure

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)
HORIZON = 32


def test_adjoint_against_central_differences_1024_worlds(oracle):
    import b2g
    from b2g import scenes
    batch = scenes.differentiable_worlds(1024, 64)
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    w.add_batched(batch.body_defs, batch.body_world, batch.fixture_body, batch.shape_defs, batch.fixture_defs)
    s0 = state6(w.body_states())
    mass = w.mass()
    w.tape_begin(16)
    for _ in range(HORIZON):
        w.step(DT, 8, 3)
    w.tape_end()
    fin = state6(w.body_states())
    cot = np.zeros((len(fin), 6), np.float32)
    b0 = batch.meta["body0"]
    cot[b0, 0:2] = 2.0 * (fin[b0, 0:2] - batch.meta["goal"])
    g, gf = w.tape_backward(cot)
    struct = tape_structure(w)
    rng = np.random.default_rng(2026)
    sample = rng.choice(1024, 32, replace=False)
    res = [fd_check_world(oracle, batch, int(wi), mass, s0, struct, g, gf, rng) for wi in sample]
    followed = [r for r in res if r["branch"]]
    smooth = sum(r["smooth"] for r in followed)
    errs = [r["rel_err_smooth"] for r in followed if r["smooth"]]
    print("worlds following the recorded features: %d of 32; smooth parameters %d of %d; largest relative "
          "error %.3g; median %.3g" % (len(followed), smooth, sum(r["params"] for r in followed), max(errs),
                                       float(np.median(errs))), flush=True)
    assert len(followed) >= 16
    assert smooth >= 150
    for r in followed:
        if r["smooth"]:
            assert r["rel_err_smooth"] <= 1e-3, r
    w.close()
