"""The torch.autograd.Function wrapper of a taped rollout (b2g.autograd.Rollout).

Forward: the final states equal the world's own states after the same steps; backward: the
gradients equal b2g_tape_backward's for the same cotangent (to rounding: the reverse pass sums
adjoints with float atomics), and agree with central differences
(h = 1e-3, f64) of the host frozen model on the recorded structure within 1e-3 of each world's
gradient scale, on a 30-step rollout of 8 worlds of 16 bodies (parameters where the central
differences at h and h / 10 disagree, a clamp switching inside the step, and worlds whose f64
re-simulation leaves the recorded manifold features are counted and left out).
"""
import numpy as np
import pytest

from adjoint_model import fd_check_world, tape_structure

pytestmark = pytest.mark.gpu


def test_rollout_function_forward_backward(oracle):
    import torch
    import b2g
    from b2g import scenes
    from b2g.autograd import Rollout, state6
    batch = scenes.differentiable_worlds(8, 16)
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    w.add_batched(batch.body_defs, batch.body_world, batch.fixture_body, batch.shape_defs, batch.fixture_defs)
    s0 = state6(w.body_states())
    mass = w.mass()
    state0 = torch.tensor(s0, dtype=torch.float64, requires_grad=True)
    fric = torch.tensor(batch.fixture_defs["friction"], dtype=torch.float64, requires_grad=True)
    steps = 30
    final = Rollout.apply(w, state0, fric, steps)
    assert final.dtype == torch.float64 and final.shape == (len(s0), 6)
    np.testing.assert_array_equal(final.detach().numpy().astype(np.float32), state6(w.body_states()))
    b0 = torch.as_tensor(batch.meta["body0"])
    goal = torch.as_tensor(batch.meta["goal"])
    loss = ((final[b0, 0:2] - goal) ** 2).sum()
    loss.backward()
    # the same cotangent through the C ABI directly
    cot = np.zeros((len(s0), 6), np.float32)
    cot[batch.meta["body0"], 0:2] = 2.0 * (final.detach().numpy()[batch.meta["body0"], 0:2] - batch.meta["goal"])
    g, gf = w.tape_backward(cot)
    # Two backward passes over one tape: the reverse pass sums per-body and per-fixture adjoints
    # with float atomics, whose order varies between launches, so they agree to rounding (a few
    # ulps of the largest gradient), not bit for bit.
    for got, ref in ((state0.grad.numpy(), g), (fric.grad.numpy(), gf)):
        scale = max(1e-6, float(np.abs(ref).max()))
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-5 * scale)
    # against f64 central differences of the frozen model
    struct = tape_structure(w)
    rng = np.random.default_rng(7)
    res = [fd_check_world(oracle, batch, wi, mass, s0, struct, g, gf, rng) for wi in range(8)]
    print(res)
    checked = [r for r in res if r["branch"] and r["smooth"]]
    assert len(checked) >= 4, res
    assert sum(r["smooth"] for r in checked) >= 40
    for r in checked:
        assert r["rel_err_smooth"] <= 1e-3, r
