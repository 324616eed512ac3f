"""torch.autograd.Function over a recorded rollout of a b2g world.

``Rollout.apply(world, state0, friction, steps)`` sets the world's body states from ``state0``
((bodies, 6): c.x, c.y, a, v.x, v.y, w; positions and angles also set the transforms and sweeps)
and its fixture frictions from ``friction`` ((fixtures,)), steps the world ``steps`` times with a
tape open (body-state checkpoints every ``checkpoint_every`` steps) and returns the final states
((bodies, 6)). The forward saves the tape; ``backward`` takes the cotangent of the final states and
returns dL/d state0 and dL/d friction through b2g_tape_backward, the adjoint of the frozen step:
the recorded contact set, manifold types and feature ids, solve order and position-iteration
counts are held fixed and everything continuous is differentiated.

The tensors may live on any device: the C ABI takes host buffers, so they are copied to the host
(float32) and the results copied back to the input's device and dtype. The world needs
continuous physics and sleeping off (a tape refuses them). box2d-rs has no derivatives; this
is the interface for callers that optimise through the simulation.
"""
import numpy as np
import torch

from . import BODY_STATE


def _states_from(world, s6):
    """Body-state records from (n, 6) rows: sweep and transform at the given centre and angle."""
    st = world.body_states()
    mass = world.mass()
    s6 = np.asarray(s6, np.float64)
    out = np.zeros(len(st), BODY_STATE)
    out[:] = st
    a = s6[:, 2].astype(np.float32)
    c = s6[:, 0:2].astype(np.float32)
    s, co = np.sin(a.astype(np.float64)).astype(np.float32), np.cos(a.astype(np.float64)).astype(np.float32)
    lc = mass[:, 4:6].astype(np.float32)
    # p = c - R(a) * local_centre
    p = np.stack([c[:, 0] - (co * lc[:, 0] - s * lc[:, 1]), c[:, 1] - (s * lc[:, 0] + co * lc[:, 1])], 1)
    out["c"], out["a"], out["c0"], out["a0"] = c, a, c, a
    out["p"], out["q"] = p, np.stack([s, co], 1)
    out["v"], out["w"] = s6[:, 3:5].astype(np.float32), s6[:, 5].astype(np.float32)
    return out


def state6(states):
    s = np.zeros((len(states), 6), np.float32)
    s[:, 0:2], s[:, 2], s[:, 3:5], s[:, 5] = states["c"], states["a"], states["v"], states["w"]
    return s


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, world, state0, friction, steps, dt=1.0 / 60.0, velocity_iterations=8,
                position_iterations=3, checkpoint_every=16):
        if world.flags.get("continuous") or world.flags.get("allow_sleep"):
            raise ValueError("Rollout: the world needs continuous physics and sleeping off (set_flags)")
        world.set_body_states(_states_from(world, state0.detach().cpu().numpy()))
        world.set_fixture_frictions(friction.detach().cpu().numpy().astype(np.float32))
        world.tape_begin(int(checkpoint_every))
        try:
            for _ in range(int(steps)):
                world.step(dt, velocity_iterations, position_iterations)
        finally:
            world.tape_end()
        final = state6(world.body_states())
        ctx.world = world
        ctx.meta = (state0.device, state0.dtype, friction.device, friction.dtype)
        return torch.from_numpy(final).to(device=state0.device, dtype=state0.dtype)

    @staticmethod
    def backward(ctx, grad_final):
        sdev, sdt, fdev, fdt = ctx.meta
        cot = grad_final.detach().cpu().numpy().astype(np.float32)
        g, gf = ctx.world.tape_backward(cot)
        gs = torch.from_numpy(np.ascontiguousarray(g, np.float32)).to(device=sdev, dtype=sdt)
        gfr = torch.from_numpy(np.ascontiguousarray(gf, np.float32)).to(device=fdev, dtype=fdt)
        return None, gs, gfr, None, None, None, None, None
