"""Adjoint pinned by central differences on the CPU.

The frozen step (recorded contact set, manifold types and feature ids, solve order and
position-iteration counts held fixed; everything continuous recomputed, including the manifold
geometry by collide) is modelled in f64 on the host from the library's templated step functions
(tests/host/adjoint_check.cpp); its reverse pass uses the library's vector-Jacobian products
(box2d-rs_amd/csrc/b2g_adjoint_t.h), the same code the GPU adjoint runs in f32. Central
differences of the f64 forward must agree with the reverse pass. The structure is taken from
the oracle's own run of one world of the differentiable-worlds scene (boxes and circles in a
20 x 20 m box, per-body friction), and the loss is |c_body0 - goal|^2 after the run.
"""
import os

import numpy as np
import pytest

from adjoint_model import HostAdjoint, LIB, oracle_structure, state6


def _setup(oracle, oracle64, steps, n=12, seed_world=0):
    if not os.path.exists(LIB):
        pytest.skip("tests/host not built")
    from b2g import scenes
    batch = scenes.differentiable_worlds(seed_world + 1, n)
    sc = scenes.world_of(batch, seed_world)
    ow = oracle64.world()
    ow.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    mass = ow.mass()
    state0 = state6(ow.bodies())
    model = HostAdjoint(oracle, sc, mass)
    for slots, lam0, params in oracle_structure(oracle64, sc, steps):
        model.add_step(slots, lam0, params)
    body0 = int(np.nonzero(sc.body_defs["type"] == 2)[0][0])
    goal = batch.meta["goal"][seed_world]
    return model, state0, body0, goal


def _loss(model, state0, body0, goal, friction=None):
    out, mism = model.forward(state0, friction)
    d = out[body0, 0:2] - goal
    return float(d @ d), out, mism


@pytest.mark.parametrize("steps", [20, 45])
def test_host_adjoint_matches_central_differences(oracle, oracle64, steps):
    model, s0, body0, goal = _setup(oracle, oracle64, steps)
    L, out, mism0 = _loss(model, s0, body0, goal)
    # the model solves in contact-list order without the island early exit, so over longer runs
    # its trajectory (and then a few manifolds' features) drifts from the oracle run the structure
    # was taken from; it remains one function, differentiated along its own branches
    assert mism0 <= (0 if steps <= 20 else 40)
    cot = np.zeros_like(out)
    cot[body0, 0:2] = 2.0 * (out[body0, 0:2] - goal)
    g, gf, _ = model.backward(s0, cot)
    # parameters: initial velocities (v.x, v.y, w) of every dynamic body and every friction
    dyn = np.nonzero(model.btype == 2)[0]
    params = [("state", b, j) for b in dyn for j in (3, 4, 5)]
    params += [("friction", f, 0) for f in range(model.nf)]
    rng = np.random.default_rng(steps)
    pick = rng.choice(len(params), size=min(40, len(params)), replace=False)
    checked, agreed, smooth = 0, 0, 0
    for idx in pick:
        kind, i, j = params[idx]

        def f(h):
            if kind == "state":
                s = s0.copy()
                s[i, j] += h
                Lh, _, m = _loss(model, s, body0, goal)
            else:
                fr = model.friction.copy()
                fr[i] += h
                Lh, _, m = _loss(model, s0, body0, goal, fr)
            return Lh, m

        an = g[i, j] if kind == "state" else gf[i]
        fd = {}
        for h in (1e-5, 1e-6):
            (lp, mp), (lm, mm) = f(h), f(-h)
            fd[h] = (lp - lm) / (2 * h) if mp == mism0 and mm == mism0 else None
        if fd[1e-5] is None or fd[1e-6] is None:
            continue
        checked += 1
        scale = max(1e-3, abs(fd[1e-5]))
        # a kink (a clamp switching between the two evaluations) shows as disagreeing step sizes
        if abs(fd[1e-5] - fd[1e-6]) > 1e-4 * scale:
            continue
        smooth += 1
        agreed += abs(an - fd[1e-5]) <= 1e-5 * scale + 1e-7
    assert checked >= 30
    assert smooth >= 0.8 * checked
    assert agreed == smooth, "%d of %d smooth parameters disagree with central differences" % (smooth - agreed, smooth)


def test_host_adjoint_friction_and_velocity_matter(oracle, oracle64):
    """The gradient is not trivially zero: body 0's initial velocity moves its final position,
    and friction of touching bodies has an effect after contacts form."""
    model, s0, body0, goal = _setup(oracle, oracle64, 45)
    L, out, _ = _loss(model, s0, body0, goal)
    cot = np.zeros_like(out)
    cot[body0, 0:2] = 2.0 * (out[body0, 0:2] - goal)
    g, gf, _ = model.backward(s0, cot)
    assert np.abs(g[body0, 3:5]).max() > 1e-3
    assert np.abs(gf).max() > 0.0


def test_world_slice_recovers_each_world(oracle64):
    """world_slice (which takes one world's part out of a batch's recorded structure) returns, for
    a batch assembled from three single-world structures, each of those structures exactly."""
    from b2g import scenes
    from adjoint_model import world_slice
    batch = scenes.differentiable_worlds(3, 10)
    per_world = [oracle_structure(oracle64, scenes.world_of(batch, wi), 12) for wi in range(3)]
    b_off = [int(np.nonzero(batch.body_world == wi)[0][0]) for wi in range(3)]
    f_off = [int(np.nonzero(np.isin(batch.fixture_body, np.nonzero(batch.body_world == wi)[0]))[0][0])
             for wi in range(3)]
    merged, seen = [], [0, 0, 0]
    total = 0
    starts = [[] for _ in range(3)]  # per world: the batch number of each of its slots
    for k in range(12):
        rows, lams = [], []
        for wi in range(3):
            slots, lam0, params = per_world[wi][k]
            s = slots.astype(np.int64).copy()
            s[:, 0:2] += b_off[wi]
            s[:, 2:4] += f_off[wi]
            for col in (5, 6):
                has = s[:, col] >= 0
                s[has, col] = 2 * np.asarray(starts[wi])[s[has, col] // 2] + s[has, col] % 2
            rows.append(s)
            lams.append(lam0)
        for wi in range(3):
            n = len(per_world[wi][k][0])
            starts[wi].extend(range(total, total + n))
            total += n
        merged.append((np.concatenate(rows).astype(np.int32), np.concatenate(lams), per_world[0][k][2]))
    for wi in range(3):
        got = world_slice(batch, merged, wi)
        for k in range(12):
            np.testing.assert_array_equal(got[k][0], per_world[wi][k][0])
            np.testing.assert_array_equal(got[k][1], per_world[wi][k][1])
