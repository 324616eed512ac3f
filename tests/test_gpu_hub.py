"""A body with more contacts than the colouring's masks hold (127 colours).

A dynamic platform on the ground carries 200 small boxes: its 201 contacts need 201 colours. The
colouring gives the contacts it cannot place the highest colour, where two constraints may share
the platform, and that colour is solved one constraint after another (the reference solves every
constraint that way, b2_contact_solver_private.rs:268-730). Checked:
  * the oracle's island solver, replaying one SOLVE phase in the GPU's constraint order, agrees
    within 1e-5 (states and stored impulses), after 5 and after 30 steps;
  * 128 colours, every one but the last conflict-free, and the last holding the overflow;
  * over 120 steps the boxes rest on the platform and the platform on the ground, their cores apart
    by the polygon skins within the allowed penetration (no capacity error, nothing sinking or
    launched), with sleeping on and off;
  * a tape refuses such a step with B2G_ERR_UNSUPPORTED (the adjoint replays colours in parallel).
"""
import numpy as np
import pytest

import b2g
from b2g import scenes
from test_gpu_solver import DT, solver_replay_check

pytestmark = pytest.mark.gpu

N_BOXES = 200
PLATFORM_TOP = 0.5 + 0.5  # ground top 0.5, platform half height 0.25 on it: top at 1.0 (cores)


def hub_scene(n=N_BOXES):
    nb = 2 + n
    bd = np.zeros(nb, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[1:]["type"] = 2
    bd[1]["position"] = (0.0, 0.75)
    xs = (np.arange(n) - 0.5 * (n - 1)) * 0.8
    bd[2:]["position"] = np.stack([xs, np.full(n, PLATFORM_TOP + 0.15 + 0.01)], 1)
    sd = np.zeros(nb, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (100.0, 0.5)
    sd[1]["params"][:2] = (0.4 * n + 1.0, 0.25)
    sd[2:]["params"][:, :2] = (0.15, 0.15)
    fd = np.zeros(nb, b2g.FIXTURE_DEF)
    fd["friction"] = 0.6
    fd["restitution_threshold"] = 1.0
    fd["category"] = 1
    fd["mask"] = 0xFFFF
    fd["density"] = 1.0
    return scenes.Scene(bd, np.arange(nb, dtype=np.int32), sd, fd, {"name": "hub"})


@pytest.mark.parametrize("warm_steps", [5, 30])
def test_hub_replay_and_colours(oracle, warm_steps):
    sc = hub_scene()
    info = solver_replay_check(oracle, sc, warm_steps)
    order, offs, cb, touch = info["order"], info["offsets"], info["contact_bodies"], info["touch"]
    assert info["constraints"] == N_BOXES + 1
    assert len(offs) - 1 == 128
    pos = {int(k): i for i, k in enumerate(touch)}
    btype = sc.body_defs["type"]
    for c in range(len(offs) - 2):
        seen = set()
        for k in order[offs[c]:offs[c + 1]]:
            for b in cb[pos[int(k)]]:
                if btype[b] == 2:
                    assert b not in seen, "colour %d uses dynamic body %d twice" % (c, b)
                    seen.add(b)
    assert offs[-1] - offs[-2] >= N_BOXES + 1 - 127


@pytest.mark.parametrize("sleep", [False, True])
def test_hub_rests(sleep):
    sc = hub_scene()
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=sleep, continuous=True, block_solve=True)
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    for _ in range(120):
        w.step(DT, 8, 3)
    st = w.body_states()
    # resting shapes keep their cores apart by at most the two polygon skins (2 x 0.01 m) and at
    # least the skins less the allowed penetration (3 linear slops)
    gap_platform = float(st["c"][1, 1]) - (0.5 + 0.25)
    gaps = st["c"][2:, 1] - (float(st["c"][1, 1]) + 0.25 + 0.15)
    for g in (np.array([gap_platform]), gaps):
        assert (g > 0.02 - 3 * 0.005).all() and (g < 0.02 + 1e-4).all(), (g.min(), g.max())
    assert np.abs(st["v"][1:]).max() < 0.05
    w.close()


def test_tape_refuses_the_overflow_colour():
    sc = hub_scene()
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    w.tape_begin(4)
    with pytest.raises(b2g.B2gError):
        for _ in range(5):
            w.step(DT, 8, 3)
    w.close()
