"""Continuous collision (the TOI pass) on the GPU against the oracle.

Scenes whose TOI events are independent (each event's island is one fast body and static
geometry, and the events have distinct times), so the order in which events are taken is the
same in the reference's list order and the GPU's contact-set order:
  * a fast ball shot at a thin static wall (2 m per step against 0.1 m): it stops in front of the
    wall with continuous physics on and passes through with it off, and every step's body state
    and event count agree with the oracle's (src/private/dynamics/b2_world.rs:534-901) within 1e-4;
  * twelve boxes falling at 60 m/s onto a thin static ground at different heights and angles: the
    same, over 40 steps, and all of them end resting on the ground;
  * fast spinning boxes onto a static edge and fast balls onto a kinematic platform moving up: the
    same comparison over 45 steps, the boxes above the edge and the balls on the platform;
  * sleeping and continuous physics on together: fast boxes landing among bouncing balls whose
    ground contacts separate in steps with TOI events; awake flags and sleep times compared
    exactly with the oracle's at every step;
  * a touching contact whose fat AABBs separate in a TOI sub-step (an asleep box launched into a
    ceiling off the slab it touches): the next collide wakes both bodies, as the reference's
    destroy does, and the slab's sleep time is reset as the oracle's;
  * fast boxes dropped into a static corner (a TOI island touching two static fixtures) and fast
    balls landing at distinct times on one moving kinematic platform (events sharing a kinematic
    body);
  * bullets (B2bodyDef::bullet, B2body::set_bullet): balls and small boxes at 100-142 m/s against
    free dynamic planks and two bullets meeting head-on, no gravity (each TOI island is one
    constraint): times of impact against dynamic bodies, states and event counts as the oracle's;
    without the flag they pass through;
  * determinism: two runs of a 40-box rain with sleeping on are bit-identical; a 1024-body mixed
    world runs 30 steps with continuous physics with finite states and nothing falling out;
  * sub-stepping (B2world::set_sub_stepping): one solved event per step, the pass resumed by the
    next steps without a solve; states, event counts and step_complete as the oracle's at every
    step (the rain, and the bouncing rain with sleeping), and every completed pass bit-identical
    to the GPU's pass without sub-stepping;
  * TOI islands beyond the event workgroup's LDS store: every island forced into the global store
    gives the same runs bit for bit, and a plank landing on 600 static fixtures (hundreds of
    island contacts) stops on them with the oracle's event counts and positions within the TOI
    position solve's 1.5-slop band.
"""
import numpy as np
import pytest

import b2g

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def _fixtures(n, friction=0.4):
    fd = np.zeros(n, b2g.FIXTURE_DEF)
    fd["friction"] = friction
    fd["restitution_threshold"] = 1.0
    fd["category"] = 1
    fd["mask"] = 0xFFFF
    fd["density"] = 1.0
    return fd


def wall_scene():
    bd = np.zeros(2, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[0]["position"] = (5.0, 0.0)
    bd[1]["type"] = 2
    bd[1]["linear_velocity"] = (120.0, 0.0)
    bd[1]["gravity_scale"] = 0.0
    sd = np.zeros(2, b2g.SHAPE_DEF)
    sd[0]["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (0.05, 2.0)
    sd[1]["type"] = b2g.SHAPE_CIRCLE
    sd[1]["radius"] = 0.1
    return bd, np.arange(2, dtype=np.int32), sd, _fixtures(2, 0.2), (0.0, 0.0)


def rain_scene(n=12, spacing=1.5):
    bd = np.zeros(1 + n, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[1:]["type"] = 2
    bd[1:]["position"] = np.stack([np.arange(n) * spacing - 0.5 * spacing * n, 3.0 + 0.37 * np.arange(n)], 1)
    bd[1:]["linear_velocity"] = (0.0, -60.0)
    bd[1:]["angle"] = np.linspace(-0.4, 0.4, n)
    sd = np.zeros(1 + n, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (spacing * n, 0.05)
    sd[1:]["params"][:, :2] = (0.3, 0.2)
    return bd, np.arange(1 + n, dtype=np.int32), sd, _fixtures(1 + n), (0.0, -10.0)


def _pair(oracle, scene, continuous, sleep=False):
    bd, fb, sd, fd, g = scene
    w = b2g.World(g)
    w.set_flags(warm_starting=True, allow_sleep=sleep, continuous=continuous, block_solve=True)
    w.add(bd, fb, sd, fd)
    o = oracle.world(g)
    o.set_flags(warm_starting=True, allow_sleep=sleep, block_solve=True, continuous=continuous)
    o.add(bd, fb, sd, fd)
    return w, o


def _compare(w, o, steps, tol=1e-4):
    for step in range(steps):
        w.step(DT, 8, 3)
        o.step(DT, 8, 3)
        assert w.profile()["toi_events"] == o.toi_events, step
        g, r = w.body_states(), o.bodies()
        for f in ("c", "a", "v", "w"):
            np.testing.assert_allclose(g[f], r[f], rtol=0, atol=tol, err_msg="step %d %s" % (step, f))
    return w.body_states()


@pytest.mark.parametrize("continuous", [False, True])
def test_fast_ball_and_thin_wall(oracle, continuous):
    w, o = _pair(oracle, wall_scene(), continuous)
    st = _compare(w, o, 6)
    x = float(st[1]["c"][0])
    if continuous:
        assert x < 5.0 - 0.05 - 0.1 + 0.02, x
    else:
        assert x > 5.05, x


@pytest.mark.parametrize("continuous", [False, True])
def test_fast_boxes_rain_on_thin_ground(oracle, continuous):
    w, o = _pair(oracle, rain_scene(), continuous)
    st = _compare(w, o, 40)
    y = st["c"][1:, 1]
    if continuous:
        assert (y > 0.05 + 0.2 - 0.02).all() and (y < 0.05 + 0.2 + 0.3).all(), y


def test_continuous_runs_are_deterministic():
    runs = []
    for _ in range(2):
        bd, fb, sd, fd, g = rain_scene(n=40, spacing=0.9)
        w = b2g.World(g)
        w.set_flags(warm_starting=True, allow_sleep=True, continuous=True, block_solve=True)
        w.add(bd, fb, sd, fd)
        events = 0
        for _ in range(60):
            w.step(DT, 8, 3)
            events += w.profile()["toi_events"]
        runs.append(w.body_states().tobytes())
        assert events >= 40
    assert runs[0] == runs[1]


def test_large_mixed_world_with_continuous():
    from b2g import scenes
    sc = scenes.large_world(n=1024, seed=4)
    ys = []
    for cont in (False, True):
        w = b2g.World()
        w.set_flags(warm_starting=True, allow_sleep=False, continuous=cont, block_solve=True)
        w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
        for _ in range(30):
            w.step(DT, 8, 3)
        st = w.body_states()
        assert np.isfinite(st["c"]).all()
        ys.append(st["c"][:, 1])
    dyn = sc.body_defs["type"] == 2
    # nothing falls out of the container either way
    assert (ys[1][dyn] > ys[0][dyn].min() - 1.0).all()


def edge_kinematic_scene():
    """Fast spinning boxes onto a static edge ground, and fast balls onto a kinematic platform
    moving up: times of impact with edges and with a moving kinematic body."""
    n = 6
    bd = np.zeros(2 + 2 * n, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[1]["type"] = 1  # kinematic platform
    bd[1]["position"] = (40.0, 0.0)
    bd[1]["linear_velocity"] = (0.0, 2.0)
    for i in range(n):
        b = bd[2 + i]
        b["type"] = 2
        b["position"] = (3.0 * i, 4.0 + 0.61 * i)
        b["linear_velocity"] = (1.5 * (i - 2), -70.0)
        b["angular_velocity"] = 9.0 - 3.0 * i
        c = bd[2 + n + i]
        c["type"] = 2
        c["position"] = (34.0 + 2.5 * i, 5.0 + 0.53 * i)
        c["linear_velocity"] = (0.0, -65.0)
    sd = np.zeros(2 + 2 * n, b2g.SHAPE_DEF)
    sd[0]["type"] = b2g.SHAPE_EDGE
    sd[0]["params"][:4] = (-10.0, 0.0, 25.0, 0.0)
    sd[1]["type"] = b2g.SHAPE_BOX
    sd[1]["params"][:2] = (9.0, 0.05)
    sd[2:2 + n]["type"] = b2g.SHAPE_BOX
    sd[2:2 + n]["params"][:, :2] = (0.4, 0.15)
    sd[2 + n:]["type"] = b2g.SHAPE_CIRCLE
    sd[2 + n:]["radius"] = 0.12
    return bd, np.arange(2 + 2 * n, dtype=np.int32), sd, _fixtures(2 + 2 * n), (0.0, -10.0)


def test_edges_and_kinematic_bodies(oracle):
    w, o = _pair(oracle, edge_kinematic_scene(), True)
    st = _compare(w, o, 45)
    n = 6
    platform_y = float(st[1]["c"][1])
    assert (st["c"][2:2 + n, 1] > 0.0).all(), st["c"][2:2 + n, 1]  # above the edge
    assert (st["c"][2 + n:, 1] > platform_y).all(), (st["c"][2 + n:, 1], platform_y)  # on the platform


def test_tape_refuses_continuous_physics():
    bd, fb, sd, fd, g = wall_scene()
    w = b2g.World(g)
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=True, block_solve=True)
    w.add(bd, fb, sd, fd)
    with pytest.raises(b2g.B2gError):
        w.tape_begin(4)
    w.step(DT, 8, 3)  # the world still steps


@pytest.mark.parametrize("continuous", [False, True])
def test_reference_style_api_set_continuous_physics(continuous):
    """B2world::set_continuous_physics through the reference-style API: a 0.1 m ball at 120 m/s
    against a 0.1 m thick static wall stops at it only with continuous physics."""
    world = b2g.B2world((0.0, 0.0))
    world.set_continuous_physics(continuous)
    wall = world.create_body(b2g.B2bodyDef(position=(5.0, 0.0)))
    box = b2g.B2polygonShape()
    box.set_as_box(0.05, 2.0)
    wall.create_fixture_by_shape(box, 0.0)
    ball = world.create_body(b2g.B2bodyDef(body_type=b2g.B2_DYNAMIC_BODY, linear_velocity=(120.0, 0.0)))
    ball.create_fixture(b2g.B2fixtureDef(shape=b2g.B2circleShape(radius=0.1), density=1.0, friction=0.2))
    for _ in range(6):
        world.step(1.0 / 60.0, 8, 3)
    x = float(ball.get_position().x)
    assert (x < 4.9) if continuous else (x > 5.05), x


def _compare_sleep(w, o, steps, tol=1e-4):
    """Per step: body states within tol, awake flags and sleep times exactly, TOI event counts
    exactly; returns (states, steps with TOI events and a destroyed touching contact)."""
    both = 0
    for step in range(steps):
        w.step(DT, 8, 3)
        o.step(DT, 8, 3)
        assert w.profile()["toi_events"] == o.toi_events, step
        g, r = w.body_states(), o.bodies()
        for f in ("c", "a", "v", "w"):
            np.testing.assert_allclose(g[f], r[f], rtol=0, atol=tol, err_msg="step %d %s" % (step, f))
        np.testing.assert_array_equal(g["awake"], r["awake"], err_msg="step %d awake" % step)
        np.testing.assert_array_equal(g["sleep_time"], r["sleep_time"], err_msg="step %d sleep_time" % step)
        if o.toi_events and len(w.destroyed_contacts()[0]):
            both += 1
    return w.body_states(), both


def bounce_rain_scene(n_fast=24, n_balls=20):
    """Fast boxes (60 m/s) landing on a thin static ground one after another over about 40 steps,
    among balls with restitution that leave the same ground upward at different speeds and bounce
    on it (their touching contacts with it separate and are destroyed in steps that also have TOI
    events), with sleeping on: everything comes to rest and falls asleep."""
    n = n_fast + n_balls
    bd = np.zeros(1 + n, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[1:]["type"] = 2
    xs = np.arange(n) * 1.1 - 0.55 * n
    bd[1:1 + n_fast]["position"] = np.stack([xs[:n_fast], 2.0 + 1.7 * np.arange(n_fast)], 1)
    bd[1:1 + n_fast]["linear_velocity"] = (0.0, -60.0)
    bd[1 + n_fast:]["position"] = np.stack([xs[n_fast:], np.full(n_balls, 0.05 + 0.2)], 1)
    bd[1 + n_fast:]["linear_velocity"] = np.stack([np.zeros(n_balls), 1.5 + 0.15 * np.arange(n_balls)], 1)
    sd = np.zeros(1 + n, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (0.6 * n + 2.0, 0.05)
    sd[1:1 + n_fast]["params"][:, :2] = (0.25, 0.15)
    sd[1 + n_fast:]["type"] = b2g.SHAPE_CIRCLE
    sd[1 + n_fast:]["radius"] = 0.2
    fd = _fixtures(1 + n)
    fd["restitution"][1 + n_fast:] = 0.6
    return bd, np.arange(1 + n, dtype=np.int32), sd, fd, (0.0, -10.0)


def test_sleep_with_continuous_matches_oracle(oracle):
    """Sleeping and continuous physics on together (B2world::new's defaults): every step's states,
    awake flags, sleep times and TOI event counts agree with the oracle while fast boxes land and
    balls bounce, and the scene ends asleep."""
    w, o = _pair(oracle, bounce_rain_scene(), True, sleep=True)
    st, _ = _compare_sleep(w, o, 150)
    assert (st["awake"][1:] == 0).sum() > 0


def launch_scene(n_fast=20):
    """A box floating with its core 0.018 m above a floating slab (no gravity on either: their
    manifold touches, no impulse flows, both fall asleep), a thin static ceiling 0.5 m above the
    box, and a rain of fast boxes landing on a static ground one after another from about step 38
    on (TOI events every step). Each floating body has one constraint, so the solve order, which
    differs between the GPU's colours and the reference's island order, does not matter."""
    m = 4 + n_fast
    bd = np.zeros(m, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[1]["position"] = (-6.0, 0.918 + 0.3 + 0.5 + 0.05)  # static ceiling
    bd[2:]["type"] = 2
    bd[2]["position"] = (-6.0, 0.5)  # slab, half height 0.1
    bd[3]["position"] = (-6.0, 0.918)  # box, half height 0.3
    bd[2:4]["gravity_scale"] = 0.0
    t = 38.0 / 60.0 + np.arange(n_fast) * 0.6 / 60.0  # arrival times
    bd[4:]["position"] = np.stack([np.arange(n_fast) * 0.9, 0.3 + 60.0 * t + 5.0 * t * t], 1)
    bd[4:]["linear_velocity"] = (0.0, -60.0)
    sd = np.zeros(m, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (30.0, 0.05)
    sd[1]["params"][:2] = (1.0, 0.05)
    sd[2]["params"][:2] = (0.5, 0.1)
    sd[3]["params"][:2] = (0.3, 0.3)
    sd[4:]["params"][:, :2] = (0.25, 0.15)
    return bd, np.arange(m, dtype=np.int32), sd, _fixtures(m), (0.0, -10.0)


def test_touching_contact_destroyed_in_a_toi_step(oracle):
    """The asleep box is launched upward at 60 m/s (awake, as B2body::set_awake(true) with
    set_linear_velocity leave it). In the next step it hits the ceiling in a TOI event, whose
    sub-step moves its fat AABB off the slab's while their contact, not updated by the event (both
    bodies dynamic), still touches: the GPU's rebuild of the contact set after the events drops it
    and hands it to the next collide, which wakes both bodies, as the reference's collide destroys
    it (src/private/dynamics/b2_contact_manager.rs:146-153, src/private/dynamics/b2_contact.rs:33-45)
    and its set_awake(true) resets the slab's sleep time (src/b2_body.rs:783-801). Every step's
    states, awake flags, sleep times and TOI event counts agree with the oracle's."""
    w, o = _pair(oracle, launch_scene(), True, sleep=True)
    _compare_sleep(w, o, 42)
    g, r = w.body_states(), o.bodies()
    assert (g["awake"][2:4] == 0).all()  # slab and box sleep
    for st in (g, r):
        st["v"][3] = (0.0, 60.0)
        st["awake"][3] = 1
        st["sleep_time"][3] = 0.0
    w.set_body_states(g)
    o.set_bodies(r)
    o.set_awake(3, True)
    _compare_sleep(w, o, 1)
    assert o.toi_events >= 1
    fab, _ = w.destroyed_contacts()
    assert (2, 3) in {tuple(sorted(map(int, p))) for p in fab}, fab
    _compare_sleep(w, o, 1)
    # the slab's sleep time: reset by the wake, then one step of rest (2 h without the wake)
    assert w.body_states()["sleep_time"][2] == DT
    _compare_sleep(w, o, 10)


def corner_scene():
    """Fast frictionless balls dropped into corners of a static floor and static walls, each ball
    pressed against its wall while it falls: in its TOI event with the floor it touches both static
    fixtures. (Frictionless balls: the wall's and the floor's normals are orthogonal and pass
    through the centre, so the two constraints on a ball do not interact and the solve order,
    which differs between the GPU's colours and the reference's island order, does not matter.)"""
    n = 5
    bd = np.zeros(1 + 2 * n, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    sd = np.zeros(1 + 2 * n, b2g.SHAPE_DEF)
    sd[0]["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (20.0, 0.05)
    for i in range(n):
        x0 = -10.0 + 4.0 * i
        bd[1 + i]["position"] = (x0 - 0.05, 3.0)  # wall, its right face at x0
        sd[1 + i]["type"] = b2g.SHAPE_BOX
        sd[1 + i]["params"][:2] = (0.05, 3.0)
        b = bd[1 + n + i]
        b["type"] = 2
        b["position"] = (x0 + 0.2 - 0.002, 2.0 + 0.9 * i)
        b["linear_velocity"] = (-0.3, -55.0 - 3.0 * i)
        sd[1 + n + i]["type"] = b2g.SHAPE_CIRCLE
        sd[1 + n + i]["radius"] = 0.2
    return bd, np.arange(1 + 2 * n, dtype=np.int32), sd, _fixtures(1 + 2 * n, 0.0), (0.0, -10.0)


def test_toi_island_touching_several_static_fixtures(oracle):
    w, o = _pair(oracle, corner_scene(), True)
    st = _compare(w, o, 50)
    n = 5
    balls = st["c"][1 + n:]
    walls = -10.0 + 4.0 * np.arange(n)
    assert (balls[:, 1] > 0.05 + 0.2 - 0.02).all() and (balls[:, 0] > walls + 0.2 - 0.02).all(), balls


def shared_platform_scene():
    """Two groups of fast balls landing at distinct times on one kinematic platform moving up
    and sideways: their TOI events share the platform, whose sweep is advanced by each event."""
    n = 6
    bd = np.zeros(1 + n, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[0]["type"] = 1
    bd[0]["linear_velocity"] = (0.7, 1.5)
    bd[1:]["type"] = 2
    bd[1:]["position"] = np.stack([-3.0 + 1.2 * np.arange(n), 2.5 + 0.83 * np.arange(n)], 1)
    bd[1:]["linear_velocity"] = (0.0, -70.0)
    sd = np.zeros(1 + n, b2g.SHAPE_DEF)
    sd[0]["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (6.0, 0.05)
    sd[1:]["type"] = b2g.SHAPE_CIRCLE
    sd[1:]["radius"] = 0.1
    return bd, np.arange(1 + n, dtype=np.int32), sd, _fixtures(1 + n), (0.0, -10.0)


def test_events_sharing_a_kinematic_platform(oracle):
    w, o = _pair(oracle, shared_platform_scene(), True)
    st = _compare(w, o, 40)
    assert (st["c"][1:, 1] > st[0]["c"][1]).all(), (st["c"][1:, 1], st[0]["c"][1])


def bullet_scene(n=6):
    """No gravity: n free dynamic planks (0.1 x 2 m) at different heights and distances, a fast
    bullet (balls and small boxes in turn, 100-142 m/s) aimed at each, and two bullets flying at
    each other. Every contact is between two bodies that touch nothing else, so each TOI island
    is one constraint and the solve order does not matter."""
    m = 2 * n + 2
    bd = np.zeros(m, b2g.BODY_DEF)
    bd["type"] = 2
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    sd = np.zeros(m, b2g.SHAPE_DEF)
    for i in range(n):
        bd[i]["position"] = (5.0 + 1.3 * i, 3.0 * i)
        sd[i]["type"] = b2g.SHAPE_BOX
        sd[i]["params"][:2] = (0.05, 1.0)
        b = bd[n + i]
        b["position"] = (0.0, 3.0 * i + 0.2 * (i % 3))
        b["linear_velocity"] = (100.0 + 7.0 * i, 0.0)
        b["angular_velocity"] = 2.0 * (i - n / 2)
        b["bullet"] = 1
        if i % 2:
            sd[n + i]["type"] = b2g.SHAPE_BOX
            sd[n + i]["params"][:2] = (0.12, 0.08)
        else:
            sd[n + i]["type"] = b2g.SHAPE_CIRCLE
            sd[n + i]["radius"] = 0.1
    for k, (x, v) in enumerate(((-3.0, 90.0), (3.0, -80.0))):
        b = bd[2 * n + k]
        b["position"] = (x, -4.0)
        b["linear_velocity"] = (v, 0.0)
        b["bullet"] = 1
        sd[2 * n + k]["type"] = b2g.SHAPE_CIRCLE
        sd[2 * n + k]["radius"] = 0.15
    fd = _fixtures(m, 0.3)
    return bd, np.arange(m, dtype=np.int32), sd, fd, (0.0, 0.0)


def test_bullets_against_dynamic_bodies(oracle):
    """Bullets get times of impact against dynamic bodies (src/private/dynamics/b2_world.rs:616-621):
    every step's states and TOI event counts agree with the oracle's (every bullet hits its plank,
    the two bullets meet), and without the flag they pass through on both."""
    scene = bullet_scene()
    n = 6
    w, o = _pair(oracle, scene, True)
    events = 0
    for _ in range(12):
        _compare(w, o, 1)
        events += o.toi_events
    assert events >= n + 1  # every bullet hits its plank, and the two bullets meet
    bd = scene[0].copy()
    bd["bullet"] = 0
    w2, o2 = _pair(oracle, (bd,) + tuple(scene[1:]), True)
    st2 = _compare(w2, o2, 12)
    assert (st2["c"][n:2 * n, 0] > st2["c"][:n, 0]).all()  # all passed through


def test_reference_style_bullet():
    """B2bodyDef::bullet and B2body::set_bullet through the reference-style API."""
    def run(flag, late):
        world = b2g.B2world((0.0, 0.0))
        plank = world.create_body(b2g.B2bodyDef(body_type=b2g.B2_DYNAMIC_BODY, position=(5.0, 0.0)))
        box = b2g.B2polygonShape()
        box.set_as_box(0.05, 1.0)
        plank.create_fixture_by_shape(box, 1.0)
        ball = world.create_body(b2g.B2bodyDef(body_type=b2g.B2_DYNAMIC_BODY, linear_velocity=(120.0, 0.0),
                                               bullet=flag and not late))
        circle = b2g.B2circleShape()
        circle.m_radius = 0.1
        ball.create_fixture_by_shape(circle, 1.0)
        world.step(1.0 / 60.0, 8, 3)  # 2 m: still short of the plank
        if late:
            ball.set_bullet(flag)
        assert ball.is_bullet() == flag
        for _ in range(5):
            world.step(1.0 / 60.0, 8, 3)
        return float(ball.get_position().x), float(plank.get_position().x)
    for late in (False, True):
        x, px = run(True, late)
        assert x < px, (x, px)
        x, px = run(False, late)
        assert x > px, (x, px)


@pytest.mark.parametrize("scene", ["rain", "bounce"])
def test_sub_stepping_matches_oracle(oracle, scene):
    """B2world::set_sub_stepping: each step solves at most one TOI event and the next steps resume
    the pass without a solve until a pass finds no event. Every step's states, event counts and
    step_complete agree with the oracle's (awake flags and sleep times exactly with sleeping on);
    and each completed sub-stepped pass leaves the GPU world equal, bit for bit, to a GPU world
    stepped without sub-stepping, with the same events."""
    sc, sleep, passes = (rain_scene(), False, 30) if scene == "rain" else (bounce_rain_scene(), True, 120)
    w, o = _pair(oracle, sc, True, sleep=sleep)
    full = b2g.World(sc[4])
    full.set_flags(warm_starting=True, allow_sleep=sleep, continuous=True, block_solve=True)
    full.add(*sc[:4])
    w.set_sub_stepping(True)
    o.set_sub_stepping(True)
    assert w.sub_stepping() == (True, True)
    resumed = 0
    for p in range(passes):
        full.step(DT, 8, 3)
        events = 0
        while True:
            w.step(DT, 8, 3)
            o.step(DT, 8, 3)
            ev = w.profile()["toi_events"]
            assert ev == o.toi_events and ev <= 1, p
            assert w.sub_stepping() == (True, o.step_complete), p
            events += ev
            g, r = w.body_states(), o.bodies()
            for f in ("c", "a", "v", "w"):
                np.testing.assert_allclose(g[f], r[f], rtol=0, atol=1e-4, err_msg="pass %d %s" % (p, f))
            if sleep:
                np.testing.assert_array_equal(g["awake"], r["awake"], err_msg="pass %d awake" % p)
                np.testing.assert_array_equal(g["sleep_time"], r["sleep_time"], err_msg="pass %d sleep" % p)
            if o.step_complete:
                break
            resumed += 1
        assert events == full.profile()["toi_events"], p
        a, b = full.body_states(), w.body_states()
        for f in ("c", "a", "v", "w", "awake", "sleep_time"):
            assert a[f].tobytes() == b[f].tobytes(), (p, f)
    assert resumed >= 10
    full.close()


def test_sub_stepping_refusals():
    """While a sub-stepped pass is unfinished, bodies cannot be added and no tape can start; batched
    worlds refuse sub-stepping."""
    bd, fb, sd, fd, g = rain_scene()
    w = b2g.World(g)
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=True, block_solve=True)
    w.add(bd, fb, sd, fd)
    w.set_sub_stepping(True)
    for _ in range(60):
        w.step(DT, 8, 3)
        if not w.sub_stepping()[1]:
            break
    assert not w.sub_stepping()[1]
    with pytest.raises(RuntimeError):
        w.add(bd[1:2], np.zeros(1, np.int32), sd[1:2], fd[1:2])
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    with pytest.raises(RuntimeError):
        w.tape_begin(4)
    w.close()
    batch = b2g.World(g)
    batch.set_flags(warm_starting=True, allow_sleep=False, continuous=True, block_solve=True)
    batch.add_batched(np.concatenate([bd, bd]), np.repeat(np.arange(2, dtype=np.int32), len(bd)),
                      np.concatenate([fb, fb + len(bd)]), np.concatenate([sd, sd]), np.concatenate([fd, fd]))
    batch.set_sub_stepping(True)
    with pytest.raises(RuntimeError):
        batch.step(DT, 8, 3)
    batch.close()


def _run_tuned(scene, steps, sleep, **tune):
    bd, fb, sd, fd, g = scene
    w = b2g.World(g)
    w.set_flags(warm_starting=True, allow_sleep=sleep, continuous=True, block_solve=True)
    for k, v in tune.items():
        w.tune(getattr(b2g, k), v)
    w.add(bd, fb, sd, fd)
    events = []
    for _ in range(steps):
        w.step(DT, 8, 3)
        events.append(w.profile()["toi_events"])
    out = w.body_states().tobytes(), events
    w.close()
    return out


@pytest.mark.parametrize("scene", ["bounce", "bullets"])
def test_global_island_store_matches_lds(scene):
    """TOI islands too large for the event workgroup's LDS run in a global-memory store (the same
    event code); forcing every island there gives the same run bit for bit."""
    sc, sleep, steps = (bounce_rain_scene(), True, 150) if scene == "bounce" else (bullet_scene(), False, 12)
    lds = _run_tuned(sc, steps, sleep)
    glob = _run_tuned(sc, steps, sleep, TUNE_TOI_GLOBAL_ISLAND=1)
    assert sum(lds[1]) >= 6
    assert glob == lds


def plank_on_many_fixtures(n=600):
    """A 40 m plank falling at 60 m/s onto one static body made of n thin boxes side by side (its
    contact list holds every one of them: beyond the LDS store's 64 contacts and 512 contacts of
    one body)."""
    bd = np.zeros(2, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[1]["type"] = 2
    bd[1]["position"] = (0.0, 2.0)
    bd[1]["linear_velocity"] = (0.0, -60.0)
    sd = np.zeros(n + 1, b2g.SHAPE_DEF)
    xs = (np.arange(n) - 0.5 * (n - 1)) * 0.05
    sd[:n]["type"] = b2g.SHAPE_POLYGON
    sd[:n]["count"] = 4
    for i, x in enumerate(xs):
        sd[i]["params"][:8] = (x - 0.025, -0.25, x + 0.025, -0.25, x + 0.025, 0.25, x - 0.025, 0.25)
    sd[n]["type"] = b2g.SHAPE_BOX
    sd[n]["params"][:2] = (20.0, 0.1)
    fb = np.zeros(n + 1, np.int32)
    fb[n] = 1
    return bd, fb, sd, _fixtures(n + 1), (0.0, -10.0)


def test_toi_island_beyond_the_lds_store(oracle):
    """The plank's event gathers hundreds of contacts with the static body: the island runs in the
    global store and the plank stops on the boxes instead of passing into them, with the oracle's
    event counts at every step. The island's contacts are solved in the GPU's set order, the
    reference's in its contact-list order (which follows its dynamic tree's history); the TOI
    position solve stops once every contact is within 1.5 linear slops of its target, so the two
    orders may leave the plank anywhere in that band: positions are compared at 1.5 slops."""
    w, o = _pair(oracle, plank_on_many_fixtures(), True)
    events = 0
    for step in range(20):
        w.step(DT, 8, 3)
        o.step(DT, 8, 3)
        assert w.profile()["toi_events"] == o.toi_events, step
        events += o.toi_events
        g, r = w.body_states(), o.bodies()
        np.testing.assert_allclose(g["c"], r["c"], rtol=0, atol=1.5 * 0.005, err_msg="step %d" % step)
    assert events >= 1
    for st in (w.body_states()[1], o.bodies()[1]):
        assert 0.25 + 0.1 < float(st["c"][1]) < 0.25 + 0.1 + 0.05, st["c"]
        assert abs(float(st["v"][1])) < 0.05
