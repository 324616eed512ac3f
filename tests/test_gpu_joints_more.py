"""Friction, motor, mouse, wheel, pulley and gear joints on the GPU against the oracle.

* Friction boxes, motor-driven boxes, mouse-dragged boxes (a target set after creation), wheels on
  spring axes (plain, motorised, limited) and Atwood pulleys: every joint is alone in its island,
  so the GPU's joint order is the reference's; every step agrees with the oracle within 1e-5
  (device sin/cos may differ from glibc's by one ulp), bodies and joint impulses.
* Geared wheels and a rack (three joints in one island, solved in the GPU's colour order against
  the reference's island order): the GPU holds the gear's ratio (w_B = -w_A / r, v_B = -w_A / r)
  and coupling as the oracle's closed forms say, and stays within 1e-3 of the oracle's states.
* The reference-style API creates each kind (B2wheelJointDef ... B2gearJointDef) and moves a
  mouse joint's target.
"""
import numpy as np
import pytest

import b2g
import joint_scenes as J

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def _pair(oracle, sc, gravity=(0.0, 0.0)):
    bd, fb, sd, fd, jd = sc[:5]
    g = b2g.World(gravity)
    g.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    g.add(bd, fb, sd, fd)
    g.add_joints(jd)
    r = oracle.world(gravity)
    r.set_flags(warm_starting=True, allow_sleep=False, block_solve=True, continuous=False)
    r.add(bd, fb, sd, fd)
    r.add_joints(jd)
    return g, r


def _compare(g, r, steps, tol=1e-5, impulses=True, what=""):
    for step in range(steps):
        g.step(DT, 8, 3)
        r.step(DT, 8, 3)
        gs, rs = g.body_states(), r.bodies()
        for f in ("c", "a", "v", "w"):
            a = np.asarray(gs[f], np.float64)
            b = np.asarray(rs[f], np.float64)
            err = np.abs(a - b).max()
            assert err <= tol * max(1.0, np.abs(b).max()), "%s step %d: %s differs by %g" % (what, step, f, err)
        if impulses:
            gi = g.joint_impulses().astype(np.float64)
            ri = r.joints()["impulse"].astype(np.float64)
            scale = max(1.0, np.abs(ri).max())
            assert np.abs(gi - ri).max() <= tol * scale, "%s step %d: impulses" % (what, step)


def test_friction_joints_match_oracle(oracle):
    sc = J.friction_boxes([(3.0, 0.0), (2.0, -1.5), (0.5, 0.5)], [2.0, -1.0, 0.0], [1.0, 2.0, 0.2], [0.1, 0.3, 0.05])
    _compare(*_pair(oracle, sc), 60, what="friction")


def test_motor_joints_match_oracle(oracle):
    sc = J.motor_boxes([(1.0, 2.0), (-3.0, 0.5), (4.0, -2.0)], [0.3, -0.2, 1.0], [(0.0, 0.0), (-1.0, 1.0), (5.0, 0.0)],
                       [0.0, 0.5, -0.5], max_force=[1e6, 1e6, 2.0], max_torque=[1e6, 1e6, 0.5])
    _compare(*_pair(oracle, sc), 60, what="motor")


def test_mouse_joints_match_oracle(oracle):
    sc = J.mouse_boxes([(0.0, 0.0), (5.0, 0.0), (10.0, 0.0)], [(2.0, 1.0), (5.0, -1.0), (12.0, 3.0)],
                       [1.0, 1.0, 2.0], [0.05, 1.0, 0.3], max_force=[1e6, 1e6, 3.0])
    g, r = _pair(oracle, sc)
    for k, t in enumerate(sc[5]):
        g.set_joint_target(k, t)
        r.set_joint_target(k, t)
    _compare(g, r, 60, what="mouse")
    for k, t in enumerate(sc[5]):  # a second drag
        g.set_joint_target(k, (t[0] - 1.0, t[1] + 0.5))
        r.set_joint_target(k, (t[0] - 1.0, t[1] + 0.5))
    _compare(g, r, 60, what="mouse, second target")


@pytest.mark.parametrize("variant", ["spring", "motor", "limit"])
def test_wheel_joints_match_oracle(oracle, variant):
    kw = {"motor": dict(motor_speed=5.0, torque=0.4), "limit": dict(limit=(-0.15, 0.2))}.get(variant, {})
    sc = J.wheels_on_springs(n=3, stiffness=(40.0, 80.0, 160.0), **kw)
    sc[4]["damping"] = (0.0, 0.5, 2.0)
    _compare(*_pair(oracle, sc, (0.0, -10.0)), 120, what="wheel " + variant)


@pytest.mark.parametrize("ma,mb,r", [(2.0, 1.0, 1.0), (3.0, 1.0, 2.0)])
def test_pulleys_match_oracle(oracle, ma, mb, r):
    _compare(*_pair(oracle, J.atwood(ma, mb, r), (0.0, -10.0)), 40, what="pulley")


@pytest.mark.parametrize("r,rack", [(2.0, False), (0.5, False), (1.0, True)])
def test_gears_hold_the_ratio(oracle, r, rack):
    g, o = _pair(oracle, J.geared_wheels(r, 3.0, prismatic_b=rack))
    st0 = g.body_states()
    for _ in range(120):
        g.step(DT, 8, 3)
        o.step(DT, 8, 3)
        st = g.body_states()
        coord_b = float(st[2]["c"][0] - st0[2]["c"][0]) if rack else float(st[2]["a"] - st0[2]["a"])
        assert abs(float(st[1]["a"] - st0[1]["a"]) + r * coord_b) < 1e-4
        rs = o.bodies()
        for f in ("c", "a", "v", "w"):
            assert np.abs(st[f].astype(np.float64) - rs[f]).max() <= 1e-3 * max(1.0, np.abs(rs[f]).max()), f
    vb = float(st[2]["v"][0]) if rack else float(st[2]["w"])
    assert abs(float(st[1]["w"]) - 3.0) < 1e-4
    assert abs(vb + 3.0 / r) < 1e-4 * 3.0 / r


def test_reference_style_new_joints():
    world = b2g.B2world((0.0, -10.0))
    ground = world.create_body(b2g.B2bodyDef(position=(0.0, 0.0)))
    circle = b2g.B2circleShape()
    circle.m_radius = 0.4
    box = b2g.B2polygonShape()
    box.set_as_box(0.5, 0.25)

    def dyn(pos, shape=box):
        b = world.create_body(b2g.B2bodyDef(body_type=b2g.B2_DYNAMIC_BODY, position=pos))
        b.create_fixture_by_shape(shape, 1.0)
        return b

    car, wheel = dyn((0.0, 5.0)), dyn((0.0, 4.0), circle)
    wd = b2g.B2wheelJointDef()
    wd.initialize(car, wheel, (0.0, 4.0), (0.0, 1.0))
    wd.stiffness, wd.enable_motor, wd.motor_speed, wd.max_motor_torque = 50.0, True, -4.0, 100.0
    world.create_joint(wd)
    slider = dyn((5.0, 5.0))
    fd = b2g.B2frictionJointDef()
    fd.initialize(ground, slider, (5.0, 5.0))
    fd.max_force, fd.max_torque = 1000.0, 1000.0
    world.create_joint(fd)
    follower = dyn((8.0, 5.0))
    md = b2g.B2motorJointDef()
    md.initialize(ground, follower)
    md.max_force, md.max_torque = 1000.0, 1000.0
    world.create_joint(md)
    dragged = dyn((12.0, 5.0))
    ms = b2g.B2mouseJointDef()
    ms.body_a, ms.body_b, ms.target = ground, dragged, (12.0, 5.0)
    ms.max_force, ms.stiffness, ms.damping = 1000.0, 50.0, 5.0
    mouse = world.create_joint(ms)
    pa, pb = dyn((16.0, 4.0)), dyn((19.0, 4.0))
    pd = b2g.B2pulleyJointDef()
    pd.initialize(pa, pb, (16.0, 8.0), (19.0, 8.0), (16.0, 4.25), (19.0, 4.25), 1.0)
    world.create_joint(pd)
    g1, g2 = dyn((23.0, 5.0), circle), dyn((25.0, 5.0), circle)
    r1, r2 = b2g.B2revoluteJointDef(), b2g.B2revoluteJointDef()
    r1.initialize(ground, g1, (23.0, 5.0))
    r2.initialize(ground, g2, (25.0, 5.0))
    j1, j2 = world.create_joint(r1), world.create_joint(r2)
    gd = b2g.B2gearJointDef()
    gd.body_a, gd.body_b, gd.joint1, gd.joint2, gd.ratio = g1, g2, j1, j2, 2.0
    world.create_joint(gd)
    world.step(1.0 / 60.0, 8, 3)
    mouse.set_target((13.0, 6.0))
    for _ in range(60):
        world.step(1.0 / 60.0, 8, 3)
    # the wheel's motor (the free-falling car turns the other way by the reaction)
    assert abs(wheel.get_angular_velocity() - car.get_angular_velocity() + 4.0) < 1e-3
    assert abs(slider.get_position().y - 5.0) < 0.05  # friction holds it against gravity
    assert abs(follower.get_position().y - 5.0) < 0.05  # the motor joint holds its offset
    p = dragged.get_position()
    assert abs(p.x - 13.0) < 0.3 and abs(p.y - 6.0) < 0.5  # pulled to the new target (sagging)
    assert abs((4.0 - pa.get_position().y) + (4.0 - pb.get_position().y)) < 1e-3  # one rope
    w1, w2 = g1.get_angular_velocity(), g2.get_angular_velocity()
    assert abs(w1 + 2.0 * w2) < 1e-4  # geared
