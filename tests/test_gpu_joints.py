"""Distance and revolute joints on the GPU against the oracle.

* Independent pendulums (each joint alone in its island, so the GPU's joint order is the
  reference's): every step of 150 agrees with the oracle's reference-order step within 1e-5
  (device sin/cos may differ from glibc's by one ulp), bodies and joint impulses.
* A hanging 12-link chain (joints in two colours against the reference's island order) stays
  within a band of the oracle's run over one second.
* Welded cantilevers (rigid and soft) and prismatic sliders (inclined, limited, motorised), each
  alone in its island, agree with the oracle at every step, as the pendulums do.
* Motor and limit behave as the closed forms of the oracle's pins say; collide_connected filters
  pairs, also for a joint created between touching bodies; a pendulum at rest falls asleep with
  sleeping on; runs are deterministic.
"""
import numpy as np
import pytest

import b2g
from joint_scenes import chain, pendulums, sliders, welds

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def _gpu(bd, fb, sd, fd, jd, gravity=(0.0, -10.0), sleep=False):
    w = b2g.World(gravity)
    w.set_flags(warm_starting=True, allow_sleep=sleep, continuous=False, block_solve=True)
    w.add(bd, fb, sd, fd)
    w.add_joints(jd)
    return w


def _cpu(oracle, bd, fb, sd, fd, jd, gravity=(0.0, -10.0), sleep=False):
    w = oracle.world(gravity)
    w.set_flags(warm_starting=True, allow_sleep=sleep, block_solve=True, continuous=False)
    w.add(bd, fb, sd, fd)
    w.add_joints(jd)
    return w


def _close(g, r, tol, what):
    for f in ("c", "a", "v", "w"):
        a = np.asarray(g[f], np.float64)
        b = np.asarray(r[f], np.float64)
        err = np.abs(a - b).max()
        assert err <= tol * max(1.0, np.abs(b).max()), "%s: %s differs by %g" % (what, f, err)


def test_pendulums_match_oracle(oracle):
    kinds = ["revolute", "distance"] * 4
    bd, fb, sd, fd, jd = pendulums(kinds, length=2.0, theta0=0.6)
    # different amplitudes and a soft, limited distance joint and a motorised, limited revolute
    for i in range(len(kinds)):
        bd[1 + i]["linear_velocity"] = (0.5 * i, 0.0)
    jd[3]["min_length"], jd[3]["max_length"], jd[3]["stiffness"], jd[3]["damping"] = 1.5, 2.5, 30.0, 0.5
    jd[2]["enable_limit"], jd[2]["lower_angle"], jd[2]["upper_angle"] = 1, -0.3, 0.4
    jd[4]["enable_motor"], jd[4]["motor_speed"], jd[4]["max_motor_torque"] = 1, 2.0, 0.02
    g = _gpu(bd, fb, sd, fd, jd)
    r = _cpu(oracle, bd, fb, sd, fd, jd)
    for step in range(150):
        g.step(DT, 8, 3)
        r.step(DT, 8, 3)
        _close(g.body_states(), r.bodies(), 1e-5, "step %d" % step)
        gi = g.joint_impulses()
        ri = np.asarray(r.joints()["impulse"], np.float64)
        err = np.abs(gi - ri).max()
        assert err <= 1e-5 * max(1.0, np.abs(ri).max()), "step %d: joint impulses differ by %g" % (step, err)


def test_chain_band(oracle):
    bd, fb, sd, fd, jd = chain(links=12)
    g = _gpu(bd, fb, sd, fd, jd)
    r = _cpu(oracle, bd, fb, sd, fd, jd)
    for _ in range(60):
        g.step(DT, 8, 3)
        r.step(DT, 8, 3)
    gs, rs = g.body_states(), r.bodies()
    assert np.abs(gs["c"] - rs["c"]).max() < 0.02
    # the chain holds: consecutive hinge points coincide within a few slops
    st = gs
    for i in range(1, 12):
        def world_point(b, local):
            q = st["q"][b]
            return st["p"][b] + np.array([q[1] * local[0] - q[0] * local[1], q[0] * local[0] + q[1] * local[1]])
        pa = world_point(jd[i]["body_a"], jd[i]["local_anchor_a"])
        pb = world_point(jd[i]["body_b"], jd[i]["local_anchor_b"])
        assert np.linalg.norm(pa - pb) < 0.02


def test_motor_and_limit():
    bd, fb, sd, fd, jd = pendulums(["revolute"], length=0.0, theta0=0.0, radius=0.5)
    jd["enable_motor"], jd["motor_speed"], jd["max_motor_torque"] = 1, 3.0, 1000.0
    w = _gpu(bd, fb, sd, fd, jd, gravity=(0.0, 0.0))
    for _ in range(5):
        w.step(DT, 8, 3)
        assert w.body_states()[1]["w"] == np.float32(3.0)
    bd, fb, sd, fd, jd = chain(links=1)
    jd["enable_limit"], jd["lower_angle"], jd["upper_angle"] = 1, -0.4, 0.3
    w = _gpu(bd, fb, sd, fd, jd)
    slop = 2.0 / 180.0 * np.pi
    lo = 1.0
    for _ in range(240):
        w.step(DT, 8, 3)
        a = float(w.body_states()[1]["a"])
        assert -0.4 - 2 * slop <= a <= 0.3 + 2 * slop, a
        lo = min(lo, a)
    assert lo < -0.4 + slop


@pytest.mark.parametrize("connected", [False, True])
def test_collide_connected(connected):
    bd, fb, sd, fd, jd = chain(links=2, collide_connected=connected)
    sd["params"][:, 0] = 0.3
    w = _gpu(bd, fb, sd, fd, jd)
    w.step(DT, 8, 3)
    assert (w.contact_count > 0) == connected


def test_joint_added_later_filters_the_contact():
    bd, fb, sd, fd, jd = chain(links=2, collide_connected=True)
    sd["params"][:, 0] = 0.3
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    w.add(bd, fb, sd, fd)
    w.add_joints(jd[:1])
    w.step(DT, 8, 3)
    fab, touching, _ = w.contacts()
    assert touching.sum() == 1
    j2 = jd[1:].copy()
    j2["collide_connected"] = 0
    w.add_joints(j2)
    w.step(DT, 8, 3)
    assert w.contact_count == 0


def test_pendulum_at_rest_sleeps(oracle):
    bd, fb, sd, fd, jd = pendulums(["revolute", "distance"], length=1.0, theta0=0.0)
    g = _gpu(bd, fb, sd, fd, jd, sleep=True)
    r = _cpu(oracle, bd, fb, sd, fd, jd, sleep=True)
    n_g = n_r = None
    for step in range(120):
        g.step(DT, 8, 3)
        r.step(DT, 8, 3)
        if n_g is None and not np.any(g.body_states()["awake"][1:]):
            n_g = step
        if n_r is None and not np.any(r.bodies()["awake"][1:]):
            n_r = step
    assert n_g is not None and n_g == n_r


def test_chain_on_ground_is_deterministic():
    bd, fb, sd, fd, jd = chain(links=16, y=3.0)
    # a ground box under the chain: contacts and joints in one island
    bd2 = np.concatenate([bd, np.zeros(1, b2g.BODY_DEF)])
    bd2[-1]["gravity_scale"] = 1.0
    sd2 = np.concatenate([sd, np.zeros(1, b2g.SHAPE_DEF)])
    sd2[-1]["type"] = b2g.SHAPE_BOX
    sd2[-1]["params"][:5] = (20.0, 0.5, 0.0, -0.5, 0.0)
    sd2[-1]["flags"] = 1
    fd2 = np.concatenate([fd, fd[:1]])
    fb2 = np.concatenate([fb, [len(bd)]]).astype(np.int32)
    runs = []
    for _ in range(2):
        w = _gpu(bd2, fb2, sd2, fd2, jd)
        for _ in range(240):
            w.step(DT, 8, 3)
        runs.append(w.body_states())
    for f in ("c", "a", "v", "w"):
        np.testing.assert_array_equal(runs[0][f], runs[1][f])
    assert np.all(np.isfinite(runs[0]["c"]))
    assert runs[0]["c"][1:17, 1].min() > -0.2  # the chain rests on the ground


def test_reference_style_api():
    world = b2g.B2world((0.0, -10.0))
    ground = world.create_body(b2g.B2bodyDef(position=(0.0, 0.0)))
    bob = world.create_body(b2g.B2bodyDef(body_type=b2g.B2_DYNAMIC_BODY, position=(2.0, 10.0)))
    circle = b2g.B2circleShape(0.1)
    bob.create_fixture_by_shape(circle, 1.0)
    jd = b2g.B2revoluteJointDef()
    jd.initialize(ground, bob, (0.0, 10.0))
    world.create_joint(jd)
    rod = world.create_body(b2g.B2bodyDef(body_type=b2g.B2_DYNAMIC_BODY, position=(5.0, 10.0)))
    rod.create_fixture_by_shape(circle, 1.0)
    dd = b2g.B2distanceJointDef()
    dd.initialize(ground, rod, (3.0, 10.0), (5.0, 10.0))
    world.create_joint(dd)
    for _ in range(60):
        world.step(1.0 / 60.0, 8, 3)
    p = np.array(bob.get_position())
    assert abs(np.linalg.norm(p - (0.0, 10.0)) - 2.0) < 0.01
    assert p[1] < 9.0  # it swung down
    q = np.array(rod.get_position())
    assert abs(np.linalg.norm(q - (3.0, 10.0)) - 2.0) < 0.01


def test_reference_style_prismatic_and_weld():
    world = b2g.B2world((0.0, -10.0))
    ground = world.create_body(b2g.B2bodyDef())
    box = b2g.B2polygonShape()
    box.set_as_box(0.5, 0.25)
    slider = world.create_body(b2g.B2bodyDef(body_type=b2g.B2_DYNAMIC_BODY, position=(0.0, 10.0)))
    slider.create_fixture_by_shape(box, 1.0)
    pj = b2g.B2prismaticJointDef()
    pj.initialize(ground, slider, (0.0, 10.0), (1.0, 0.0))
    pj.enable_motor, pj.motor_speed, pj.max_motor_force = True, 2.0, 1e4
    world.create_joint(pj)
    beam = world.create_body(b2g.B2bodyDef(body_type=b2g.B2_DYNAMIC_BODY, position=(6.0, 10.0)))
    beam.create_fixture_by_shape(box, 1.0)
    wj = b2g.B2weldJointDef()
    wj.initialize(ground, beam, (5.5, 10.0))
    world.create_joint(wj)
    for _ in range(60):
        world.step(1.0 / 60.0, 8, 3)
    assert abs(slider.get_linear_velocity()[0] - 2.0) < 1e-4
    assert abs(slider.get_position()[1] - 10.0) < 0.005
    assert abs(beam.get_angle()) < 0.05 and abs(beam.get_position()[1] - 10.0) < 0.01


def test_welds_match_oracle(oracle):
    bd, fb, sd, fd, jd = welds([0.0, 5.0, 0.0, 40.0], omega=0.5)
    jd[1]["damping"] = 0.3
    bd[3]["angle"] = 0.2  # welded at an angle: the reference angle is 0, the weld pulls it back
    g = _gpu(bd, fb, sd, fd, jd)
    r = _cpu(oracle, bd, fb, sd, fd, jd)
    for step in range(150):
        g.step(DT, 8, 3)
        r.step(DT, 8, 3)
        _close(g.body_states(), r.bodies(), 1e-5, "step %d" % step)
        gi = g.joint_impulses()[:, :3]
        ri = np.asarray(r.joints()["impulse"], np.float64)[:, :3]
        err = np.abs(gi - ri).max()
        assert err <= 1e-5 * max(1.0, np.abs(ri).max()), "step %d: weld impulses differ by %g" % (step, err)


def test_sliders_match_oracle(oracle):
    bd, fb, sd, fd, jd = sliders([-np.pi / 6, np.pi / 2 + 0.3, -np.pi / 4, 0.2])
    bd[2]["angular_velocity"] = 1.0  # starts turning: the joint takes it back
    jd[2]["enable_limit"], jd[2]["lower_angle"], jd[2]["upper_angle"] = 1, -1.0, 0.5
    jd[3]["enable_motor"], jd[3]["motor_speed"], jd[3]["max_motor_torque"] = 1, -1.5, 30.0
    jd[3]["reference_angle"] = 0.1
    g = _gpu(bd, fb, sd, fd, jd)
    r = _cpu(oracle, bd, fb, sd, fd, jd)
    for step in range(150):
        g.step(DT, 8, 3)
        r.step(DT, 8, 3)
        _close(g.body_states(), r.bodies(), 1e-5, "step %d" % step)
        gi = g.joint_impulses()
        ri = np.asarray(r.joints()["impulse"], np.float64)
        err = np.abs(gi - ri).max()
        assert err <= 1e-5 * max(1.0, np.abs(ri).max()), "step %d: slider impulses differ by %g" % (step, err)

