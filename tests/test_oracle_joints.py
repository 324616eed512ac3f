"""Pins of the oracle's joints (distance and revolute) against closed forms and invariants.

The reference has no joint tests of its own; these pin the restatement
(src/private/dynamics/joints/b2_distance_joint.rs, b2_revolute_joint.rs, the island's joint
loops in src/private/dynamics/b2_island_private.rs:196-275 and the joint filter of
src/private/dynamics/b2_body.rs:391-416) by what mechanics fixes:
  * pendulum period: a bob on a revolute or a rigid distance joint swings with the physical
    pendulum's small-angle period 2 pi sqrt((I + m L^2) / (m g L)) (within 1.5%, the
    amplitude correction at 0.15 rad and the integrator's error), and the joint holds: the
    anchors stay together (revolute) or the length stays L (distance) to within the linear slop;
  * soft distance joint: with gravity off, a bob on a spring (stiffness k, no damping) oscillates
    with period 2 pi sqrt(m / k) (within 2%);
  * motor: a wheel pinned at its centre with a motor of ample torque turns at the motor speed
    after the first step, exactly; with a small torque it accelerates at max_torque / I;
  * limit: a hinged bar released horizontally stays within its angle limits (plus the angular
    slop);
  * collide_connected: overlapping jointed boxes get no contact unless the joint allows it, and
    a joint created between touching bodies removes their contact at the next step;
  * weld: a rigid cantilever holds its angle and anchor within the slops under gravity; a soft
    weld (angular stiffness k) swings about its anchor with period 2 pi sqrt(I_anchor / k)
    (within 2%), I_anchor = I + m d^2 about the anchor;
  * prismatic: a box on an axis inclined by theta slides down it as the integrator's closed form
    of frictionless acceleration g sin(theta) says (within 0.5%), without turning or leaving the
    axis; its translation limits hold within the slop; a motor of ample force drives it along
    the axis at the motor speed.
A slip such as a wrong sign in the block solve, a dropped inertia term or a swapped limit
fails one of these.
"""
import numpy as np
import pytest

import b2g
from joint_scenes import chain, pendulums, sliders, welds

DT = np.float32(1.0 / 60.0)
G = 10.0


def _oracle_world(oracle, bd, fb, sd, fd, jd, gravity=(0.0, -10.0)):
    w = oracle.world(gravity)
    w.set_flags(warm_starting=True, allow_sleep=False, block_solve=True)
    w.add(bd, fb, sd, fd)
    w.add_joints(jd)
    return w


def _zero_crossings(t, x):
    s = np.sign(x)
    idx = np.flatnonzero(s[1:] * s[:-1] < 0)
    # linear interpolation of each crossing
    return np.array([t[i] - x[i] * (t[i + 1] - t[i]) / (x[i + 1] - x[i]) for i in idx])


@pytest.mark.parametrize("kind", ["revolute", "distance"])
def test_pendulum_period_and_constraint(oracle, kind):
    L, theta0, r = 2.0, 0.15, 0.1
    bd, fb, sd, fd, jd = pendulums([kind], length=L, theta0=theta0, radius=r)
    w = _oracle_world(oracle, bd, fb, sd, fd, jd)
    m = fd["density"][0] * np.pi * r * r
    I = 0.5 * m * r * r
    T = 2 * np.pi * np.sqrt((I + m * L * L) / (m * G * L))
    pivot = np.array([0.0, 10.0])
    steps = int(3.2 * T / DT)
    xs, err = [], 0.0
    for _ in range(steps):
        w.step(DT, 8, 3)
        st = w.bodies()[1]
        c = np.array(st["c"], np.float64)
        xs.append(c[0] - pivot[0])
        err = max(err, abs(np.linalg.norm(c - pivot) - L))
    t = np.arange(1, steps + 1) * float(DT)
    zc = _zero_crossings(t, np.array(xs))
    assert len(zc) >= 5
    period = 2 * np.mean(np.diff(zc))
    assert abs(period / T - 1) < 0.015, (period, T)
    assert err < 0.005, err


def test_soft_distance_joint_spring_period(oracle):
    L, k, r = 2.0, 0.5, 0.1
    bd, fb, sd, fd, jd = pendulums(["distance"], length=L, theta0=0.0, radius=r)
    jd["min_length"] = 0.5
    jd["max_length"] = 4.0
    jd["stiffness"] = k
    jd["length"] = 1.8  # rest length: the bob starts stretched by 0.2
    w = _oracle_world(oracle, bd, fb, sd, fd, jd, gravity=(0.0, 0.0))
    m = fd["density"][0] * np.pi * r * r
    T = 2 * np.pi * np.sqrt(m / k)
    steps = int(4.2 * T / DT)
    ys = []
    for _ in range(steps):
        w.step(DT, 8, 3)
        ys.append(10.0 - w.bodies()[1]["c"][1] - 1.8)
    t = np.arange(1, steps + 1) * float(DT)
    zc = _zero_crossings(t, np.array(ys))
    assert len(zc) >= 6
    period = 2 * np.mean(np.diff(zc))
    assert abs(period / T - 1) < 0.02, (period, T)


def _wheel(oracle, torque, speed):
    bd, fb, sd, fd, jd = pendulums(["revolute"], length=0.0, theta0=0.0, radius=0.5)
    jd["enable_motor"] = 1
    jd["motor_speed"] = speed
    jd["max_motor_torque"] = torque
    w = _oracle_world(oracle, bd, fb, sd, fd, jd, gravity=(0.0, 0.0))
    m = fd["density"][0] * np.pi * 0.25
    return w, 0.5 * m * 0.25


def test_revolute_motor(oracle):
    w, I = _wheel(oracle, torque=1000.0, speed=3.0)
    for _ in range(5):
        w.step(DT, 8, 3)
        assert w.bodies()[1]["w"] == np.float32(3.0)
    w, I = _wheel(oracle, torque=0.05, speed=30.0)
    n = 30
    for _ in range(n):
        w.step(DT, 8, 3)
    omega = float(w.bodies()[1]["w"])
    assert abs(omega - 0.05 / I * n * float(DT)) < 1e-3 * omega, (omega, 0.05 / I * n * float(DT))


def test_revolute_limit(oracle):
    bd, fb, sd, fd, jd = chain(links=1)
    lower, upper = -0.4, 0.3
    jd["enable_limit"] = 1
    jd["lower_angle"] = lower
    jd["upper_angle"] = upper
    w = _oracle_world(oracle, bd, fb, sd, fd, jd)
    slop = 2.0 / 180.0 * np.pi
    lo = 1.0
    for _ in range(240):
        w.step(DT, 8, 3)
        a = float(w.bodies()[1]["a"])
        assert lower - 2 * slop <= a <= upper + 2 * slop, a
        lo = min(lo, a)
    assert lo < lower + slop  # it reached the lower limit


@pytest.mark.parametrize("connected", [False, True])
def test_collide_connected(oracle, connected):
    bd, fb, sd, fd, jd = chain(links=2, collide_connected=connected)
    sd["params"][:, 0] = 0.3  # links overlap near the hinge
    w = _oracle_world(oracle, bd, fb, sd, fd, jd)
    w.step(DT, 8, 3)
    assert (w.contact_count > 0) == connected


def test_joint_filters_an_existing_contact(oracle):
    bd, fb, sd, fd, jd = chain(links=2, collide_connected=True)
    sd["params"][:, 0] = 0.3
    w = oracle.world((0.0, -10.0))
    w.set_flags(warm_starting=True, allow_sleep=False, block_solve=True)
    w.add(bd, fb, sd, fd)
    w.add_joints(jd[:1])
    w.step(DT, 8, 3)
    assert w.contact_count == 1  # links 1 and 2 overlap, nothing joins them yet
    j2 = jd[1:].copy()
    j2["collide_connected"] = 0
    w.add_joints(j2)
    w.step(DT, 8, 3)
    assert w.contact_count == 0


def test_weld_cantilever_rigid(oracle):
    bd, fb, sd, fd, jd = welds([0.0])
    w = _oracle_world(oracle, bd, fb, sd, fd, jd)
    slop = 2.0 / 180.0 * np.pi
    for _ in range(120):
        w.step(DT, 8, 3)
        st = w.bodies()[1]
        assert abs(float(st["a"])) < 2 * slop
        # the welded end stays at the anchor
        q = st["q"]
        end = np.array(st["p"], np.float64) + np.array([q[1] * -1.0, q[0] * -1.0])
        assert np.linalg.norm(end - (0.0, 10.0)) < 0.01


def test_weld_soft_angular_period(oracle):
    k = 2.0
    half = (1.0, 0.1)
    bd, fb, sd, fd, jd = welds([k], half=half, omega=0.2)
    w = _oracle_world(oracle, bd, fb, sd, fd, jd, gravity=(0.0, 0.0))
    m = fd["density"][0] * 4 * half[0] * half[1]
    I = m * (4 * half[0] ** 2 + 4 * half[1] ** 2) / 12.0
    T = 2 * np.pi * np.sqrt((I + m * half[0] ** 2) / k)
    steps = int(3.2 * T / DT)
    angles = []
    for _ in range(steps):
        w.step(DT, 8, 3)
        angles.append(float(w.bodies()[1]["a"]))
    t = np.arange(1, steps + 1) * float(DT)
    zc = _zero_crossings(t, np.array(angles))
    assert len(zc) >= 5
    period = 2 * np.mean(np.diff(zc))
    assert abs(period / T - 1) < 0.02, (period, T)


def _slide(st, angle, anchor):
    d = np.array(st["c"], np.float64) - anchor
    return d @ np.array([np.cos(angle), np.sin(angle)]), d @ np.array([-np.sin(angle), np.cos(angle)])


def test_prismatic_incline(oracle):
    angles = [-np.pi / 6, np.pi / 2 + 0.3]
    bd, fb, sd, fd, jd = sliders(angles)
    w = _oracle_world(oracle, bd, fb, sd, fd, jd)
    h = float(DT)
    n = 60
    for _ in range(n):
        w.step(DT, 8, 3)
    for i, ang in enumerate(angles):
        st = w.bodies()[1 + i]
        s_along, s_perp = _slide(st, ang, np.array([4.0 * i, 10.0]))
        # symplectic Euler from rest under the axial component of gravity: s_n = h^2 a n (n + 1) / 2
        expect = h * h * (-G * np.sin(ang)) * n * (n + 1) / 2
        assert abs(s_along / expect - 1) < 0.005, (s_along, expect)
        assert abs(s_perp) < 0.005 and abs(float(st["a"])) < 2.0 / 180.0 * np.pi


def test_prismatic_limit_and_motor(oracle):
    bd, fb, sd, fd, jd = sliders([-np.pi / 4, 0.0])
    jd[0]["enable_limit"], jd[0]["lower_angle"], jd[0]["upper_angle"] = 1, -1.0, 0.5
    jd[1]["enable_motor"], jd[1]["motor_speed"], jd[1]["max_motor_torque"] = 1, 1.5, 1e4
    w = _oracle_world(oracle, bd, fb, sd, fd, jd)
    for _ in range(120):
        w.step(DT, 8, 3)
        st = w.bodies()
        s0, _ = _slide(st[1], -np.pi / 4, np.array([0.0, 10.0]))
        assert -1.0 - 0.01 <= s0 <= 0.5 + 0.01, s0
        assert abs(float(st[2]["v"][0]) - 1.5) < 1e-5
    assert s0 > 0.5 - 0.02  # it slid down to the upper limit (the axis points down the slope)

