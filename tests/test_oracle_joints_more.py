"""Pins of the oracle's friction, motor, mouse, wheel, pulley and gear joints against closed forms.

The reference has no tests of its own for these joints. The restatements
(src/private/dynamics/joints/b2_friction_joint.rs, b2_motor_joint.rs, b2_mouse_joint.rs,
b2_wheel_joint.rs, b2_pulley_joint.rs, b2_gear_joint.rs) are pinned by what mechanics fixes:
  * friction: a box sliding and spinning without gravity, held by a friction joint at its centre,
    loses h F / m of speed and h T / I of spin per step (Coulomb friction at max force F and
    torque T) along its unchanged direction until it stops (1e-4);
  * motor: with ample force and torque, a box's position and angle errors to the motor's offsets
    shrink by (1 - correction factor) per step, exactly as the velocity target
    -correction / h * error makes them (1e-4);
  * mouse: a box pulled to a new target with stiffness m w^2 swings about it with the damped
    period 2 pi / (w sqrt(1 - zeta^2)) (2%, the soft step's discretisation), and critically
    damped it reaches the target without overshoot;
  * wheel: a wheel on a vertical spring axis settles at the static stretch m g / k, swings with
    period 2 pi sqrt(m / k) (2%) without leaving the axis, its motor turns it at the motor speed
    from the first step, and its translation limit holds within the linear slop;
  * pulley: an Atwood machine accelerates at g (r^2 m_A - r m_B) / (r^2 m_A + m_B) for ratios 1
    and 2 (0.1%), keeps length_a + ratio * length_b and its rate v_A + ratio v_B at zero;
  * gear: two revolute wheels geared with ratio r turn at w_B = -w_A / r, and a revolute wheel
    geared to a prismatic rack drives it at v_B = -w_A / r, with the coupling held.
A dropped term, a wrong sign or a swapped body fails one of these.
"""
import numpy as np
import pytest

import joint_scenes as J

DT = np.float32(1.0 / 60.0)
H = float(DT)


def _world(oracle, sc, gravity=(0.0, 0.0)):
    bd, fb, sd, fd, jd = sc[:5]
    w = oracle.world(gravity)
    w.set_flags(warm_starting=True, allow_sleep=False, block_solve=True, continuous=False)
    w.add(bd, fb, sd, fd)
    w.add_joints(jd)
    return w


def _crossings(x):
    s = np.sign(x)
    idx = np.flatnonzero(s[1:] * s[:-1] < 0)
    return np.array([i - x[i] / (x[i + 1] - x[i]) for i in idx]) * H


def test_friction_joint_coulomb(oracle):
    speeds = [(3.0, 0.0), (2.0, -1.5), (0.5, 0.5)]
    spins, forces, torques = [2.0, -1.0, 0.0], [1.0, 2.0, 0.2], [0.1, 0.3, 0.05]
    sc = J.friction_boxes(speeds, spins, forces, torques)
    w = _world(oracle, sc)
    mass = w.mass()
    v0 = np.asarray(speeds, np.float64)
    for n in range(1, 41):
        w.step(DT, 8, 3)
        st = w.bodies()
        for i in range(3):
            m, I = float(mass[1 + i, 0]), float(mass[1 + i, 2])
            sp0 = np.linalg.norm(v0[i])
            exp = max(sp0 - n * H * forces[i] / m, 0.0)
            v = st[1 + i]["v"].astype(np.float64)
            assert abs(np.linalg.norm(v) - exp) <= 1e-4 * sp0 + 1e-6, (n, i)
            if exp > 0:
                np.testing.assert_allclose(v / np.linalg.norm(v), v0[i] / sp0, atol=1e-5)
            w_exp = np.sign(spins[i]) * max(abs(spins[i]) - n * H * torques[i] / I, 0.0)
            assert abs(float(st[1 + i]["w"]) - w_exp) <= 1e-4 * max(abs(spins[i]), 1.0) + 1e-6, (n, i)


def test_motor_joint_error_decay(oracle):
    starts, angles = [(1.0, 2.0), (-3.0, 0.5)], [0.3, -0.2]
    offsets, aoffs = [(0.0, 0.0), (-1.0, 1.0)], [0.0, 0.5]
    w = _world(oracle, J.motor_boxes(starts, angles, offsets, aoffs))
    e0 = np.asarray(starts, np.float64) - np.asarray(offsets, np.float64)
    a0 = np.asarray(angles, np.float64) - np.asarray(aoffs, np.float64)
    for n in range(1, 11):
        w.step(DT, 8, 3)
        st = w.bodies()
        e = st[1:]["c"].astype(np.float64) - np.asarray(offsets, np.float64)
        a = st[1:]["a"].astype(np.float64) - np.asarray(aoffs, np.float64)
        np.testing.assert_allclose(e, 0.7 ** n * e0, atol=1e-4 * np.abs(e0).max())
        np.testing.assert_allclose(a, 0.7 ** n * a0, atol=1e-4 * np.abs(a0).max())


def test_mouse_joint_spring(oracle):
    sc = J.mouse_boxes([(0.0, 0.0), (5.0, 0.0)], [(2.0, 1.0), (5.0, -1.0)], [1.0, 1.0], [0.05, 1.0])
    targets = sc[5]
    w = _world(oracle, sc)
    w.set_joint_target(0, targets[0])
    w.set_joint_target(1, targets[1])
    xs = []
    for _ in range(300):
        w.step(DT, 8, 3)
        xs.append(w.bodies()["c"][1:].astype(np.float64))
    xs = np.array(xs)
    z = _crossings(xs[:, 0, 0] - targets[0][0])
    period = 2.0 * np.mean(np.diff(z[:6]))
    expected = 1.0 / np.sqrt(1.0 - 0.05 ** 2)  # 1 Hz spring, damping ratio 0.05
    assert abs(period / expected - 1.0) < 0.02, (period, expected)
    # critically damped: no overshoot past the target, and there at the end
    assert (xs[:, 1, 1] - targets[1][1]).min() > -1e-5
    assert np.abs(xs[-1, 1] - targets[1]).max() < 1e-4


def test_wheel_joint_spring_motor_limit(oracle):
    k = 40.0
    w = _world(oracle, J.wheels_on_springs(n=1, stiffness=(k,)), (0.0, -10.0))
    m = float(w.mass()[1, 0])
    ys = []
    for _ in range(400):
        w.step(DT, 8, 3)
        ys.append(w.bodies()[1]["c"].astype(np.float64))
    ys = np.array(ys)
    eq = 9.0 - m * 10.0 / k
    z = _crossings(ys[:, 1] - eq)
    period = 2.0 * np.mean(np.diff(z[:6]))
    assert abs(period / (2 * np.pi * np.sqrt(m / k)) - 1.0) < 0.02, period
    assert np.abs(ys[:, 0]).max() < 1e-5  # on the axis
    assert abs(ys[:, 1].mean() - eq) < 1e-3
    w = _world(oracle, J.wheels_on_springs(n=1, stiffness=(k,), motor_speed=5.0, torque=1000.0), (0.0, -10.0))
    for _ in range(3):
        w.step(DT, 8, 3)
        assert abs(float(w.bodies()[1]["w"]) - 5.0) < 1e-6
    w = _world(oracle, J.wheels_on_springs(n=1, stiffness=(k,), limit=(-0.15, 0.2)), (0.0, -10.0))
    tr = []
    for _ in range(200):
        w.step(DT, 8, 3)
        tr.append(float(w.bodies()[1]["c"][1]) - 9.0)
    assert min(tr) > -0.15 - 0.005 and max(tr) < 0.2 + 0.005
    assert min(tr) < -0.14  # the limit is reached


@pytest.mark.parametrize("ma,mb,r", [(2.0, 1.0, 1.0), (3.0, 1.0, 2.0)])
def test_pulley_atwood(oracle, ma, mb, r):
    w = _world(oracle, J.atwood(ma, mb, r), (0.0, -10.0))
    ys = []
    for _ in range(40):
        w.step(DT, 8, 3)
        st = w.bodies()
        ys.append((float(st[1]["c"][1]), float(st[2]["c"][1])))
        # the rope's rate: d/dt (length_a + ratio * length_b) = -(v_A + ratio v_B) = 0
        va, vb = float(st[1]["v"][1]), float(st[2]["v"][1])
        assert abs(va + r * vb) < 1e-4 * max(1.0, abs(va)), (va, vb)
    ys = np.array(ys)
    acc = np.diff(ys[:, 0], 2)[5:30].mean() / H ** 2
    expected = -10.0 * (r * r * ma - r * mb) / (r * r * ma + mb)
    assert abs(acc / expected - 1.0) < 1e-3, (acc, expected)
    la, lb = 9.5 - (ys[:, 0] + 0.5), 9.5 - (ys[:, 1] + 0.5)
    assert np.abs(la + r * lb - (4.0 + r * 4.0)).max() < 1e-4


@pytest.mark.parametrize("r,rack", [(2.0, False), (0.5, False), (1.0, True)])
def test_gear_ratio(oracle, r, rack):
    w = _world(oracle, J.geared_wheels(r, 3.0, prismatic_b=rack))
    st0 = w.bodies()
    for _ in range(120):
        w.step(DT, 8, 3)
        st = w.bodies()
        coord_b = float(st[2]["c"][0] - st0[2]["c"][0]) if rack else float(st[2]["a"] - st0[2]["a"])
        assert abs(float(st[1]["a"] - st0[1]["a"]) + r * coord_b) < 1e-4
    st = w.bodies()
    vb = float(st[2]["v"][0]) if rack else float(st[2]["w"])
    assert abs(float(st[1]["w"]) - 3.0) < 1e-5
    assert abs(vb + 3.0 / r) < 1e-5 * 3.0 / r
