"""Small joint scenes shared by the oracle and GPU joint tests (definitions in the library's
layouts: BODY_DEF, SHAPE_DEF, FIXTURE_DEF, JOINT_DEF)."""
import numpy as np

import b2g


def _fixture(n):
    fd = np.zeros(n, b2g.FIXTURE_DEF)
    fd["friction"] = 0.2
    fd["restitution_threshold"] = 1.0
    fd["category"] = 1
    fd["mask"] = 0xFFFF
    fd["density"] = 1.0
    return fd


def _bodies(n):
    bd = np.zeros(n, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    return bd


def pendulums(kinds, length=2.0, theta0=0.3, radius=0.1, spacing=3.0):
    """One static pivot body and, per entry of kinds ("revolute" or "distance"), a circle bob hung
    from its own pivot point at angle theta0 from the vertical (each pendulum is its own island)."""
    n = len(kinds)
    bd = _bodies(1 + n)
    bd[0]["type"] = 0
    sd = np.zeros(n, b2g.SHAPE_DEF)
    jd = np.zeros(n, b2g.JOINT_DEF)
    fb = np.arange(1, n + 1, dtype=np.int32)
    for i, kind in enumerate(kinds):
        pivot = np.array([spacing * i, 10.0])
        bob = pivot + length * np.array([np.sin(theta0), -np.cos(theta0)])
        bd[1 + i]["type"] = 2
        bd[1 + i]["position"] = bob
        sd[i]["type"] = b2g.SHAPE_CIRCLE
        sd[i]["radius"] = radius
        j = jd[i]
        j["body_a"], j["body_b"] = 0, 1 + i
        j["local_anchor_a"] = pivot
        if kind == "revolute":
            j["type"] = b2g.JOINT_REVOLUTE
            j["local_anchor_b"] = pivot - bob
        else:
            j["type"] = b2g.JOINT_DISTANCE
            j["local_anchor_b"] = (0.0, 0.0)
            j["length"] = j["min_length"] = j["max_length"] = length
    return bd, fb, sd, _fixture(n), jd


def chain(links=12, half=0.25, y=10.0, collide_connected=False):
    """The testbed's chain: boxes of half-length `half` hinged end to end by revolute joints,
    the first hinged to a static ground body, hanging sideways."""
    bd = _bodies(1 + links)
    bd[0]["type"] = 0
    sd = np.zeros(links, b2g.SHAPE_DEF)
    jd = np.zeros(links, b2g.JOINT_DEF)
    fb = np.arange(1, links + 1, dtype=np.int32)
    for i in range(links):
        bd[1 + i]["type"] = 2
        bd[1 + i]["position"] = (half + 2 * half * i, y)
        sd[i]["type"] = b2g.SHAPE_BOX
        sd[i]["params"][:2] = (half, 0.0625)
        j = jd[i]
        j["type"] = b2g.JOINT_REVOLUTE
        j["collide_connected"] = int(collide_connected)
        anchor = np.array([2 * half * i, y])
        j["body_a"], j["body_b"] = i, 1 + i
        a_pos = np.zeros(2) if i == 0 else np.array(bd[i]["position"], np.float64)
        j["local_anchor_a"] = anchor - a_pos
        j["local_anchor_b"] = anchor - np.array(bd[1 + i]["position"], np.float64)
    return bd, fb, sd, _fixture(links), jd


def welds(stiffness, half=(1.0, 0.1), spacing=4.0, omega=0.0):
    """One static body and, per entry of stiffness, a box welded at its left end to a point of the
    static body (a cantilever; stiffness 0 is the rigid weld), each its own island."""
    n = len(stiffness)
    bd = _bodies(1 + n)
    bd[0]["type"] = 0
    sd = np.zeros(n, b2g.SHAPE_DEF)
    jd = np.zeros(n, b2g.JOINT_DEF)
    fb = np.arange(1, n + 1, dtype=np.int32)
    for i, k in enumerate(stiffness):
        anchor = np.array([spacing * i, 10.0])
        bd[1 + i]["type"] = 2
        bd[1 + i]["position"] = anchor + np.array([half[0], 0.0])
        bd[1 + i]["angular_velocity"] = omega
        sd[i]["type"] = b2g.SHAPE_BOX
        sd[i]["params"][:2] = half
        j = jd[i]
        j["type"] = b2g.JOINT_WELD
        j["body_a"], j["body_b"] = 0, 1 + i
        j["local_anchor_a"] = anchor
        j["local_anchor_b"] = (-half[0], 0.0)
        j["stiffness"] = k
    return bd, fb, sd, _fixture(n), jd


def sliders(angles, half=(0.5, 0.25), spacing=4.0):
    """One static body and, per entry of angles, a box on a prismatic joint whose axis (in the
    static body's frame) makes that angle with the x axis, anchored at the box's centre (each its
    own island)."""
    n = len(angles)
    bd = _bodies(1 + n)
    bd[0]["type"] = 0
    sd = np.zeros(n, b2g.SHAPE_DEF)
    jd = np.zeros(n, b2g.JOINT_DEF)
    fb = np.arange(1, n + 1, dtype=np.int32)
    for i, ang in enumerate(angles):
        anchor = np.array([spacing * i, 10.0])
        bd[1 + i]["type"] = 2
        bd[1 + i]["position"] = anchor
        sd[i]["type"] = b2g.SHAPE_BOX
        sd[i]["params"][:2] = half
        j = jd[i]
        j["type"] = b2g.JOINT_PRISMATIC
        j["body_a"], j["body_b"] = 0, 1 + i
        j["local_anchor_a"] = anchor
        j["local_anchor_b"] = (0.0, 0.0)
        j["local_axis_a"] = (np.cos(ang), np.sin(ang))
    return bd, fb, sd, _fixture(n), jd


def _box_shapes(n, hx, hy):
    sd = np.zeros(n, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_BOX
    sd["params"][:, 0] = hx
    sd["params"][:, 1] = hy
    return sd


def friction_boxes(speeds, spins, forces, torques):
    """No gravity: one static ground and, per entry, a 1 x 0.5 m box sliding and spinning, held by
    a friction joint to the ground at its centre (each box its own island)."""
    n = len(speeds)
    bd = _bodies(1 + n)
    bd[1:]["type"] = 2
    bd[1:]["position"] = np.stack([np.arange(n) * 4.0, np.zeros(n)], 1)
    bd[1:]["linear_velocity"] = np.asarray(speeds, np.float32)
    bd[1:]["angular_velocity"] = spins
    jd = np.zeros(n, b2g.JOINT_DEF)
    jd["type"] = b2g.JOINT_FRICTION
    jd["body_a"] = 0
    jd["body_b"] = np.arange(1, n + 1)
    jd["local_anchor_a"] = bd[1:]["position"]
    jd["max_force"] = forces
    jd["max_torque"] = torques
    return bd, np.arange(1, n + 1, dtype=np.int32), _box_shapes(n, 0.5, 0.25), _fixture(n), jd


def motor_boxes(starts, angles, offsets, angle_offsets, max_force=1e6, max_torque=1e6, cf=0.3):
    """No gravity: boxes at rest driven by motor joints from the static ground (at the origin) to
    the given offsets (each box its own island)."""
    n = len(starts)
    bd = _bodies(1 + n)
    bd[1:]["type"] = 2
    bd[1:]["position"] = starts
    bd[1:]["angle"] = angles
    jd = np.zeros(n, b2g.JOINT_DEF)
    jd["type"] = b2g.JOINT_MOTOR
    jd["body_a"] = 0
    jd["body_b"] = np.arange(1, n + 1)
    jd["linear_offset"] = offsets
    jd["angular_offset"] = angle_offsets
    jd["max_force"] = max_force
    jd["max_torque"] = max_torque
    jd["correction_factor"] = cf
    return bd, np.arange(1, n + 1, dtype=np.int32), _box_shapes(n, 0.5, 0.25), _fixture(n), jd


def mouse_boxes(starts, targets, hz, zetas, max_force=1e6):
    """No gravity: boxes pulled by mouse joints (body A the static ground) toward targets with
    spring frequency hz and damping ratios zetas (stiffness m w^2, damping 2 m zeta w)."""
    n = len(starts)
    bd = _bodies(1 + n)
    bd[1:]["type"] = 2
    bd[1:]["position"] = starts
    fd = _fixture(n)
    m = float(fd["density"][0]) * 1.0 * 0.5
    w = 2.0 * np.pi * np.asarray(hz)
    jd = np.zeros(n, b2g.JOINT_DEF)
    jd["type"] = b2g.JOINT_MOUSE
    jd["body_a"] = 0
    jd["body_b"] = np.arange(1, n + 1)
    jd["target"] = starts  # at creation: the target is the grabbed point (B's anchor)
    jd["stiffness"] = m * w * w
    jd["damping"] = 2.0 * m * np.asarray(zetas) * w
    jd["max_force"] = max_force
    return bd, np.arange(1, n + 1, dtype=np.int32), _box_shapes(n, 0.5, 0.25), fd, jd, np.asarray(targets), m


def wheels_on_springs(n=3, stiffness=(40.0, 80.0, 160.0), motor_speed=0.0, torque=0.0, limit=None):
    """Circle wheels of radius 0.4 on wheel joints to the static ground, anchored at their centres
    with the vertical as the axis and springs at rest at the start, gravity on."""
    bd = _bodies(1 + n)
    bd[1:]["type"] = 2
    xs = np.arange(n) * 3.0
    bd[1:]["position"] = np.stack([xs, np.full(n, 9.0)], 1)
    sd = np.zeros(n, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_CIRCLE
    sd["radius"] = 0.4
    jd = np.zeros(n, b2g.JOINT_DEF)
    jd["type"] = b2g.JOINT_WHEEL
    jd["body_a"] = 0
    jd["body_b"] = np.arange(1, n + 1)
    jd["local_anchor_a"] = np.stack([xs, np.full(n, 9.0)], 1)
    jd["local_anchor_b"] = (0.0, 0.0)
    jd["local_axis_a"] = (0.0, 1.0)
    jd["stiffness"] = stiffness
    if torque:
        jd["enable_motor"] = 1
        jd["motor_speed"] = motor_speed
        jd["max_motor_torque"] = torque
    if limit is not None:
        jd["enable_limit"] = 1
        jd["lower_angle"], jd["upper_angle"] = limit
    return bd, np.arange(1, n + 1, dtype=np.int32), sd, _fixture(n), jd


def atwood(mass_a, mass_b, ratio):
    """Two boxes (1 x 1 m, densities giving the masses) hanging from ground anchors 4 m above their
    tops on one pulley joint, gravity on, rotation fixed."""
    bd = _bodies(3)
    bd[1:]["type"] = 2
    bd[1:]["fixed_rotation"] = 1
    bd[1]["position"] = (-2.0, 5.0)
    bd[2]["position"] = (2.0, 5.0)
    fd = _fixture(2)
    fd["density"] = (mass_a, mass_b)
    jd = np.zeros(1, b2g.JOINT_DEF)
    jd["type"] = b2g.JOINT_PULLEY
    jd["body_a"], jd["body_b"] = 1, 2
    jd["collide_connected"] = 1
    jd["ground_anchor_a"] = (-2.0, 9.5)
    jd["ground_anchor_b"] = (2.0, 9.5)
    jd["local_anchor_a"] = (0.0, 0.5)
    jd["local_anchor_b"] = (0.0, 0.5)
    jd["length_a"] = 4.0
    jd["length_b"] = 4.0
    jd["ratio"] = ratio
    return bd, np.arange(1, 3, dtype=np.int32), _box_shapes(2, 0.5, 0.5), fd, jd


def geared_wheels(ratio, motor_speed, torque=1e4, prismatic_b=False):
    """Two circles pinned to the static ground by revolute joints (or the second on a horizontal
    prismatic joint: a rack), coupled by a gear joint; a motor drives the first. No gravity."""
    bd = _bodies(3)
    bd[1:]["type"] = 2
    bd[1]["position"] = (0.0, 5.0)
    bd[2]["position"] = (3.0, 5.0)
    sd = np.zeros(2, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_CIRCLE
    sd["radius"] = (1.0, 0.5)
    jd = np.zeros(3, b2g.JOINT_DEF)
    jd[0]["type"] = b2g.JOINT_REVOLUTE
    jd[0]["body_a"], jd[0]["body_b"] = 0, 1
    jd[0]["local_anchor_a"] = (0.0, 5.0)
    jd[0]["enable_motor"], jd[0]["motor_speed"], jd[0]["max_motor_torque"] = 1, motor_speed, torque
    jd[1]["body_a"], jd[1]["body_b"] = 0, 2
    jd[1]["local_anchor_a"] = (3.0, 5.0)
    if prismatic_b:
        jd[1]["type"] = b2g.JOINT_PRISMATIC
        jd[1]["local_axis_a"] = (1.0, 0.0)
    else:
        jd[1]["type"] = b2g.JOINT_REVOLUTE
    jd[2]["type"] = b2g.JOINT_GEAR
    jd[2]["body_a"], jd[2]["body_b"] = 1, 2
    jd[2]["joint1"], jd[2]["joint2"] = 0, 1
    jd[2]["ratio"] = ratio
    return bd, np.arange(1, 3, dtype=np.int32), sd, _fixture(2), jd
