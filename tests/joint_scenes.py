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
