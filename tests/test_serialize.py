"""The serde world format (CPU): layout, number round trips, exact polygons.

* Layout: tests/golden/serde_world_small.json is written by hand from the reference's serde
  definitions -- world members (src/serialize/serialize_b2_world.rs:131-179), body record
  (src/serialize/serialize_b2_body.rs:57-70), B2bodyDef field order (src/b2_body.rs:89-148),
  fixture record (src/serialize/serialize_b2_fixture.rs:23-52), B2filter (src/b2_fixture.rs:28-42),
  shapes (src/shapes/b2_circle_shape.rs:10-16, src/shapes/b2_edge_shape.rs:14-28,
  src/shapes/serialize_b2_polygon_shape.rs:40-90, base src/b2_shape.rs:42-50), enums by variant
  name -- in serde_json's pretty layout (two-space indent), bodies newest first. The reader must
  give the definitions it holds, and the writer must give the file back byte for byte. No file
  written by the reference itself exists in its tree (its testbed writes them at run time,
  examples/testbed/test_serialize.rs:22-54), so the byte layout against serde_json's own output
  is parity unpinned; values are pinned (f32 text round trip below).
* Joints: tests/golden/serde_world_joints.json (a distance and a revolute joint, written by hand
  from src/serialize/serialize_b2_joint.rs:33-53 and src/joints/serialize/serialize_b2_distance_joint.rs:44-51,
  serialize_b2_revolute_joint.rs:47-58) reads into the expected joint definitions and writes back
  byte for byte; its body indices resolve as the reference's reader resolves them (mirrored in
  the newest-first list of the world it builds, serialize_b2_joint.rs:88-150) or as the writer
  meant them.
* Numbers: every finite f32 bit pattern drawn at random survives text -> value exactly.
* Exact polygons: a polygon stored as vertices, normals and centroid (what the format holds)
  builds the same oracle shape and mass as the hull it came from (B2polygonShape::set,
  src/private/collision/b2_polygon_shape.rs:96-211), bit for bit.
* What the GPU path does not hold (wheel and other joint kinds, chains, disabled bodies) is refused.
"""
import json
import re
import os

import numpy as np
import pytest

from b2g import serialize as ser
import b2g

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "serde_world_small.json")
GOLDEN_JOINTS = os.path.join(os.path.dirname(__file__), "golden", "serde_world_joints.json")


def test_golden_layout_reads_and_writes_back():
    text = open(GOLDEN).read()
    doc = ser.read_world(text)
    bd, fb, sd, fd = doc["body_defs"], doc["fixture_body"], doc["shape_defs"], doc["fixture_defs"]
    assert doc["gravity"] == (0.0, -10.0) and doc["checkpoint"] is None
    # file order: the box (created last), the circle, the ground (created first)
    assert list(bd["type"]) == [2, 2, 0]
    np.testing.assert_array_equal(bd["position"], np.array([[0.25, 4.0], [1.5, 2.0], [0.0, 0.0]], np.float32))
    np.testing.assert_array_equal(bd["angle"], np.float32([0.5, 0.0, 0.0]))
    np.testing.assert_array_equal(bd["linear_velocity"][0], np.float32([1.0, 0.0]))
    assert bd["angular_velocity"][0] == np.float32(-2.0)
    assert bd["linear_damping"][1] == np.float32(0.1) and bd["angular_damping"][1] == np.float32(0.05)
    assert list(bd["allow_sleep"]) == [1, 0, 1] and list(bd["awake"]) == [1, 1, 0]
    assert list(bd["fixed_rotation"]) == [0, 1, 0] and bd["gravity_scale"][1] == np.float32(0.5)
    assert list(fb) == [0, 1, 2]
    assert list(sd["type"]) == [b2g.SHAPE_POLYGON_EXACT, b2g.SHAPE_CIRCLE, b2g.SHAPE_EDGE]
    assert sd["count"][0] == 4 and sd["radius"][0] == np.float32(0.01) and sd["radius"][1] == np.float32(0.5)
    np.testing.assert_array_equal(sd["params"][0][:8], np.float32([-0.5, -0.5, 0.5, -0.5, 0.5, 0.5, -0.5, 0.5]))
    np.testing.assert_array_equal(sd["params"][0][16:24], np.float32([0, -1, 1, 0, 0, 1, -1, 0]))
    np.testing.assert_array_equal(sd["params"][2][:4], np.float32([-10, 0, 10, 0]))
    assert sd["flags"][2] == 0
    np.testing.assert_array_equal(fd["friction"], np.float32([0.6, 0.2, 0.6]))
    np.testing.assert_array_equal(fd["restitution"], np.float32([0.0, 0.3, 0.0]))
    np.testing.assert_array_equal(fd["density"], np.float32([1.0, 2.5, 0.0]))
    assert (fd["category"][1], fd["mask"][1], fd["group"][1]) == (2, 65533, -3)
    # written back in creation order (the reverse of the file's), the text is the file's
    order = np.arange(len(bd))[::-1]
    remap = np.empty_like(order)
    remap[order] = np.arange(len(order))
    out = ser.defs_to_json(doc["gravity"], bd[order], remap[fb][::-1], sd[::-1], fd[::-1])
    assert out == text


def test_write_read_round_trip_is_exact():
    from b2g import scenes
    scene = scenes.large_world(n=256, seed=3)
    # polygons as stored: build exact records from the oracle's hull of every polygon
    import b2o
    o = b2o.Oracle("f32")
    shapes, status = o.make_shapes(scene.shape_defs)
    assert np.all(status == 0)
    sd = exact_shape_defs(scene.shape_defs, shapes)
    text = ser.defs_to_json((0.0, -10.0), scene.body_defs, scene.fixture_body, sd, scene.fixture_defs)
    doc = ser.read_world(text)
    nb = len(scene.body_defs)
    back = np.arange(nb)[::-1]
    for f in ("type", "position", "angle", "linear_velocity", "angular_velocity", "linear_damping",
              "angular_damping", "allow_sleep", "fixed_rotation", "gravity_scale"):
        np.testing.assert_array_equal(doc["body_defs"][f][back], scene.body_defs[f], err_msg=f)
    # fixtures come back newest first: map them to their creation index
    fb = np.asarray(scene.fixture_body)
    per_body = [np.flatnonzero(fb == b) for b in range(nb)]
    file_order = np.concatenate([per_body[b][::-1] for b in range(nb)[::-1]])
    for f in ("friction", "restitution", "restitution_threshold", "density", "category", "mask", "group"):
        np.testing.assert_array_equal(doc["fixture_defs"][f], scene.fixture_defs[f][file_order], err_msg=f)
    got = doc["shape_defs"]
    want = sd[file_order]
    for f in ("type", "count", "flags", "radius", "params"):
        np.testing.assert_array_equal(got[f], want[f], err_msg=f)


def exact_shape_defs(shape_defs, oracle_shapes):
    """The shapes as the serde format stores them (polygons: hull vertices, normals, centroid)."""
    out = np.zeros(len(shape_defs), b2g.SHAPE_DEF)
    for i, (d, s) in enumerate(zip(shape_defs, oracle_shapes)):
        out[i]["radius"] = s["radius"]
        if int(s["type"]) == 0:
            out[i]["type"] = b2g.SHAPE_CIRCLE
            out[i]["params"][:2] = s["p"]
        elif int(s["type"]) == 1:
            out[i]["type"] = b2g.SHAPE_EDGE
            out[i]["flags"] = int(s["one_sided"])
            out[i]["params"][:8] = np.concatenate([s["v1"], s["v2"], s["v0"], s["v3"]])
        else:
            n = int(s["count"])
            out[i]["type"] = b2g.SHAPE_POLYGON_EXACT
            out[i]["count"] = n
            out[i]["params"][:2 * n] = s["vertices"][:2 * n]
            out[i]["params"][16:16 + 2 * n] = s["normals"][:2 * n]
            out[i]["params"][32:34] = s["centroid"]
    return out


def test_exact_polygon_equals_its_hull(oracle):
    from b2g import scenes
    scene = scenes.large_world(n=512, seed=5)
    poly = np.flatnonzero(np.isin(scene.shape_defs["type"], (b2g.SHAPE_POLYGON, b2g.SHAPE_BOX)))
    src = scene.shape_defs[poly]
    hull, st = oracle.make_shapes(src)
    assert np.all(st == 0)
    exact = exact_shape_defs(src, hull)
    again, st2 = oracle.make_shapes(exact)
    assert np.all(st2 == 0)
    for f in hull.dtype.names:
        np.testing.assert_array_equal(again[f], hull[f], err_msg=f)
    np.testing.assert_array_equal(oracle.shape_mass(again, 1.0), oracle.shape_mass(hull, 1.0))


def test_f32_text_round_trip():
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 2 ** 32, 60000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = vals[np.isfinite(vals)]
    for x in vals[:20000]:
        y = ser._parse_f32(ser._f32_text(x))
        assert np.float32(y).view(np.uint32) == np.float32(x).view(np.uint32), repr(x)
    # decimal text read with correct rounding (ties to even): the exact midpoint between 1 and its
    # successor 1 + 2^-23 goes to 1 (even), a hair below to 1, a hair above to the successor
    one_up = np.nextafter(np.float32(1.0), np.float32(2.0))
    assert ser._parse_f32("1.000000059604644775390625") == np.float32(1.0)
    assert ser._parse_f32("1.0000000596046447") == np.float32(1.0)
    assert ser._parse_f32("1.0000000596046448") == one_up
    # the same above 1 + 2^-23 (odd): its midpoint with 1 + 2^-22 goes up to the even one
    assert ser._parse_f32("1.000000178813934326171875") == np.float32(1.0 + 2.0 ** -22)


@pytest.mark.parametrize("what", ["joint", "chain", "disabled"])
def test_unsupported_members_are_refused(what):
    doc = json.loads(open(GOLDEN).read())
    if what == "joint":
        doc["m_joints_list"] = [{"jtype": "EMouseJoint", "joint_def": {}}]  # never written by the reference
    elif what == "chain":
        doc["m_bodies_list"][1]["m_fixture_list"][0]["m_shape_type"] = "EChain"
    else:
        doc["m_bodies_list"][0]["m_definition"]["enabled"] = False
    with pytest.raises(b2g.B2gError):
        ser.read_world(json.dumps(doc))


def test_golden_joints_read_and_write_back():
    text = open(GOLDEN_JOINTS).read()
    doc = ser.read_world(text, joint_bodies="file")
    bd, fb, sd, fd, jd = doc["body_defs"], doc["fixture_body"], doc["shape_defs"], doc["fixture_defs"], doc["joint_defs"]
    assert list(bd["type"]) == [2, 2, 0] and list(fb) == [0, 1]
    assert list(jd["type"]) == [b2g.JOINT_DISTANCE, b2g.JOINT_REVOLUTE]
    assert (jd["body_a"][0], jd["body_b"][0], jd["body_a"][1], jd["body_b"][1]) == (1, 0, 2, 1)
    assert list(jd["collide_connected"]) == [0, 1]
    np.testing.assert_array_equal(jd["length"][0], np.float32(2.828427))
    assert (jd["min_length"][0], jd["max_length"][0]) == (np.float32(2.0), np.float32(3.5))
    assert (jd["stiffness"][0], jd["damping"][0]) == (np.float32(4.0), np.float32(0.5))
    np.testing.assert_array_equal(jd["local_anchor_b"][1], np.float32([-1.0, 0.0]))
    assert jd["enable_limit"][1] == 1 and jd["enable_motor"][1] == 0
    assert (jd["lower_angle"][1], jd["upper_angle"][1]) == (np.float32(-0.5), np.float32(0.25))
    # the reference's reader: index i names the i-th body of the new world's newest-first list
    ref = ser.read_world(text)["joint_defs"]
    assert (ref["body_a"][0], ref["body_b"][0], ref["body_a"][1], ref["body_b"][1]) == (1, 2, 0, 1)
    # written back in creation order (bodies, fixtures and joints the reverse of the file's)
    n = len(bd)
    order = np.arange(n)[::-1]
    remap = np.empty_like(order)
    remap[order] = np.arange(n)
    jw = jd[::-1].copy()
    jw["body_a"] = remap[jw["body_a"]]
    jw["body_b"] = remap[jw["body_b"]]
    out = ser.defs_to_json(doc["gravity"], bd[order], remap[fb][::-1], sd[::-1], fd[::-1], joints=jw)
    assert out == text



def test_f32_text_is_serde_json_layout():
    """serde_json writes an f32 with ryu: the shortest decimal digits that read back as the same
    f32, in plain notation when -6 < kk <= 13 (10^(kk-1) <= |x| < 10^kk; integral values end in
    ".0"), else as d.ddd e(kk-1) with no plus sign and no leading exponent zeros."""
    cases = [(1.0, "1.0"), (0.5, "0.5"), (100.0, "100.0"), (1e10, "10000000000.0"), (1e13, "1e13"),
             (1.5e13, "1.5e13"), (123456789.0, "123456790.0"), (0.001234, "0.001234"), (1e-5, "0.00001"),
             (1e-6, "0.000001"), (1e-7, "1e-7"), (3.4028235e38, "3.4028235e38"), (1.17549435e-38, "1.1754944e-38"),
             (-0.0, "-0.0"), (0.0, "0.0"), (0.1, "0.1"), (-2.5, "-2.5"), (9.999999e12, "9999999000000.0")]
    for x, want in cases:
        assert ser._f32_text(x) == want, (x, ser._f32_text(x), want)
    rng = np.random.default_rng(12)
    bits = rng.integers(0, 2 ** 32, 40000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = np.concatenate([vals[np.isfinite(vals)][:8000], rng.uniform(-50, 50, 4000).astype(np.float32),
                           (10.0 ** rng.uniform(-9, 16, 4000)).astype(np.float32)])
    plain = re.compile(r"^-?(0\.0*[1-9][0-9]*|[1-9][0-9]*\.[0-9]+)$")
    sci = re.compile(r"^-?[1-9](\.[0-9]*[1-9])?e-?[1-9][0-9]?$")
    for x in vals:
        t = ser._f32_text(x)
        assert np.float32(float(t)).view(np.uint32) == x.view(np.uint32), (x, t)
        if x == 0:
            continue
        kk = int(np.floor(np.log10(abs(float(np.float32(float(t))))))) + 1
        if -6 < kk <= 13:
            assert plain.match(t), (x, t)
        else:
            assert sci.match(t), (x, t)
        # shortest: no decimal with one significant digit fewer reads back as x
        mant = t.lstrip("-").split("e")[0].replace(".", "").strip("0")
        n = len(mant)
        if n > 1:
            for cand in ("%.*e" % (n - 2, float(x)), "%.*e" % (n - 2, np.nextafter(float(x), 0.0)),
                         "%.*e" % (n - 2, np.nextafter(float(x), np.inf))):
                assert np.float32(float(cand)).view(np.uint32) != x.view(np.uint32), (x, t, cand)


def _every_joint_kind():
    """Bodies (two static, eight dynamic) and one joint of each kind, two gears (revolute-revolute,
    revolute-prismatic) after the joints they couple, and a mouse joint."""
    import joint_scenes as J  # noqa: F401  (the scenes module sits next to this file)
    bd = np.zeros(10, b2g.BODY_DEF)
    bd["type"] = 2
    bd[:2]["type"] = 0
    bd["position"] = np.stack([np.arange(10) * 2.0, np.full(10, 3.0)], 1)
    sd = np.zeros(10, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_CIRCLE
    sd["radius"] = 0.4
    fd = np.zeros(10, b2g.FIXTURE_DEF)
    fd["density"], fd["friction"], fd["mask"], fd["category"] = 1.0, 0.3, 0xFFFF, 1
    jd = np.zeros(11, b2g.JOINT_DEF)
    kinds = [b2g.JOINT_DISTANCE, b2g.JOINT_REVOLUTE, b2g.JOINT_WELD, b2g.JOINT_PRISMATIC, b2g.JOINT_WHEEL,
             b2g.JOINT_FRICTION, b2g.JOINT_MOTOR, b2g.JOINT_PULLEY, b2g.JOINT_MOUSE, b2g.JOINT_GEAR, b2g.JOINT_GEAR]
    jd["type"] = kinds
    jd["body_a"] = [0, 0, 2, 1, 3, 0, 1, 4, 0, 3, 3]
    jd["body_b"] = [2, 3, 4, 5, 6, 7, 8, 9, 9, 3, 5]
    jd["local_anchor_a"] = np.float32(np.arange(22).reshape(11, 2) * 0.25)
    jd["local_anchor_b"] = np.float32(-np.arange(22).reshape(11, 2) * 0.125)
    jd["length"], jd["min_length"], jd["max_length"] = 2.5, 1.0, 3.0
    jd["stiffness"], jd["damping"], jd["reference_angle"] = 4.0, 0.5, 0.125
    jd["lower_angle"], jd["upper_angle"], jd["max_motor_torque"], jd["motor_speed"] = -0.5, 0.75, 20.0, 1.5
    jd["enable_limit"], jd["enable_motor"] = 1, 1
    jd["local_axis_a"] = (0.6, 0.8)
    jd["max_force"], jd["max_torque"], jd["angular_offset"], jd["correction_factor"] = 30.0, 5.0, 0.25, 0.3
    jd["linear_offset"] = (1.0, -2.0)
    jd["ground_anchor_a"], jd["ground_anchor_b"] = (0.0, 10.0), (8.0, 10.0)
    jd["length_a"], jd["length_b"], jd["ratio"] = 6.0, 7.0, 2.0
    jd["target"] = (18.0, 3.0)
    jd[9]["body_a"], jd[9]["body_b"], jd[9]["joint1"], jd[9]["joint2"] = 3, 5, 1, 3  # revolute and prismatic
    jd[10]["body_a"], jd[10]["body_b"], jd[10]["joint1"], jd[10]["joint2"] = 3, 9, 1, 1
    jd["collide_connected"] = [0, 1, 0, 0, 0, 1, 0, 1, 0, 0, 0]
    return bd, np.arange(10, dtype=np.int32), sd, fd, jd


def test_every_joint_kind_round_trip():
    """Every kind the reference's writer writes (all but the mouse joint) reads back with its
    fields; gears go in their own list and name their joints by position in the joint list; the
    reference's reader resolves those positions mirrored, as it does bodies. The writer numbers
    the joint list before it leaves the mouse joints out (src/serialize/serialize_b2_world.rs:
    149-166 and 64-66), so with a mouse joint a gear's positions count it, as the reference's."""
    bd, fb, sd, fd, jd = _every_joint_kind()
    text = ser.defs_to_json((0.0, -10.0), bd, fb, sd, fd, joints=jd)
    doc = json.loads(text)
    assert [r["jtype"] for r in doc["m_joints_list"]] == [
        "EPulleyJoint", "EMotorJoint", "EFrictionJoint", "EWheelJoint", "EPrismaticJoint", "EWeldJoint",
        "ERevoluteJoint", "EDistanceJoint"]  # newest first, the mouse joint not written
    assert [r["jtype"] for r in doc["m_gear_joints_list"]] == ["EGearJoint", "EGearJoint"]
    # positions in the numbered list (newest first, the mouse joint at 0): revolute 7, prismatic 5
    assert [(r["joint_def"]["joint1"], r["joint_def"]["joint2"]) for r in doc["m_gear_joints_list"]] == [(7, 7), (7, 5)]
    assert list(doc["m_joints_list"][1]["joint_def"]) == ["base", "linear_offset", "angular_offset", "max_force",
                                                          "max_torque", "correction_factor"]
    assert list(doc["m_joints_list"][0]["joint_def"])[:3] == ["base", "ground_anchor_a", "ground_anchor_b"]
    # without the mouse joint the positions read back as the joints they name
    keep = [i for i in range(len(jd)) if jd[i]["type"] != b2g.JOINT_MOUSE]
    jd = jd[keep].copy()  # gears now 8 and 9, coupling creation indices 1 and 3 as before
    text = ser.defs_to_json((0.0, -10.0), bd, fb, sd, fd, joints=jd)
    got = ser.read_world(text, joint_bodies="file")["joint_defs"]
    nb = len(bd)
    written = [7, 6, 5, 4, 3, 2, 1, 0, 9, 8]  # creation index of each read joint
    assert list(got["type"]) == list(jd["type"][written])
    np.testing.assert_array_equal(got["body_a"], nb - 1 - jd["body_a"][written])
    np.testing.assert_array_equal(got["body_b"], nb - 1 - jd["body_b"][written])
    for k, i in enumerate(written[:8]):
        t = int(jd[i]["type"])
        fields = {b2g.JOINT_WHEEL: ["local_anchor_a", "local_anchor_b", "local_axis_a", "lower_angle", "upper_angle",
                                    "motor_speed", "max_motor_torque", "stiffness", "damping"],
                  b2g.JOINT_FRICTION: ["local_anchor_a", "local_anchor_b", "max_force", "max_torque"],
                  b2g.JOINT_MOTOR: ["linear_offset", "angular_offset", "max_force", "max_torque", "correction_factor"],
                  b2g.JOINT_PULLEY: ["ground_anchor_a", "ground_anchor_b", "local_anchor_a", "local_anchor_b",
                                     "length_a", "length_b", "ratio"]}.get(t, [])
        for f in fields:
            np.testing.assert_array_equal(got[k][f], jd[i][f], err_msg="%d %s" % (t, f))
    # the gears' joints as creation indices of the read world: the revolute (creation 1) is read
    # 7th (index 6), the prismatic (creation 3) 5th (index 4)
    assert (got[8]["joint1"], got[8]["joint2"], got[9]["joint1"], got[9]["joint2"]) == (6, 6, 6, 4)
    assert got[9]["ratio"] == np.float32(2.0)
    ref = ser.read_world(text)["joint_defs"]
    assert (ref[8]["joint1"], ref[9]["joint2"]) == (8 - 1 - 6, 8 - 1 - 4)
