"""The reference's serde world format (JSON): read, write, and a checkpoint extension.

A world is written the way ``serde_json::to_writer_pretty(&*world)`` writes a ``B2world``
(examples/testbed/test_serialize.rs:57-60) and read the way ``B2worldDeserializeResult`` reads it:

* world: ``{"m_gravity", "m_bodies_list", "m_joints_list", "m_gear_joints_list"}``
  (src/serialize/serialize_b2_world.rs:131-179; joints are written as ``{"jtype", "joint_def"}``);
* body: ``{"m_definition": B2bodyDef, "m_fixture_list": [...]}`` (src/serialize/serialize_b2_body.rs:37-70);
  the definition holds the body's current transform position ``m_xf.p``, sweep angle, velocities
  and flags, with the B2bodyDef fields in declaration order (src/b2_body.rs:89-148);
* fixture: ``{"m_friction", "m_restitution", "m_restitution_threshold", "m_density",
  "m_is_sensor", "m_filter", "m_shape_type", "m_shape"}`` (src/serialize/serialize_b2_fixture.rs:23-52);
* shapes: circle ``{"base", "m_p"}`` (src/shapes/b2_circle_shape.rs:10-16), edge ``{"base",
  "m_vertex1", "m_vertex2", "m_vertex0", "m_vertex3", "m_one_sided"}`` (src/shapes/b2_edge_shape.rs:14-28),
  polygon ``{"base", "m_centroid", "m_count", "m_vertices", "m_normals"}``
  (src/shapes/serialize_b2_polygon_shape.rs:40-90), ``base`` = ``{"m_type", "m_radius"}``
  (src/b2_shape.rs:42-50); enums by variant name (``B2DynamicBody``, ``EPolygon``).

The reference's body list and each body's fixture list are newest first (``push_front``), so a
file lists them newest first, and a reader creates them in file order (so a world read back has
its lists in the opposite order). Polygons are taken as stored (vertices, normals, centroid:
``SHAPE_POLYGON_EXACT``), not re-hulled.

Numbers are written as serde_json writes an f32 (ryu's shortest round-trip digits in ryu's f32
layout, see _f32_text) and read with correct rounding to f32, so values survive a round trip
exactly and the text is what serde_json would write for the same values.

Checkpoint extension (a ``"b2g_checkpoint"`` member of the world object): the state a GPU world
needs to continue exactly as the written one would: every body's full state (transform, sweep
with c0/a0, velocities, sleep timer and flag), every proxy's fat AABB, every contact's manifold
with its impulses and the solve colour its predecessor had, the touching contacts that left the
set (the next collide wakes their bodies), the solver scalars (warm start ratio, colour-reuse
counters) and the world flags, with the creation order of bodies and fixtures. A world read with
the extension is created in the original order and continues bit-identically. The reference's
reader rejects unknown members (its field enums have no catch-all variant,
src/serialize/serialize_b2_world.rs:189-198), so only a file written without the extension is
for the reference.

Joints are ``{"jtype", "joint_def"}`` records in the world's joint list, newest first (src/serialize/serialize_b2_world.rs:58-129), each definition with its
``base`` (``jtype``, ``user_data``, ``body_a``, ``body_b``, ``collide_connected``,
src/serialize/serialize_b2_joint.rs:33-53) and the kind's fields in their serializer's order
(src/joints/serialize/serialize_b2_*_joint.rs). A joint names its bodies by index: the writer
gives a body's position in the written body list (serialize_b2_world.rs:135-138); the
reference's reader looks that index up in the list of the world it has just built, which is in
the opposite order (serialize_b2_joint.rs:88-150 over a newest-first list), so a round trip
through the reference attaches a joint to the mirrored body. ``read_world`` follows the
reference's reader by default (``joint_bodies="reference"``); ``joint_bodies="file"`` takes the
index as the writer meant it. Gear joints go in their own list after the others and name their
two joints by position in the written joint list (src/serialize/serialize_b2_world.rs:144-174),
which the reference's reader also looks up mirrored; a mouse joint is not written (the
reference's writer skips it). A checkpoint does not hold gear or mouse joints. Chain shapes and
disabled bodies are not part of the GPU path: reading them raises.
"""
import json
from fractions import Fraction

import numpy as np

from . import (BODY_DEF, FIXTURE_DEF, JOINT_DEF, JOINT_DISTANCE, JOINT_FRICTION, JOINT_GEAR, JOINT_MOTOR, JOINT_MOUSE,
               JOINT_PRISMATIC, JOINT_PULLEY, JOINT_REVOLUTE, JOINT_WELD, JOINT_WHEEL, MANIFOLD, SHAPE_DEF, SHAPE_CIRCLE,
               SHAPE_EDGE, SHAPE_POLYGON_EXACT, B2gError, ERR_UNSUPPORTED, World)

BODY_TYPES = ("B2StaticBody", "B2KinematicBody", "B2DynamicBody")          # src/b2_body.rs:25-31
SHAPE_TYPES = ("ECircle", "EEdge", "EPolygon", "EChain")                  # src/b2_shape.rs:23-31
MANIFOLD_TYPES = ("ECircles", "EFaceA", "EFaceB")                         # src/b2_collision.rs:63-67
JOINT_TYPES = {JOINT_DISTANCE: "EDistanceJoint", JOINT_PRISMATIC: "EPrismaticJoint",  # src/b2_joint.rs:46-58
               JOINT_REVOLUTE: "ERevoluteJoint", JOINT_WELD: "EWeldJoint", JOINT_WHEEL: "EWheelJoint",
               JOINT_FRICTION: "EFrictionJoint", JOINT_MOTOR: "EMotorJoint", JOINT_PULLEY: "EPulleyJoint",
               JOINT_GEAR: "EGearJoint", JOINT_MOUSE: "EMouseJoint"}
LINEAR_SLOP = np.float32(0.005)


# ---- f32 numbers ------------------------------------------------------------------------------
class F32(float):
    """A float written as the shortest decimal that reads back as the same f32."""


def _f32_text(x):
    """An f32 as serde_json writes it: ryu's shortest decimal digits that read back as the same
    f32, laid out as ryu's f32 formatter does (the ryu crate, pretty::format32): with the digits d
    and the power of ten kk such that 10^(kk-1) <= |x| < 10^kk, plain notation when
    -6 < kk <= 13 ("1234.5", "0.001234", integral values with ".0"), otherwise d.ddd followed by
    "e" and the exponent kk - 1 without a plus sign or leading zeros ("1e-7", "1.5e13")."""
    v = np.float32(x)
    if not np.isfinite(v):
        raise ValueError("serde world: non-finite number %r" % x)
    sign = "-" if np.signbit(v) else ""
    if v == 0:
        return sign + "0.0"
    mant, exp = np.format_float_scientific(abs(v), unique=True, trim="-").split("e")
    digits = mant.replace(".", "")
    length = len(digits)
    kk = int(exp) + 1
    k = kk - length  # |x| = digits x 10^k
    if 0 <= k and kk <= 13:
        text = digits + "0" * (kk - length) + ".0"
    elif 0 < kk <= 13:
        text = digits[:kk] + "." + digits[kk:]
    elif -6 < kk <= 0:
        text = "0." + "0" * (-kk) + digits
    elif length == 1:
        text = digits + "e" + str(kk - 1)
    else:
        text = digits[0] + "." + digits[1:] + "e" + str(kk - 1)
    return sign + text


def _parse_f32(text):
    """Decimal text -> the nearest f32 (ties to even), exactly."""
    x = float(text)
    f = np.float32(x)
    if not np.isfinite(f):
        return f
    # a double-rounding error is only possible when x is next to a midpoint between f32 values
    lo, hi = np.nextafter(f, np.float32(-np.inf)), np.nextafter(f, np.float32(np.inf))
    near = False
    for m in ((float(f) + float(lo)) / 2.0, (float(f) + float(hi)) / 2.0):
        if abs(x - m) <= np.spacing(abs(m)) * 2:
            near = True
    if not near:
        return f
    exact = Fraction(text)
    best = min((abs(Fraction(float(c)) - exact), int(np.float32(c).view(np.uint32)) & 1, c) for c in (lo, f, hi))
    return np.float32(best[2])


def _dump(v, indent=0):
    """serde_json's pretty layout: two-space indent, one member or element per line."""
    pad, inner = "  " * indent, "  " * (indent + 1)
    if isinstance(v, dict):
        if not v:
            return "{}"
        items = ['%s"%s": %s' % (inner, k, _dump(x, indent + 1)) for k, x in v.items()]
        return "{\n" + ",\n".join(items) + "\n" + pad + "}"
    if isinstance(v, (list, tuple)):
        if not v:
            return "[]"
        return "[\n" + ",\n".join(inner + _dump(x, indent + 1) for x in v) + "\n" + pad + "]"
    if isinstance(v, bool):
        return "true" if v else "false"
    if v is None:
        return "null"
    if isinstance(v, F32):
        return _f32_text(v)
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, str):
        return json.dumps(v)
    raise TypeError("serde world: cannot write %r" % (v,))


def _vec(x, y):
    return {"x": F32(x), "y": F32(y)}


def _getvec(d):
    return _parse_f32_any(d["x"]), _parse_f32_any(d["y"])


def _parse_f32_any(x):
    return x if isinstance(x, np.float32) else np.float32(x)


# ---- writing -----------------------------------------------------------------------------------
def _shape_json(sd):
    t = int(sd["type"])
    r = F32(sd["radius"])
    p = sd["params"]
    if t == SHAPE_CIRCLE:
        return "ECircle", {"base": {"m_type": "ECircle", "m_radius": r}, "m_p": _vec(p[0], p[1])}
    if t == SHAPE_EDGE:
        return "EEdge", {"base": {"m_type": "EEdge", "m_radius": r}, "m_vertex1": _vec(p[0], p[1]),
                         "m_vertex2": _vec(p[2], p[3]), "m_vertex0": _vec(p[4], p[5]), "m_vertex3": _vec(p[6], p[7]),
                         "m_one_sided": bool(int(sd["flags"]) & 1)}
    if t == SHAPE_POLYGON_EXACT:
        n = int(sd["count"])
        return "EPolygon", {"base": {"m_type": "EPolygon", "m_radius": r}, "m_centroid": _vec(p[32], p[33]),
                            "m_count": n, "m_vertices": [_vec(p[2 * i], p[2 * i + 1]) for i in range(n)],
                            "m_normals": [_vec(p[16 + 2 * i], p[17 + 2 * i]) for i in range(n)]}
    raise B2gError(ERR_UNSUPPORTED, "serde world: shape kind %d" % t)


def _manifold_json(m):
    pts = []
    for i in range(2):
        q = m["points"][i]
        ident = int(q["id"])
        pts.append({"local_point": _vec(*q["local_point"]), "normal_impulse": F32(q["normal_impulse"]),
                    "tangent_impulse": F32(q["tangent_impulse"]),
                    "id": {"index_a": ident & 0xff, "index_b": (ident >> 8) & 0xff, "type_a": (ident >> 16) & 0xff,
                           "type_b": (ident >> 24) & 0xff}})
    return {"points": pts, "local_normal": _vec(*m["local_normal"]), "local_point": _vec(*m["local_point"]),
            "m_type": MANIFOLD_TYPES[int(m["type"])], "point_count": int(m["point_count"])}


def _joint_json(d, nb, joint_pos=None):
    """A joint record; body indices are positions in the written (newest-first) body list, a
    gear's joints positions in the written joint list (joint_pos: creation index -> position)."""
    t = int(d["type"])
    base = {"jtype": JOINT_TYPES[t], "user_data": None, "body_a": nb - 1 - int(d["body_a"]),
            "body_b": nb - 1 - int(d["body_b"]), "collide_connected": bool(d["collide_connected"])}
    out = {"base": base}
    if t not in (JOINT_MOTOR, JOINT_GEAR, JOINT_PULLEY):
        out.update({"local_anchor_a": _vec(*d["local_anchor_a"]), "local_anchor_b": _vec(*d["local_anchor_b"])})
    if t == JOINT_DISTANCE:
        # the joint's lengths as its constructor clamps them (b2_distance_joint.rs:44-78)
        length = max(np.float32(d["length"]), LINEAR_SLOP)
        lo = max(np.float32(d["min_length"]), LINEAR_SLOP)
        hi = max(np.float32(d["max_length"]), lo)
        out.update({"length": F32(length), "min_length": F32(lo), "max_length": F32(hi),
                    "stiffness": F32(d["stiffness"]), "damping": F32(d["damping"])})
    elif t == JOINT_REVOLUTE:
        out.update({"reference_angle": F32(d["reference_angle"]), "enable_limit": bool(d["enable_limit"]),
                    "lower_angle": F32(d["lower_angle"]), "upper_angle": F32(d["upper_angle"]),
                    "enable_motor": bool(d["enable_motor"]), "motor_speed": F32(d["motor_speed"]),
                    "max_motor_torque": F32(d["max_motor_torque"])})
    elif t == JOINT_WELD:
        out.update({"reference_angle": F32(d["reference_angle"]), "stiffness": F32(d["stiffness"]),
                    "damping": F32(d["damping"])})
    elif t == JOINT_PRISMATIC:
        # the axis as the constructor normalises it (B2vec2::normalize, src/b2_math.rs:82-92)
        ax = np.array(d["local_axis_a"], np.float32)
        length = np.sqrt(ax[0] * ax[0] + ax[1] * ax[1])
        if length >= np.finfo(np.float32).eps:
            inv = np.float32(1.0) / length
            ax = np.array([ax[0] * inv, ax[1] * inv], np.float32)
        out.update({"local_axis_a": _vec(*ax), "reference_angle": F32(d["reference_angle"]),
                    "enable_limit": bool(d["enable_limit"]), "lower_translation": F32(d["lower_angle"]),
                    "upper_translation": F32(d["upper_angle"]), "enable_motor": bool(d["enable_motor"]),
                    "motor_speed": F32(d["motor_speed"]), "max_motor_force": F32(d["max_motor_torque"])})
    elif t == JOINT_WHEEL:  # src/joints/serialize/serialize_b2_wheel_joint.rs:44-61 (the axis as given)
        out.update({"local_axis_a": _vec(*d["local_axis_a"]), "enable_limit": bool(d["enable_limit"]),
                    "lower_translation": F32(d["lower_angle"]), "upper_translation": F32(d["upper_angle"]),
                    "enable_motor": bool(d["enable_motor"]), "motor_speed": F32(d["motor_speed"]),
                    "max_motor_torque": F32(d["max_motor_torque"]), "stiffness": F32(d["stiffness"]),
                    "damping": F32(d["damping"])})
    elif t == JOINT_FRICTION:  # serialize_b2_friction_joint.rs:37-47
        out.update({"max_force": F32(d["max_force"]), "max_torque": F32(d["max_torque"])})
    elif t == JOINT_MOTOR:  # serialize_b2_motor_joint.rs:38-49
        out.update({"linear_offset": _vec(*d["linear_offset"]), "angular_offset": F32(d["angular_offset"]),
                    "max_force": F32(d["max_force"]), "max_torque": F32(d["max_torque"]),
                    "correction_factor": F32(d["correction_factor"])})
    elif t == JOINT_PULLEY:  # serialize_b2_pulley_joint.rs:40-53
        out.update({"ground_anchor_a": _vec(*d["ground_anchor_a"]), "ground_anchor_b": _vec(*d["ground_anchor_b"]),
                    "local_anchor_a": _vec(*d["local_anchor_a"]), "local_anchor_b": _vec(*d["local_anchor_b"]),
                    "length_a": F32(d["length_a"]), "length_b": F32(d["length_b"]), "ratio": F32(d["ratio"])})
    elif t == JOINT_GEAR:  # serialize_b2_gear_joint.rs:36-55: the coupled joints by list position
        out.update({"joint1": joint_pos[int(d["joint1"])], "joint2": joint_pos[int(d["joint2"])],
                    "ratio": F32(d["ratio"])})
    else:
        raise B2gError(ERR_UNSUPPORTED, "serde world: joint kind %d" % t)
    return {"jtype": JOINT_TYPES[t], "joint_def": out}


def _joint_lists(joints, nb):
    """The world's joint list and gear joint list as the reference writes them
    (src/serialize/serialize_b2_world.rs:144-174): both newest first; a joint's position in the
    non-gear list is its index (a mouse joint counts but writes nothing, serialize_b2_world.rs:64-66)."""
    if joints is None:
        return [], [], []
    newest_first = list(reversed(range(len(joints))))
    non_gear = [i for i in newest_first if int(joints[i]["type"]) != JOINT_GEAR]
    gears = [i for i in newest_first if int(joints[i]["type"]) == JOINT_GEAR]
    pos = {i: k for k, i in enumerate(non_gear)}
    written = [i for i in non_gear if int(joints[i]["type"]) != JOINT_MOUSE]
    jl = [_joint_json(joints[i], nb) for i in written]
    gl = [_joint_json(joints[i], nb, pos) for i in gears]
    return jl, gl, written + gears


def defs_to_json(gravity, body_defs, fixture_body, shape_defs, fixture_defs, states=None, extension=None, joints=None):
    """serde world JSON of bodies and fixtures given in creation order. shape_defs hold the shapes
    as stored (circles, edges, SHAPE_POLYGON_EXACT polygons); states (BODY_STATE, optional) give
    the current transform, velocities and awake flags, else the definitions' own."""
    bd, fbody, fd = body_defs, np.asarray(fixture_body), fixture_defs
    per_body = [[] for _ in range(len(bd))]
    for f in range(len(fbody)):
        per_body[int(fbody[f])].append(f)
    bodies = []
    for b in reversed(range(len(bd))):  # the body list is newest first
        d = bd[b]
        if states is not None:
            s = states[b]
            pos, angle, vel, omega, awake = s["p"], s["a"], s["v"], s["w"], s["awake"] != 0
        else:
            pos, angle, vel, omega = d["position"], d["angle"], d["linear_velocity"], d["angular_velocity"]
            awake = bool(d["awake"])
        definition = {
            "body_type": BODY_TYPES[int(d["type"])],
            "position": _vec(*pos),
            "angle": F32(angle),
            "linear_velocity": _vec(*vel),
            "angular_velocity": F32(omega),
            "linear_damping": F32(d["linear_damping"]),
            "angular_damping": F32(d["angular_damping"]),
            "allow_sleep": bool(d["allow_sleep"]),
            # B2body::new sets the awake flag only on non-static bodies (src/private/dynamics/b2_body.rs:34-36)
            "awake": bool(awake) and int(d["type"]) != 0,
            "fixed_rotation": bool(d["fixed_rotation"]),
            "bullet": bool(d["bullet"]),
            "enabled": True,
            "user_data": None,
            "gravity_scale": F32(d["gravity_scale"]),
        }
        fixtures = []
        for f in reversed(per_body[b]):  # the fixture list is newest first
            kind, shape = _shape_json(shape_defs[f])
            x = fd[f]
            fixtures.append({
                "m_friction": F32(x["friction"]), "m_restitution": F32(x["restitution"]),
                "m_restitution_threshold": F32(x["restitution_threshold"]), "m_density": F32(x["density"]),
                "m_is_sensor": bool(x["is_sensor"]),
                "m_filter": {"category_bits": int(x["category"]), "mask_bits": int(x["mask"]),
                             "group_index": int(x["group"])},
                "m_shape_type": kind, "m_shape": shape})
        bodies.append({"m_definition": definition, "m_fixture_list": fixtures})
    jl, gl, _ = _joint_lists(joints, len(bd))
    out = {"m_gravity": _vec(*gravity), "m_bodies_list": bodies, "m_joints_list": jl, "m_gear_joints_list": gl}
    if extension is not None:
        out["b2g_checkpoint"] = extension(per_body)
    return _dump(out) + "\n"


def world_to_json(world, checkpoint=False):
    """The serde JSON text of a GPU world (one world, not a batch); with checkpoint=True the
    b2g_checkpoint extension is added."""
    if world._batched:
        raise B2gError(ERR_UNSUPPORTED, "serde world: a batch of worlds has no single-world record")
    bd, fbody, fd = world.definitions()
    return defs_to_json(world.gravity, bd, fbody, world.shapes(), fd, states=world.body_states(),
                        extension=(lambda per_body: _checkpoint(world, fbody, per_body)) if checkpoint else None,
                        joints=world.joint_definitions() if world.joint_count else None)


def _checkpoint(world, fbody, per_body):
    if world.joint_count:
        kinds = world.joint_definitions()["type"]
        if (kinds == JOINT_GEAR).any() or (kinds == JOINT_MOUSE).any():
            # a gear's constant comes from the poses when it was created, and the reference's format
            # writes no mouse joint: neither could be restored exactly
            raise B2gError(ERR_UNSUPPORTED, "serde checkpoint: a world with gear or mouse joints")
    st = world.body_states()
    fat = world.fat_aabbs()
    dfab, dman = world.destroyed_contacts()
    fab, _, man = world.contacts()
    col = world.contact_colours()
    solver = world.solver_state()
    # file position -> creation index, for bodies and (per body, in file order) fixtures
    body_order = list(reversed(range(len(per_body))))
    fixture_order = [f for b in body_order for f in reversed(per_body[b])]
    return {
        "version": 1,
        "flags": dict(world.flags),
        "body_index": body_order,
        "fixture_index": fixture_order,
        "bodies": [{"p": _vec(*s["p"]), "q": {"s": F32(s["q"][0]), "c": F32(s["q"][1])}, "c": _vec(*s["c"]),
                    "a": F32(s["a"]), "v": _vec(*s["v"]), "w": F32(s["w"]), "c0": _vec(*s["c0"]), "a0": F32(s["a0"]),
                    "sleep_time": F32(s["sleep_time"]), "awake": bool(s["awake"] != 0)} for s in st],
        "fat_aabbs": [[F32(v) for v in a] for a in np.asarray(fat).reshape(-1, 4)],
        "contacts": [{"fixture_a": int(fab[k, 0]), "fixture_b": int(fab[k, 1]), "colour": int(col[k]),
                      "manifold": _manifold_json(man[k])} for k in range(len(fab))],
        "destroyed": [{"fixture_a": int(dfab[k, 0]), "fixture_b": int(dfab[k, 1]), "manifold": _manifold_json(dman[k])}
                      for k in range(len(dfab))],
        "joint_index": _joint_lists(world.joint_definitions(), len(per_body))[2] if world.joint_count else [],
        # prismatic axes as given (the record holds the normalised axis; normalising it again could
        # move a last bit)
        "joint_axes": [_vec(*a) for a in world.joint_definitions()["local_axis_a"]] if world.joint_count else [],
        "joint_impulses": [[F32(v) for v in row] for row in (world.joint_impulses() if world.joint_count else [])],
        "solver": {"inv_dt0": F32(solver["inv_dt0"]), "n_colours": int(solver["n_colours"]),
                   "scratch_colours": int(solver["scratch_colours"]),
                   "steps_since_scratch": int(solver["steps_since_scratch"])},
    }


# ---- reading -----------------------------------------------------------------------------------
def _loads(text):
    return json.loads(text, parse_float=_parse_f32, parse_int=int)


def _read_joint(rec, nb, joint_bodies):
    kind = rec["jtype"]
    names = {v: k for k, v in JOINT_TYPES.items() if k != JOINT_MOUSE}
    if kind not in names:
        raise B2gError(ERR_UNSUPPORTED, "serde world: %s is not on the GPU path" % kind)
    d = rec["joint_def"]
    base = d["base"]
    j = np.zeros((), JOINT_DEF)
    j["type"] = names[kind]
    ia, ib = int(base["body_a"]), int(base["body_b"])
    if joint_bodies == "reference":
        # the reference's reader: the index into the newly built world's newest-first body list,
        # whose bodies were created in file order
        ia, ib = nb - 1 - ia, nb - 1 - ib
    j["body_a"], j["body_b"] = ia, ib
    j["collide_connected"] = int(bool(base["collide_connected"]))
    if "local_anchor_a" in d:
        j["local_anchor_a"] = _getvec(d["local_anchor_a"])
        j["local_anchor_b"] = _getvec(d["local_anchor_b"])
    f = _parse_f32_any
    if kind == "EDistanceJoint":
        for key in ("length", "min_length", "max_length", "stiffness", "damping"):
            j[key] = f(d[key])
    elif kind == "ERevoluteJoint":
        for key in ("reference_angle", "lower_angle", "upper_angle", "motor_speed", "max_motor_torque"):
            j[key] = f(d[key])
        j["enable_limit"], j["enable_motor"] = int(bool(d["enable_limit"])), int(bool(d["enable_motor"]))
    elif kind == "EWeldJoint":
        for key in ("reference_angle", "stiffness", "damping"):
            j[key] = f(d[key])
    elif kind == "EWheelJoint":
        j["local_axis_a"] = _getvec(d["local_axis_a"])
        j["lower_angle"], j["upper_angle"] = f(d["lower_translation"]), f(d["upper_translation"])
        for key in ("motor_speed", "max_motor_torque", "stiffness", "damping"):
            j[key] = f(d[key])
        j["enable_limit"], j["enable_motor"] = int(bool(d["enable_limit"])), int(bool(d["enable_motor"]))
    elif kind == "EFrictionJoint":
        j["max_force"], j["max_torque"] = f(d["max_force"]), f(d["max_torque"])
    elif kind == "EMotorJoint":
        j["linear_offset"] = _getvec(d["linear_offset"])
        for key in ("angular_offset", "max_force", "max_torque", "correction_factor"):
            j[key] = f(d[key])
    elif kind == "EPulleyJoint":
        j["ground_anchor_a"], j["ground_anchor_b"] = _getvec(d["ground_anchor_a"]), _getvec(d["ground_anchor_b"])
        for key in ("length_a", "length_b", "ratio"):
            j[key] = f(d[key])
    elif kind == "EGearJoint":
        # list positions; read_world turns them into creation indices
        j["joint1"], j["joint2"], j["ratio"] = int(d["joint1"]), int(d["joint2"]), f(d["ratio"])
    else:
        j["local_axis_a"] = _getvec(d["local_axis_a"])
        j["reference_angle"] = f(d["reference_angle"])
        j["lower_angle"], j["upper_angle"] = f(d["lower_translation"]), f(d["upper_translation"])
        j["motor_speed"], j["max_motor_torque"] = f(d["motor_speed"]), f(d["max_motor_force"])
        j["enable_limit"], j["enable_motor"] = int(bool(d["enable_limit"])), int(bool(d["enable_motor"]))
    return j


def read_world(text, joint_bodies="reference"):
    """Parse serde world JSON -> dict with gravity, body_defs, fixture_body, shape_defs,
    fixture_defs, joint_defs (in the order the reference's reader creates them: file order) and
    the checkpoint extension (or None). joint_bodies: "reference" resolves a joint's body index
    as the reference's reader does, "file" as the writer meant it (a position in the file's body
    list)."""
    if joint_bodies not in ("reference", "file"):
        raise ValueError("joint_bodies: 'reference' or 'file'")
    doc = _loads(text)
    bodies = doc.get("m_bodies_list", [])
    bd = np.zeros(len(bodies), BODY_DEF)
    fixture_body, sds, fds = [], [], []
    for i, body in enumerate(bodies):
        d = body["m_definition"]
        if not d.get("enabled", True):
            raise B2gError(ERR_UNSUPPORTED, "serde world: disabled bodies are not on the GPU path (body %d)" % i)
        bd[i]["type"] = BODY_TYPES.index(d["body_type"])
        bd[i]["position"] = _getvec(d["position"])
        bd[i]["angle"] = _parse_f32_any(d["angle"])
        bd[i]["linear_velocity"] = _getvec(d["linear_velocity"])
        bd[i]["angular_velocity"] = _parse_f32_any(d["angular_velocity"])
        bd[i]["linear_damping"] = _parse_f32_any(d["linear_damping"])
        bd[i]["angular_damping"] = _parse_f32_any(d["angular_damping"])
        bd[i]["allow_sleep"] = int(bool(d["allow_sleep"]))
        bd[i]["awake"] = int(bool(d["awake"]))
        bd[i]["fixed_rotation"] = int(bool(d["fixed_rotation"]))
        bd[i]["bullet"] = int(bool(d.get("bullet", False)))
        bd[i]["gravity_scale"] = _parse_f32_any(d["gravity_scale"])
        for fx in body.get("m_fixture_list", []):
            sds.append(_read_shape(fx["m_shape_type"], fx["m_shape"]))
            f = np.zeros((), FIXTURE_DEF)
            f["friction"] = _parse_f32_any(fx["m_friction"])
            f["restitution"] = _parse_f32_any(fx["m_restitution"])
            f["restitution_threshold"] = _parse_f32_any(fx["m_restitution_threshold"])
            f["density"] = _parse_f32_any(fx["m_density"])
            f["is_sensor"] = int(bool(fx["m_is_sensor"]))
            flt = fx["m_filter"]
            f["category"] = int(flt["category_bits"])
            f["mask"] = int(flt["mask_bits"])
            f["group"] = int(flt["group_index"])
            fds.append(f)
            fixture_body.append(i)
    sd = np.array(sds, SHAPE_DEF) if sds else np.zeros(0, SHAPE_DEF)
    fd = np.array(fds, FIXTURE_DEF) if fds else np.zeros(0, FIXTURE_DEF)
    jl = doc.get("m_joints_list", [])
    gl = doc.get("m_gear_joints_list", [])
    recs = [_read_joint(r, len(bodies), joint_bodies) for r in jl + gl]
    jd = np.array(recs, JOINT_DEF) if recs else np.zeros(0, JOINT_DEF)
    for g in range(len(jl), len(recs)):
        # a gear's joints: the reference's reader looks the positions up in the newest-first joint
        # list of the world it has built from m_joints_list in file order (mirrored, as bodies are)
        for key in ("joint1", "joint2"):
            k = int(jd[g][key])
            if not 0 <= k < len(jl):
                raise B2gError(ERR_UNSUPPORTED, "serde world: gear joint %d names joint %d" % (g - len(jl), k))
            jd[g][key] = len(jl) - 1 - k if joint_bodies == "reference" else k
    return {"gravity": _getvec(doc["m_gravity"]), "body_defs": bd, "fixture_body": np.array(fixture_body, np.int32),
            "shape_defs": sd, "fixture_defs": fd, "joint_defs": jd, "checkpoint": doc.get("b2g_checkpoint")}


def _read_shape(kind, s):
    out = np.zeros((), SHAPE_DEF)
    out["radius"] = _parse_f32_any(s["base"]["m_radius"])
    p = out["params"]
    if kind == "ECircle":
        out["type"] = SHAPE_CIRCLE
        p[0:2] = _getvec(s["m_p"])
    elif kind == "EEdge":
        out["type"] = SHAPE_EDGE
        out["flags"] = int(bool(s["m_one_sided"]))
        for i, key in enumerate(("m_vertex1", "m_vertex2", "m_vertex0", "m_vertex3")):
            p[2 * i:2 * i + 2] = _getvec(s[key])
    elif kind == "EPolygon":
        verts, norms = s["m_vertices"], s["m_normals"]
        # the reader's count is the length of the last array read (serialize_b2_polygon_shape.rs:176-186)
        n = len(norms)
        if len(verts) < n or not 3 <= n <= 8:
            raise B2gError(ERR_UNSUPPORTED, "serde world: polygon with %d vertices, %d normals" % (len(verts), n))
        out["type"] = SHAPE_POLYGON_EXACT
        out["count"] = n
        for i in range(n):
            p[2 * i:2 * i + 2] = _getvec(verts[i])
            p[16 + 2 * i:18 + 2 * i] = _getvec(norms[i])
        p[32:34] = _getvec(s["m_centroid"])
    else:
        raise B2gError(ERR_UNSUPPORTED, "serde world: %s shapes are not on the GPU path" % kind)
    return out


def world_from_json(text, joint_bodies="reference"):
    """A GPU world from serde world JSON. Without the checkpoint extension: the reference's
    reader (bodies, fixtures and joints created in file order, fresh proxies and contacts; joint
    body indices resolved as joint_bodies says). With it: the written world's creation order and
    its full state, so that stepping continues exactly."""
    ck_probe = _loads(text).get("b2g_checkpoint")
    doc = read_world(text, joint_bodies="file" if ck_probe is not None else joint_bodies)
    ck = doc["checkpoint"]
    w = World(doc["gravity"])
    if ck is None:  # the reference's reader makes a B2world::new world: the library's defaults
        w.add(doc["body_defs"], doc["fixture_body"], doc["shape_defs"], doc["fixture_defs"])
        if len(doc["joint_defs"]):
            w.add_joints(doc["joint_defs"])
        return w
    fl = ck["flags"]
    w.set_flags(warm_starting=fl["warm_starting"], allow_sleep=fl["allow_sleep"], continuous=fl["continuous"],
                block_solve=fl["block_solve"])
    nb, nf = len(doc["body_defs"]), len(doc["fixture_defs"])
    body_index = np.asarray(ck["body_index"], np.int64)
    fixture_index = np.asarray(ck["fixture_index"], np.int64)
    if sorted(body_index.tolist()) != list(range(nb)) or sorted(fixture_index.tolist()) != list(range(nf)):
        raise ValueError("serde world: checkpoint creation orders are not permutations")
    bd = np.zeros(nb, BODY_DEF)
    bd[body_index] = doc["body_defs"]
    sd = np.zeros(nf, SHAPE_DEF)
    fd = np.zeros(nf, FIXTURE_DEF)
    fb = np.zeros(nf, np.int32)
    sd[fixture_index] = doc["shape_defs"]
    fd[fixture_index] = doc["fixture_defs"]
    fb[fixture_index] = body_index[doc["fixture_body"]]
    w.add(bd, fb, sd, fd)
    jdf = doc["joint_defs"]
    if len(jdf):
        # joints in their creation order, bodies as creation indices
        joint_index = np.asarray(ck["joint_index"], np.int64)
        jd = np.zeros(len(jdf), JOINT_DEF)
        jd[joint_index] = jdf
        jd["body_a"] = body_index[jd["body_a"]]
        jd["body_b"] = body_index[jd["body_b"]]
        jd["local_axis_a"] = np.array([_getvec(a) for a in ck["joint_axes"]], np.float32).reshape(-1, 2)
        w.add_joints(jd)
        w.set_joint_impulses(np.array([[_parse_f32_any(v) for v in row] for row in ck["joint_impulses"]],
                                      np.float32).reshape(-1, 5))
    st = w.body_states()
    for i, s in enumerate(ck["bodies"]):
        st[i]["p"] = _getvec(s["p"])
        st[i]["q"] = (_parse_f32_any(s["q"]["s"]), _parse_f32_any(s["q"]["c"]))
        st[i]["c"] = _getvec(s["c"])
        st[i]["a"] = _parse_f32_any(s["a"])
        st[i]["v"] = _getvec(s["v"])
        st[i]["w"] = _parse_f32_any(s["w"])
        st[i]["c0"] = _getvec(s["c0"])
        st[i]["a0"] = _parse_f32_any(s["a0"])
        st[i]["sleep_time"] = _parse_f32_any(s["sleep_time"])
        st[i]["awake"] = 1.0 if s["awake"] else 0.0
    w.set_body_states(st)
    w.set_fat_aabbs(np.array([[_parse_f32_any(v) for v in a] for a in ck["fat_aabbs"]], np.float32).reshape(-1, 4))
    # the contacts, then the destroyed touching contacts the next collide wakes bodies for
    contacts = ck["contacts"] + [dict(c, colour=-1) for c in ck.get("destroyed", [])]
    fab = np.array([[c["fixture_a"], c["fixture_b"]] for c in contacts], np.int32).reshape(-1, 2)
    man = np.zeros(len(contacts), MANIFOLD)
    for k, c in enumerate(contacts):
        m = c["manifold"]
        man[k]["type"] = MANIFOLD_TYPES.index(m["m_type"])
        man[k]["point_count"] = int(m["point_count"])
        man[k]["local_normal"] = _getvec(m["local_normal"])
        man[k]["local_point"] = _getvec(m["local_point"])
        for i in range(2):
            q = m["points"][i]
            man[k]["points"][i]["local_point"] = _getvec(q["local_point"])
            man[k]["points"][i]["normal_impulse"] = _parse_f32_any(q["normal_impulse"])
            man[k]["points"][i]["tangent_impulse"] = _parse_f32_any(q["tangent_impulse"])
            ident = q["id"]
            man[k]["points"][i]["id"] = (int(ident["index_a"]) | int(ident["index_b"]) << 8 |
                                         int(ident["type_a"]) << 16 | int(ident["type_b"]) << 24)
    col = np.array([c["colour"] for c in contacts], np.int32)
    w.set_contacts(fab, man, col)
    sv = ck["solver"]
    w.set_solver_state({"inv_dt0": _parse_f32_any(sv["inv_dt0"]), "n_colours": sv["n_colours"],
                        "scratch_colours": sv["scratch_colours"], "steps_since_scratch": sv["steps_since_scratch"]})
    return w
