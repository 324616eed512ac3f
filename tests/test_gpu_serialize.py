"""The serde world format on the GPU: reading a world, writing a live one, checkpoints.

* Reading: a world written as serde JSON and read back is the reference reader's world (bodies
  and fixtures created in file order, polygons as stored). It must step bit-identically with the
  same bodies and fixtures handed over directly in that order, with the polygons given as their
  raw points (the hull the stored polygons came from).
* Writing a live world: the body definitions hold the current transforms, velocities and awake
  flags (src/serialize/serialize_b2_body.rs:37-56).
* Checkpoint: a world written with the checkpoint extension after some steps and read back
  continues bit-identically with the written one (body states, contacts and impulses after more
  steps), with sleeping off and on, and with joints (hanging chains; prismatic and weld joints;
  friction, motor, wheel and pulley joints); written again right after reading, its text is the
  same. A world with gear or mouse joints has no checkpoint.
* Joints read from a file (indices as the writer meant them) step bit-identically with the same
  joints created directly in file order, every kind the format writes (gears naming their joints
  by list position).
* Against the oracle: scenes written as serde JSON (the reference's hello_world box, a fast ball
  at a thin wall) and read back step on the GPU like the oracle stepping the same file's world with
  B2world::new's flags (warm starting, sleeping and continuous physics on), each step within 1e-5
  (1e-4 with TOI events, the tolerance of the continuous-collision parity tests), awake flags and
  sleep times exactly.
"""
import numpy as np
import pytest

import b2g
from b2g import serialize as ser

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def _scene(kind):
    from b2g import scenes
    if kind == "pyramid":
        return scenes.pyramid(rows=10)
    return scenes.large_world(n=1024, seed=2)


def _world(scene, sleep):
    import b2g
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=sleep, continuous=False, block_solve=True)
    w.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
    return w


def _states_equal(a, b, what):
    for f in ("p", "q", "c", "a", "v", "w", "c0", "a0", "sleep_time", "awake"):
        np.testing.assert_array_equal(a[f], b[f], err_msg="%s: %s" % (what, f))


@pytest.mark.parametrize("kind", ["pyramid", "large"])
def test_read_world_steps_like_direct_creation(kind):
    import b2g
    scene = _scene(kind)
    w0 = _world(scene, sleep=True)
    text = ser.world_to_json(w0)
    w1 = ser.world_from_json(text)
    # the same bodies and fixtures in file order (newest first), handed over directly, with the
    # polygons as raw points
    nb = len(scene.body_defs)
    order = np.arange(nb)[::-1]
    remap = np.empty(nb, np.int64)
    remap[order] = np.arange(nb)
    fb = np.asarray(scene.fixture_body)
    per_body = [np.flatnonzero(fb == b) for b in range(nb)]
    forder = np.concatenate([per_body[b][::-1] for b in order])
    w2 = b2g.World()  # the reader's world: B2world::new's flags, the library's defaults
    w2.add(scene.body_defs[order], remap[fb[forder]].astype(np.int32), scene.shape_defs[forder],
           scene.fixture_defs[forder])
    assert w1.shapes().tobytes() == w2.shapes().tobytes()
    np.testing.assert_array_equal(w1.mass(), w2.mass())
    for _ in range(90):
        w1.step(DT, 8, 3)
        w2.step(DT, 8, 3)
    _states_equal(w1.body_states(), w2.body_states(), "read vs direct")


def test_live_world_definitions_hold_current_state():
    scene = _scene("pyramid")
    w = _world(scene, sleep=False)
    for _ in range(30):
        w.step(DT, 8, 3)
    st = w.body_states()
    doc = ser.read_world(ser.world_to_json(w))
    back = np.arange(len(st))[::-1]
    bd = doc["body_defs"][back]
    np.testing.assert_array_equal(bd["position"], st["p"])
    np.testing.assert_array_equal(bd["angle"], st["a"])
    np.testing.assert_array_equal(bd["linear_velocity"], st["v"])
    np.testing.assert_array_equal(bd["angular_velocity"], st["w"])
    dyn = scene.body_defs["type"] == 2
    np.testing.assert_array_equal(bd["awake"][dyn], (st["awake"][dyn] != 0).astype(np.int32))


@pytest.mark.parametrize("kind,sleep", [("large", False), ("pyramid", True)])
def test_checkpoint_continues_bit_identically(kind, sleep):
    scene = _scene(kind)
    w1 = _world(scene, sleep)
    for _ in range(45):
        w1.step(DT, 8, 3)
    text = ser.world_to_json(w1, checkpoint=True)
    w2 = ser.world_from_json(text)
    assert ser.world_to_json(w2, checkpoint=True) == text
    _states_equal(w1.body_states(), w2.body_states(), "restored")
    for step in range(60):
        w1.step(DT, 8, 3)
        w2.step(DT, 8, 3)
    _states_equal(w1.body_states(), w2.body_states(), "after 60 more steps")
    fa, ta, ma = w1.contacts()
    fb_, tb, mb = w2.contacts()
    np.testing.assert_array_equal(fa, fb_)
    np.testing.assert_array_equal(ta, tb)
    assert ma.tobytes() == mb.tobytes()


def _joint_world(kind):
    import b2g
    from b2g import scenes
    if kind == "chains":
        scene, jd = scenes.chains(32, 8, spacing=1.5, rows=4)
        return scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs, jd
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from joint_scenes import sliders, welds
    a = sliders([-0.5, 0.3, 1.2])
    b = welds([0.0, 5.0])
    nb = len(a[0])
    bd = np.concatenate([a[0], b[0]])
    fb = np.concatenate([a[1], b[1] + nb]).astype(np.int32)
    jd = np.concatenate([a[4], b[4]])
    jd["body_a"][len(a[4]):] += nb
    jd["body_b"][len(a[4]):] += nb
    jd[0]["enable_limit"], jd[0]["lower_angle"], jd[0]["upper_angle"] = 1, -0.5, 0.4
    jd[1]["enable_motor"], jd[1]["motor_speed"], jd[1]["max_motor_torque"] = 1, 0.7, 50.0
    return bd, fb, np.concatenate([a[2], b[2]]), np.concatenate([a[3], b[3]]), jd


def _merge(parts, dx=40.0):
    """Joint scenes side by side, part k shifted by k * dx along x (world-frame joint data moved
    with it; a gear's joints renumbered)."""
    bds, fbs, sds, fds, jds = [], [], [], [], []
    nb = nj = 0
    for k, (bd, fb, sd, fd, jd) in enumerate(parts):
        bd, jd = bd.copy(), jd.copy()
        bd["position"][:, 0] += k * dx
        for key in ("ground_anchor_a", "ground_anchor_b", "target"):
            jd[key][:, 0] += k * dx
        gear = jd["type"] == b2g.JOINT_GEAR
        jd["joint1"][gear] += nj
        jd["joint2"][gear] += nj
        jd["body_a"] += nb
        jd["body_b"] += nb
        bds.append(bd)
        fbs.append(np.asarray(fb) + nb)
        sds.append(sd)
        fds.append(fd)
        jds.append(jd)
        nb += len(bd)
        nj += len(jd)
    return (np.concatenate(bds), np.concatenate(fbs).astype(np.int32), np.concatenate(sds), np.concatenate(fds),
            np.concatenate(jds))


def _more_kinds(gears=False):
    """Friction, motor, wheel (motor and limit) and pulley joints; with gears=True also geared
    wheels and a geared rack."""
    import joint_scenes as J
    parts = [J.friction_boxes([(3.0, 0.0), (2.0, -1.5)], [2.0, -1.0], [1.0, 2.0], [0.1, 0.3]),
             J.motor_boxes([(1.0, 2.0), (-3.0, 0.5)], [0.3, -0.2], [(0.0, 0.0), (-1.0, 1.0)], [0.0, 0.5],
                           max_force=[1e6, 2.0], max_torque=[1e6, 0.5])[:5],
             J.wheels_on_springs(n=2, stiffness=(40.0, 80.0), motor_speed=5.0, torque=0.4),
             J.wheels_on_springs(n=2, stiffness=(40.0, 80.0), limit=(-0.15, 0.2)),
             J.atwood(3.0, 1.0, 2.0)]
    if gears:
        parts += [J.geared_wheels(2.0, 3.0), J.geared_wheels(1.0, 3.0, prismatic_b=True)]
    return _merge(parts)


@pytest.mark.parametrize("kind", ["chains", "mixed", "more"])
def test_checkpoint_with_joints(kind):
    import b2g
    bd, fb, sd, fd, jd = _more_kinds() if kind == "more" else _joint_world(kind)
    w1 = b2g.World((0.0, -10.0))
    w1.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    w1.add(bd, fb, sd, fd)
    w1.add_joints(jd)
    for _ in range(45):
        w1.step(DT, 8, 3)
    text = ser.world_to_json(w1, checkpoint=True)
    w2 = ser.world_from_json(text)
    assert ser.world_to_json(w2, checkpoint=True) == text
    for _ in range(60):
        w1.step(DT, 8, 3)
        w2.step(DT, 8, 3)
    _states_equal(w1.body_states(), w2.body_states(), "after 60 more steps")
    np.testing.assert_array_equal(w1.joint_impulses(), w2.joint_impulses())


def test_read_joints_step_like_direct_creation():
    import b2g
    bd, fb, sd, fd, jd = _joint_world("chains")
    w0 = b2g.World()
    w0.add(bd, fb, sd, fd)
    w0.add_joints(jd)
    text = ser.world_to_json(w0)
    w1 = ser.world_from_json(text, joint_bodies="file")
    doc = ser.read_world(text, joint_bodies="file")
    w2 = b2g.World()  # the reader's world: B2world::new's flags, the library's defaults
    w2.add(doc["body_defs"], doc["fixture_body"], doc["shape_defs"], doc["fixture_defs"])
    w2.add_joints(doc["joint_defs"])
    # the file's joints are the world's, newest first, on the same bodies
    nb = len(bd)
    back = jd[::-1]
    np.testing.assert_array_equal(doc["joint_defs"]["body_a"], nb - 1 - back["body_a"])
    np.testing.assert_array_equal(doc["joint_defs"]["body_b"], nb - 1 - back["body_b"])
    for _ in range(60):
        w1.step(DT, 8, 3)
        w2.step(DT, 8, 3)
    _states_equal(w1.body_states(), w2.body_states(), "read vs direct")


def test_every_kind_read_steps_like_direct_creation():
    """Wheel, friction, motor, pulley and gear joints read from a file (the writer's indices) step
    bit-identically with the written world's bodies, fixtures and joints created directly in file
    order (bodies and joints newest first, gears after the others naming their joints in that
    order)."""
    bd, fb, sd, fd, jd = _more_kinds(gears=True)
    w0 = b2g.World((0.0, -10.0))
    w0.add(bd, fb, sd, fd)
    w0.add_joints(jd)
    text = ser.world_to_json(w0)
    w1 = ser.world_from_json(text, joint_bodies="file")
    nb = len(bd)
    border = np.arange(nb)[::-1]
    remap = np.empty(nb, np.int64)
    remap[border] = np.arange(nb)
    per_body = [np.flatnonzero(fb == b) for b in range(nb)]
    forder = np.concatenate([per_body[b][::-1] for b in border])
    gear = jd["type"] == b2g.JOINT_GEAR
    jorder = np.concatenate([np.flatnonzero(~gear)[::-1], np.flatnonzero(gear)[::-1]])
    jpos = np.empty(len(jd), np.int64)
    jpos[jorder] = np.arange(len(jd))
    j2 = jd[jorder].copy()
    j2["body_a"], j2["body_b"] = remap[j2["body_a"]], remap[j2["body_b"]]
    g2 = j2["type"] == b2g.JOINT_GEAR
    j2["joint1"][g2], j2["joint2"][g2] = jpos[j2["joint1"][g2]], jpos[j2["joint2"][g2]]
    w2 = b2g.World((0.0, -10.0))
    w2.add(bd[border], remap[fb[forder]].astype(np.int32), sd[forder], fd[forder])
    w2.add_joints(j2)
    np.testing.assert_array_equal(w1.mass(), w2.mass())
    for _ in range(90):
        w1.step(DT, 8, 3)
        w2.step(DT, 8, 3)
    _states_equal(w1.body_states(), w2.body_states(), "read vs direct")
    np.testing.assert_array_equal(w1.joint_impulses(), w2.joint_impulses())


@pytest.mark.parametrize("what", ["gear", "mouse"])
def test_checkpoint_refuses_gear_and_mouse_joints(what):
    """A gear's constant comes from the poses at its creation and the format writes no mouse joint:
    a checkpoint of a world holding either is refused (the plain file is still written, a mouse
    joint left out as the reference's writer leaves it out)."""
    import joint_scenes as J
    sc = J.geared_wheels(2.0, 3.0) if what == "gear" else J.mouse_boxes([(0.0, 0.0)], [(1.0, 1.0)], [1.0], [0.5])
    w = b2g.World()
    w.add(*sc[:4])
    w.add_joints(sc[4])
    w.step(DT, 8, 3)
    with pytest.raises(b2g.B2gError):
        ser.world_to_json(w, checkpoint=True)
    doc = ser.read_world(ser.world_to_json(w))
    assert len(doc["joint_defs"]) == (3 if what == "gear" else 0)


def _oracle_from_doc(oracle, doc):
    ow = oracle.world(doc["gravity"])
    ow.set_flags(warm_starting=True, allow_sleep=True, block_solve=True, continuous=True)
    ow.add(doc["body_defs"], doc["fixture_body"], doc["shape_defs"], doc["fixture_defs"])
    return ow


@pytest.mark.parametrize("kind", ["hello_world", "wall"])
def test_read_world_steps_like_the_oracle(oracle, kind):
    import b2g
    from b2g import scenes
    from test_gpu_toi import wall_scene
    if kind == "hello_world":
        sc = scenes.hello_world()
        bd, fb, sd, fd, g = sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs, (0.0, -10.0)
    else:
        bd, fb, sd, fd, g = wall_scene()
    src = b2g.World(g)
    src.add(bd, fb, sd, fd)
    text = ser.world_to_json(src)
    w = ser.world_from_json(text)
    assert w.flags == {"warm_starting": True, "allow_sleep": True, "continuous": True, "block_solve": True}
    ow = _oracle_from_doc(oracle, ser.read_world(text))
    tol = 1e-5 if kind == "hello_world" else 1e-4
    for k in range(90):
        w.step(DT, 8, 3)
        ow.step(DT, 8, 3)
        gs, os_ = w.body_states(), ow.bodies()
        for f in ("c", "a", "v", "w"):
            np.testing.assert_allclose(gs[f], os_[f], rtol=0, atol=tol, err_msg="step %d %s" % (k, f))
        np.testing.assert_array_equal(gs["awake"], os_["awake"], err_msg="step %d" % k)
        np.testing.assert_array_equal(gs["sleep_time"], os_["sleep_time"], err_msg="step %d" % k)
