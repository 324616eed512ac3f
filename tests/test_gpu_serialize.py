"""The serde world format on the GPU: reading a world, writing a live one, checkpoints.

* Reading: a world written as serde JSON and read back is the reference reader's world (bodies
  and fixtures created in file order, polygons as stored). It must step bit-identically with the
  same bodies and fixtures handed over directly in that order, with the polygons given as their
  raw points (the hull the stored polygons came from).
* Writing a live world: the body definitions hold the current transforms, velocities and awake
  flags (src/serialize/serialize_b2_body.rs:37-56).
* Checkpoint: a world written with the checkpoint extension after some steps and read back
  continues bit-identically with the written one (body states, contacts and impulses after more
  steps), with sleeping off and on; written again right after reading, its text is the same.
"""
import numpy as np
import pytest

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
    w2 = b2g.World()
    w2.set_flags(warm_starting=True, allow_sleep=True, continuous=False, block_solve=True)
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
