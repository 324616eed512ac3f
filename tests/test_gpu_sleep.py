"""Sleep on the GPU (allow_sleep on): island sleep, waking, and the skipped work of sleeping bodies.

* One-step replay: from the GPU's state after its COLLIDE phase, the oracle's island solver
  (src/private/dynamics/b2_island_private.rs:129-328, with the sleep rule of :291-327) replays
  the GPU's constraints in the GPU's colour order, solving only islands with an awake body.
  Every step of a small pyramid settling to sleep must agree: awake flags and sleep times
  exactly, positions and velocities within 1e-5.
* Against the oracle run in the reference's sequential order: the pyramid falls asleep within a
  band of the same step and comes to rest at the same poses within a band.
* Waking: a body created asleep stays put until a contact touches it; a box dropped on the
  sleeping pyramid wakes it; switching sleeping off wakes every body; runs are deterministic.
"""
import numpy as np
import pytest

from test_gpu_solver import DT, _world, replay_inputs

pytestmark = pytest.mark.gpu


def _sleepy_world(scene):
    w = _world(scene)
    w.set_flags(warm_starting=True, allow_sleep=True, continuous=False, block_solve=True)
    return w


def _pyramid(rows=5):
    from b2g import scenes
    return scenes.pyramid(rows=rows)


def _dynamic(scene):
    return scene.body_defs["type"] == 2


def test_sleep_replay_matches_oracle(oracle):
    import b2g
    scene = _pyramid()
    dyn = _dynamic(scene)
    w = _sleepy_world(scene)
    auto = scene.body_defs["allow_sleep"].astype(np.int32)
    btype = scene.body_defs["type"]
    fell_asleep = 0
    for step in range(500):
        w.step_phase(b2g.PHASE_COLLIDE, DT, 8, 3)
        st0, props, touch, cb, params, man = replay_inputs(scene, w)
        w.step_phase(b2g.PHASE_SOLVE, DT, 8, 3)
        st1 = w.body_states()
        order_c, _ = w.solver_order()
        # the solved constraints: touching contacts of awake islands
        pos = {int(k): i for i, k in enumerate(touch)}
        sel = np.array(sorted(pos[int(k)] for k in order_c), np.int64)
        remap = {int(touch[i]): j for j, i in enumerate(sel)}
        order = np.array([remap[int(k)] for k in order_c], np.int32)
        flags = (st0["awake"] != 0).astype(np.int32) | (auto << 1)
        inv_dt0 = np.float32(1.0) / DT if step > 0 else np.float32(0.0)
        ref, _, _, _ = oracle.solver_replay(btype, props, st0, cb[sel], params[sel], man[sel], order, DT,
                                            np.float32(inv_dt0 * DT), w.gravity, 8, 3, True, True,
                                            body_flags=flags, allow_sleep=True)
        np.testing.assert_array_equal(st1["awake"][dyn], ref["awake"][dyn], err_msg="step %d: awake" % step)
        np.testing.assert_array_equal(st1["sleep_time"][dyn], ref["sleep_time"][dyn],
                                      err_msg="step %d: sleep time" % step)
        for f in ("c", "a", "v", "w"):
            err = np.abs(st1[f][dyn].astype(np.float64) - ref[f][dyn]).max()
            assert err <= 1e-5 * max(1.0, float(np.abs(ref[f][dyn]).max())), "step %d: %s differs by %g" % (step, f, err)
        fell_asleep += int(np.count_nonzero((st0["awake"][dyn] != 0) & (st1["awake"][dyn] == 0)))
        w.step_phase(b2g.PHASE_FIND, DT, 8, 3)
        if not np.any(st1["awake"][dyn]):
            break
    assert fell_asleep == int(dyn.sum()), "every box falls asleep within 500 steps"
    # asleep: nothing moves any more
    before = w.body_states()
    for _ in range(30):
        w.step(DT, 8, 3)
    after = w.body_states()
    for f in ("c", "a", "v", "w"):
        np.testing.assert_array_equal(before[f], after[f])


def _steps_to_sleep(step, awake_of, limit):
    for n in range(limit):
        step()
        if not np.any(awake_of()):
            return n + 1
    return None


def test_sleep_band_vs_reference_order(oracle):
    scene = _pyramid()
    dyn = _dynamic(scene)
    w = _sleepy_world(scene)
    ow = oracle.world()
    ow.set_flags(warm_starting=True, allow_sleep=True, block_solve=True)
    ow.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
    n_gpu = _steps_to_sleep(lambda: w.step(DT, 8, 3), lambda: w.body_states()["awake"][dyn], 800)
    n_ref = _steps_to_sleep(lambda: ow.step(DT, 8, 3), lambda: ow.bodies()["awake"][dyn], 800)
    assert n_gpu is not None and n_ref is not None
    # the Gauss-Seidel order differs (colours vs contact list), so the velocity history differs
    # in the last bits and the 0.5 s sleep timers start a few steps apart
    assert abs(n_gpu - n_ref) <= 30, (n_gpu, n_ref)
    g, r = w.body_states(), ow.bodies()
    assert np.abs(g["c"][dyn] - r["c"][dyn]).max() <= 0.02
    assert np.abs(g["a"][dyn] - r["a"][dyn]).max() <= 0.02


def test_body_created_asleep_waits_for_a_touch():
    import b2g
    from b2g import scenes
    scene = _pyramid(rows=2)
    bd = scene.body_defs.copy()
    # a box hanging asleep 3 m above the pyramid's top box, and one asleep far away
    extra = np.zeros(2, b2g.BODY_DEF)
    extra["type"] = 2
    extra["allow_sleep"] = 1
    extra["awake"] = 0
    extra["gravity_scale"] = 1.0
    top = bd["position"][np.argmax(bd["position"][:, 1])]
    extra["position"][0] = (top[0], top[1] + 3.0)
    extra["position"][1] = (30.0, 5.0)
    sd = np.concatenate([scene.shape_defs, scene.shape_defs[-1:].repeat(2)])
    fd = np.concatenate([scene.fixture_defs, scene.fixture_defs[-1:].repeat(2)])
    fb = np.concatenate([scene.fixture_body, [len(bd), len(bd) + 1]]).astype(np.int32)
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=True, continuous=False, block_solve=True)
    w.add(np.concatenate([bd, extra]), fb, sd, fd)
    st = w.body_states()
    assert st["awake"][-1] == 0 and st["awake"][-2] == 0
    for _ in range(120):
        w.step(DT, 8, 3)
    st = w.body_states()
    # nothing touches them: they hang where they were created
    np.testing.assert_array_equal(st["c"][-2:], np.array([extra["position"][0], extra["position"][1]], np.float32))
    assert st["awake"][-1] == 0 and st["awake"][-2] == 0


def test_dropped_box_wakes_sleeping_pyramid():
    import b2g
    scene = _pyramid()
    dyn = _dynamic(scene)
    w = _sleepy_world(scene)
    n = _steps_to_sleep(lambda: w.step(DT, 8, 3), lambda: w.body_states()["awake"][dyn], 800)
    assert n is not None
    rest = w.body_states()
    top = np.argmax(rest["c"][:, 1])
    box = np.zeros(1, b2g.BODY_DEF)
    box["type"] = 2
    box["allow_sleep"] = 1
    box["awake"] = 1
    box["gravity_scale"] = 1.0
    box["position"][0] = (rest["c"][top, 0], rest["c"][top, 1] + 2.0)
    w.add(box, np.zeros(1, np.int32), scene.shape_defs[-1:], scene.fixture_defs[-1:])
    woke = False
    for _ in range(120):
        w.step(DT, 8, 3)
        st = w.body_states()
        if st["awake"][top]:
            woke = True
            break
    assert woke, "the box landing on the top box wakes it"
    # and the pile, with the new box, goes back to sleep
    dyn2 = np.append(dyn, True)
    m = _steps_to_sleep(lambda: w.step(DT, 8, 3), lambda: w.body_states()["awake"][dyn2], 900)
    assert m is not None


def test_switching_sleep_off_wakes_every_body():
    scene = _pyramid(rows=3)
    dyn = _dynamic(scene)
    w = _sleepy_world(scene)
    n = _steps_to_sleep(lambda: w.step(DT, 8, 3), lambda: w.body_states()["awake"][dyn], 800)
    assert n is not None
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    st = w.body_states()
    assert np.all(st["awake"][dyn] == 1) and np.all(st["sleep_time"][dyn] == 0)
    assert np.all(st["awake"][~dyn] == 0)


def test_sleeping_run_is_deterministic():
    from b2g import scenes
    scene = scenes.pyramid(rows=8)
    runs = []
    for _ in range(2):
        w = _sleepy_world(scene)
        for _ in range(400):
            w.step(DT, 8, 3)
        runs.append(w.body_states())
    for f in ("c", "a", "v", "w", "awake", "sleep_time"):
        np.testing.assert_array_equal(runs[0][f], runs[1][f])
