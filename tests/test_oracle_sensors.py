"""Pins of the oracle's sensors (B2fixtureDef::is_sensor).

A sensor contact gets no manifold and wakes nobody; its touching flag is the shapes' overlap,
b2_test_overlap: the GJK distance with radii below ten epsilons
(src/private/dynamics/b2_contact.rs:148-162, src/private/collision/b2_collision.rs:216-245).
Sensor contacts join no island and get no time of impact (src/private/dynamics/b2_world.rs:
444-449, 590-593, 784-789). The restatement is pinned by what the geometry and mechanics fix:
  * every sensor contact's touching flag equals the brute-force overlap of the shapes placed where
    the step's collide sees them (the least point-to-edge distance of the hulls, less both radii;
    pairs within 1e-4 of touching are not judged) at every step of mixed bodies falling through
    static sensor regions with sensors of their own; both transitions happen; a sensor contact
    never has manifold points;
  * sensors exert nothing: the same world with its sensor fixtures (zero density on the falling
    bodies) steps bit for bit as the world without them, with continuous physics on (a sensor
    with a time of impact would stop a body at the sensor region);
  * a fast ball passes through a thin static sensor wall with continuous physics on (no TOI
    event) and stops at the same wall when it is solid.
A sensor flag that ignores the radii, a transform of the wrong body, a sensor in an island or a
TOI candidate fails one of these.
"""
import numpy as np

from sensor_scene import DT, sensor_expectations, sensor_scene, thin_wall_scene


def _world(oracle, scene, continuous=True):
    bd, fb, sd, fd, g = scene
    w = oracle.world(g)
    w.set_flags(warm_starting=True, allow_sleep=False, block_solve=True, continuous=continuous)
    w.add(bd, fb, sd, fd)
    return w


def test_sensor_flags_match_brute_force_overlap(oracle):
    scene = sensor_scene()
    bd, fb, sd, fd, g = scene
    recs, status = oracle.make_shapes(sd)
    assert (status == 0).all()
    w = _world(oracle, scene)
    sensor = fd["is_sensor"] != 0
    on = off = judged = 0
    was = {}
    before = set()
    for step in range(150):
        st = w.bodies()  # the transforms the step's collide sees
        w.step(DT, 8, 3)
        fab, flags, man, _ = w.contacts()
        is_s = sensor[fab[:, 0]] | sensor[fab[:, 1]]
        assert (man["point_count"][is_s] == 0).all()
        exp = sensor_expectations(recs, fb, st, fab[is_s])
        for (a, b), t, e in zip(fab[is_s], flags[is_s], exp):
            key = (int(a), int(b))
            # a contact the step's pair search created has not been through a collide yet
            if e is not None and key in before:
                assert bool(t) == e, (step, a, b, t)
                judged += 1
            if was.get(key, False) != bool(t):
                on += bool(t)
                off += not t
            was[key] = bool(t)
        before = {(int(a), int(b)) for a, b in fab}
    assert judged > 1000 and on >= 20 and off >= 10, (judged, on, off)


def test_sensors_exert_nothing(oracle):
    a = _world(oracle, sensor_scene(True))
    scene = sensor_scene(False)
    b = _world(oracle, scene)
    fixture_body = scene[1]
    for step in range(150):
        a.step(DT, 8, 3)
        b.step(DT, 8, 3)
        sa, sb = a.bodies(), b.bodies()
        for f in ("c", "a", "v", "w"):
            assert sa[f].tobytes() == sb[f].tobytes(), (step, f)
        # the scene's premise: no two falling bodies touch (each is its own island)
        fab, _, man, _ = b.contacts()
        both = (fixture_body[fab[:, 0]] >= 4) & (fixture_body[fab[:, 1]] >= 4)
        assert not (both & (man["point_count"] > 0)).any(), step
    assert a.contact_count > b.contact_count
    st = a.bodies()
    assert (st["c"][4:, 1] < 1.5).all()  # every falling body came through the regions to the ground


def test_fast_ball_and_thin_sensor_wall(oracle):
    for sensor in (True, False):
        w = _world(oracle, thin_wall_scene(sensor))
        events = 0
        for _ in range(6):
            w.step(DT, 8, 3)
            events += w.toi_events
        x = float(w.bodies()[1]["c"][0])
        if sensor:
            assert events == 0 and x > 5.05, (events, x)
        else:
            assert events >= 1 and x < 5.0 - 0.05 - 0.1 + 0.02, (events, x)
