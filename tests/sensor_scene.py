"""Scenes with sensor fixtures (B2fixtureDef::is_sensor), shared by the oracle pins and the GPU tests.

Mixed bodies (boxes, circles, random convex polygons at random angles) fall apart from each other
through three static sensor regions (a box, a circle, a hexagon) onto a static ground; each
falling body also carries a zero-density sensor circle of its own, wide enough that neighbouring
bodies' sensors overlap on the way down (sensor-sensor contacts). Without the sensor fixtures the
scene is the same bodies and solid fixtures, so each falling body is its own island (its
contacts with the ground only).
"""
import zlib

import numpy as np

import b2g

DT = np.float32(1.0 / 60.0)


def _fixture_defs(n):
    fd = np.zeros(n, b2g.FIXTURE_DEF)
    fd["friction"] = 0.4
    fd["restitution_threshold"] = 1.0
    fd["category"] = 1
    fd["mask"] = 0xFFFF
    fd["density"] = 1.0
    return fd


def sensor_scene(with_sensors=True, n=8, seed="sensor scene"):
    rng = np.random.default_rng(zlib.crc32(seed.encode()))
    nb = 4 + n  # ground, three sensor regions, the falling bodies
    bd = np.zeros(nb, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[1]["position"] = (-8.0, 4.0)
    bd[2]["position"] = (2.0, 5.0)
    bd[3]["position"] = (10.0, 4.0)
    bd[4:]["type"] = 2
    xs = np.linspace(-14.0, 14.0, n)  # 4 m apart: no two falling bodies meet, their sensors do
    bd[4:]["position"] = np.stack([xs, rng.uniform(8.0, 12.0, n)], 1)
    bd[4:]["angle"] = rng.uniform(-1.0, 1.0, n)
    fixture_body, shapes, fixtures = [], [], []

    def add(body, sd, sensor=False, density=1.0):
        fd = _fixture_defs(1)[0]
        fd["is_sensor"] = int(sensor)
        fd["density"] = density
        fixture_body.append(body)
        shapes.append(sd)
        fixtures.append(fd)

    sd = np.zeros((), b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_BOX
    sd["params"][:2] = (20.0, 0.5)
    add(0, sd.copy())
    if with_sensors:
        sd = np.zeros((), b2g.SHAPE_DEF)
        sd["type"] = b2g.SHAPE_BOX
        sd["params"][:2] = (3.0, 1.0)
        add(1, sd.copy(), sensor=True)
        sd = np.zeros((), b2g.SHAPE_DEF)
        sd["type"] = b2g.SHAPE_CIRCLE
        sd["radius"] = 1.5
        add(2, sd.copy(), sensor=True)
        sd = np.zeros((), b2g.SHAPE_DEF)
        sd["type"] = b2g.SHAPE_POLYGON
        ang = np.arange(6) * (np.pi / 3.0)
        sd["count"] = 6
        sd["params"][:12] = (1.6 * np.stack([np.cos(ang), np.sin(ang)], 1)).ravel()
        add(3, sd.copy(), sensor=True)
    for i in range(n):
        b = 4 + i
        sd = np.zeros((), b2g.SHAPE_DEF)
        k = i % 3
        if k == 0:
            sd["type"] = b2g.SHAPE_BOX
            sd["params"][:2] = rng.uniform(0.2, 0.45, 2)
        elif k == 1:
            sd["type"] = b2g.SHAPE_CIRCLE
            sd["radius"] = rng.uniform(0.2, 0.45)
        else:
            m = int(rng.integers(3, 9))
            a = np.sort(rng.uniform(0.0, 2.0 * np.pi, m))
            r = rng.uniform(0.25, 0.45, m)
            sd["type"] = b2g.SHAPE_POLYGON
            sd["count"] = m
            sd["params"][:2 * m] = np.stack([r * np.cos(a), r * np.sin(a)], 1).ravel()
        add(b, sd.copy())
        if with_sensors:
            sd = np.zeros((), b2g.SHAPE_DEF)
            sd["type"] = b2g.SHAPE_CIRCLE
            sd["radius"] = 2.1
            add(b, sd.copy(), sensor=True, density=0.0)
    return (bd, np.array(fixture_body, np.int32), np.array(shapes, b2g.SHAPE_DEF),
            np.array(fixtures, b2g.FIXTURE_DEF), (0.0, -10.0))


def thin_wall_scene(sensor):
    """A ball of radius 0.1 at 120 m/s (2 m per step) towards a static wall 0.1 m thick, no
    gravity: with continuous physics a solid wall stops it; a sensor wall gets no time of impact."""
    bd = np.zeros(2, b2g.BODY_DEF)
    bd["gravity_scale"] = 0.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[0]["position"] = (5.0, 0.0)
    bd[1]["type"] = 2
    bd[1]["linear_velocity"] = (120.0, 0.0)
    sd = np.zeros(2, b2g.SHAPE_DEF)
    sd[0]["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (0.05, 2.0)
    sd[1]["type"] = b2g.SHAPE_CIRCLE
    sd[1]["radius"] = 0.1
    fd = _fixture_defs(2)
    fd[0]["is_sensor"] = int(sensor)
    return bd, np.arange(2, dtype=np.int32), sd, fd, (0.0, 0.0)


def _core_points(rec):
    if rec["type"] == 0:
        return np.array([rec["p"]], np.float64)
    if rec["type"] == 1:
        return np.array([rec["v1"], rec["v2"]], np.float64)
    return np.array(rec["vertices"][:2 * rec["count"]], np.float64).reshape(-1, 2)


def _place(pts, p, q):
    s, c = float(q[0]), float(q[1])
    return np.stack([c * pts[:, 0] - s * pts[:, 1] + p[0], s * pts[:, 0] + c * pts[:, 1] + p[1]], 1)


def _seg_point(a, b, p):
    d = b - a
    L = d @ d
    t = 0.0 if L == 0 else np.clip((p - a) @ d / L, 0.0, 1.0)
    return np.linalg.norm(a + t * d - p)


def _edges(P):
    if len(P) == 1:
        return [(P[0], P[0])]
    return [(P[i], P[(i + 1) % len(P)]) for i in range(len(P))]


def _inside(P, q):
    if len(P) < 3:
        return False
    for a, b in _edges(P):
        if (b[0] - a[0]) * (q[1] - a[1]) - (b[1] - a[1]) * (q[0] - a[0]) < 0:
            return False
    return True


def _cross(a, b, c, d):
    def o(p, q, r):
        return np.sign((q[0] - p[0]) * (r[1] - p[1]) - (q[1] - p[1]) * (r[0] - p[0]))
    return o(a, b, c) * o(a, b, d) < 0 and o(c, d, a) * o(c, d, b) < 0


def hull_distance(PA, PB):
    """Distance between two convex point sets (a polygon counter-clockwise, or one point), by brute
    force: 0 when one holds a point of the other or their edges cross, else the least
    point-to-edge distance."""
    if any(_inside(PA, q) for q in PB) or any(_inside(PB, q) for q in PA):
        return 0.0
    if any(_cross(a, b, c, d) for a, b in _edges(PA) for c, d in _edges(PB)):
        return 0.0
    best = np.inf
    for a, b in _edges(PA):
        for q in PB:
            best = min(best, _seg_point(a, b, q))
    for c, d in _edges(PB):
        for q in PA:
            best = min(best, _seg_point(c, d, q))
    return best


def sensor_expectations(shape_recs, fixture_body, states, pairs, margin=1e-4):
    """For each fixture pair: True / False when the placed shapes overlap / are apart by more than
    margin beyond the summed radii (brute-force distance), None when the pair is within the margin
    of touching."""
    out = []
    for fa, fb in pairs:
        ra, rb = shape_recs[fa], shape_recs[fb]
        sa, sb = states[fixture_body[fa]], states[fixture_body[fb]]
        d = hull_distance(_place(_core_points(ra), sa["p"], sa["q"]), _place(_core_points(rb), sb["p"], sb["q"]))
        gap = d - (float(ra["radius"]) + float(rb["radius"]))
        out.append(None if abs(gap) < margin else bool(gap < 0.0))
    return out
