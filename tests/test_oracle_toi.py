"""Pins of the oracle's continuous collision (GJK distance, time of impact, the TOI event loop).

The reference has no tests of its own for these; the restatement
(src/private/collision/b2_distance.rs:404, src/private/collision/b2_time_of_impact.rs:232,
src/private/dynamics/b2_world.rs:534-901) is pinned by what geometry and mechanics fix:
  * GJK: the distance between random convex polygons, circles' centres and segments at random
    poses equals the brute-force distance of disjoint convex sets (the least vertex-to-edge
    distance), 0 for overlapping sets, with the witness points on the shapes and that far apart;
    with radii the distance shrinks by the two radii;
  * time of impact: two circles closing head-on and a box dropping flat onto a static box touch at
    the analytic time when the core distance reaches the target (max(slop, r_a + r_b - 3 slop))
    within the tolerance window (a quarter slop over the closing speed); for random moving and
    turning polygon pairs the distance at the reported time is within the tolerance of the target
    and no sampled earlier time is closer than target - tolerance (no tunnelling); pairs that never
    come close are separated;
  * the event loop: a fast small ball shot at a thin static wall passes through it with continuous
    physics off and stops in front of it with continuous physics on (TOI events); boxes falling
    at 60 m/s onto a thin static ground mostly fall through it without continuous physics and all
    come to rest on it (within the slops) with it.
A slip such as a wrong barycentric sign, a swapped support direction or a dropped sweep advance
fails one of these.
"""
import numpy as np
import pytest

import b2g

SLOP = 0.005


def _shape_defs(kind, rng, n):
    sd = np.zeros(n, b2g.SHAPE_DEF)
    for i in range(n):
        k = kind if kind != "mixed" else ("polygon", "circle", "edge")[i % 3]
        if k == "polygon":
            m = rng.integers(3, 9)
            ang = np.sort(rng.uniform(0, 2 * np.pi, m))
            r = rng.uniform(0.3, 1.2, m)
            pts = np.stack([r * np.cos(ang), r * np.sin(ang)], 1) + rng.uniform(-0.3, 0.3, 2)
            sd[i]["type"] = b2g.SHAPE_POLYGON
            sd[i]["count"] = m
            sd[i]["params"][:2 * m] = pts.ravel()
        elif k == "circle":
            sd[i]["type"] = b2g.SHAPE_CIRCLE
            sd[i]["radius"] = rng.uniform(0.1, 0.8)
            sd[i]["params"][:2] = rng.uniform(-0.3, 0.3, 2)
        else:
            sd[i]["type"] = b2g.SHAPE_EDGE
            sd[i]["params"][:4] = rng.uniform(-1.0, 1.0, 4)
    return sd


def _core_points(rec):
    if rec["type"] == 0:
        return np.array([rec["p"]], np.float64)
    if rec["type"] == 1:
        return np.array([rec["v1"], rec["v2"]], np.float64)
    return np.array(rec["vertices"][:2 * rec["count"]], np.float64).reshape(-1, 2)


def _place(pts, xf):
    s, c = xf[2], xf[3]
    return np.stack([c * pts[:, 0] - s * pts[:, 1] + xf[0], s * pts[:, 0] + c * pts[:, 1] + xf[1]], 1)


def _seg_point(a, b, p):
    d = b - a
    L = d @ d
    t = 0.0 if L == 0 else np.clip((p - a) @ d / L, 0.0, 1.0)
    return np.linalg.norm(a + t * d - p)


def _edges(P):
    if len(P) == 1:
        return [(P[0], P[0])]
    if len(P) == 2:
        return [(P[0], P[1])]
    return [(P[i], P[(i + 1) % len(P)]) for i in range(len(P))]


def _inside(P, q):
    """q inside (or on) the counter-clockwise convex polygon P (three or more points)."""
    if len(P) < 3:
        return False
    for a, b in _edges(P):
        if (b[0] - a[0]) * (q[1] - a[1]) - (b[1] - a[1]) * (q[0] - a[0]) < 0:
            return False
    return True


def _cross_segments(a, b, c, d):
    def o(p, q, r):
        return np.sign((q[0] - p[0]) * (r[1] - p[1]) - (q[1] - p[1]) * (r[0] - p[0]))
    return o(a, b, c) * o(a, b, d) < 0 and o(c, d, a) * o(c, d, b) < 0


def brute_distance(PA, PB):
    """Distance between the convex hulls of two point sets given in order (a polygon counter-clockwise)."""
    if any(_inside(PA, q) for q in PB) or any(_inside(PB, q) for q in PA):
        return 0.0
    for a, b in _edges(PA):
        for c, d in _edges(PB):
            if _cross_segments(a, b, c, d):
                return 0.0
    best = np.inf
    for a, b in _edges(PA):
        for q in PB:
            best = min(best, _seg_point(a, b, q))
    for c, d in _edges(PB):
        for q in PA:
            best = min(best, _seg_point(c, d, q))
    return best


def _xf(rng, n, spread):
    ang = rng.uniform(-np.pi, np.pi, n)
    return np.stack([rng.uniform(-spread, spread, n), rng.uniform(-spread, spread, n), np.sin(ang), np.cos(ang)], 1)


@pytest.mark.parametrize("kind", ["polygon", "mixed"])
def test_gjk_distance_matches_brute_force(oracle64, kind):
    rng = np.random.default_rng(11)
    n = 300
    sa, st = oracle64.make_shapes(_shape_defs(kind, rng, n))
    sb, st2 = oracle64.make_shapes(_shape_defs("polygon" if kind == "polygon" else "mixed", rng, n)[::-1].copy())
    assert (st == 0).all() and (st2 == 0).all()
    xa = _xf(rng, n, 0.2)
    xb = _xf(rng, n, 2.5)
    out = oracle64.distance(sa, xa, sb, xb)
    n_overlap = 0
    for i in range(n):
        PA = _place(_core_points(sa[i]), xa[i])
        PB = _place(_core_points(sb[i]), xb[i])
        ref = brute_distance(PA, PB)
        d = out[i, 0]
        if ref == 0.0:
            n_overlap += 1
            assert d < 1e-9, (i, d)
            continue
        assert abs(d - ref) < 1e-9 * max(1.0, ref), (i, d, ref)
        pa, pb = out[i, 1:3], out[i, 3:5]
        assert abs(np.linalg.norm(pb - pa) - d) < 1e-9
        # the witness points lie on the shapes: their distances to the other shape add up
        assert brute_distance(PA, pa[None]) < 1e-9 and brute_distance(PB, pb[None]) < 1e-9
    assert 5 <= n_overlap < n - 50  # both cases are exercised
    # with radii: the gap shrinks by the radii, overlaps report 0
    outr = oracle64.distance(sa, xa, sb, xb, use_radii=True)
    gap = out[:, 0] - sa["radius"] - sb["radius"]
    sep = (gap > 0) & (out[:, 0] > 1e-9)
    np.testing.assert_allclose(outr[sep, 0], gap[sep], atol=1e-9)
    assert (outr[~sep, 0] == 0).all()


def _sweep(lc, c0, c, a0, a):
    return np.array([lc[0], lc[1], c0[0], c0[1], c[0], c[1], a0, a, 0.0])


def test_toi_circles_head_on(oracle64):
    rng = np.random.default_rng(3)
    sd = np.zeros(2, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_CIRCLE
    for trial in range(50):
        ra, rb = rng.uniform(0.05, 0.5, 2)
        sd[0]["radius"], sd[1]["radius"] = ra, rb
        shapes, _ = oracle64.make_shapes(sd)
        x0 = ra + rb + rng.uniform(0.1, 2.0)
        x1 = rng.uniform(-1.0, ra + rb - 0.05)
        wa = _sweep((0, 0), (0, 0), (0, 0), 0, 0)
        wb = _sweep((0, 0), (x0, 0), (x1, 0), 0, 0)
        state, t = oracle64.time_of_impact(shapes[:1], wa[None], shapes[1:], wb[None])
        target = max(SLOP, ra + rb - 3 * SLOP)
        tol = 0.25 * SLOP
        t_star = (x0 - target) / (x0 - x1)
        assert state[0] == 3, (trial, state)
        assert abs(t[0] - t_star) <= tol / (x0 - x1) + 1e-12, (trial, t[0], t_star)


def test_toi_box_drops_flat(oracle64):
    sd = np.zeros(2, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (5.0, 0.5)
    sd[1]["params"][:2] = (0.5, 0.5)
    shapes, _ = oracle64.make_shapes(sd)
    r = shapes["radius"][0] + shapes["radius"][1]
    target = max(SLOP, r - 3 * SLOP)
    tol = 0.25 * SLOP
    for y0, y1 in [(2.0, 0.5), (1.3, 0.9), (1.05, -3.0), (3.0, 1.02)]:
        wa = _sweep((0, 0), (0, 0), (0, 0), 0, 0)
        wb = _sweep((0, 0), (0.3, y0), (0.3, y1), 0, 0)
        state, t = oracle64.time_of_impact(shapes[:1], wa[None], shapes[1:], wb[None])
        gap0, gap1 = y0 - 1.0, y1 - 1.0
        if gap1 > target + tol:
            assert state[0] == 4 and t[0] == 1.0
            continue
        t_star = (gap0 - target) / (gap0 - gap1)
        assert state[0] == 3 and abs(t[0] - t_star) <= tol / (gap0 - gap1) + 1e-12, (y0, y1, t[0], t_star)


def _sweep_xf(w, t):
    c = (1 - t) * w[2:4] + t * w[4:6]
    a = (1 - t) * w[6] + t * w[7]
    s, co = np.sin(a), np.cos(a)
    p = c - np.array([co * w[0] - s * w[1], s * w[0] + co * w[1]])
    return np.array([p[0], p[1], s, co])


def test_toi_random_motion_no_tunnelling(oracle64):
    rng = np.random.default_rng(5)
    n = 120
    sa, _ = oracle64.make_shapes(_shape_defs("polygon", rng, n))
    sb, _ = oracle64.make_shapes(_shape_defs("mixed", rng, n))
    counts = {3: 0, 4: 0}
    for i in range(n):
        wa = _sweep((0.1, 0.0), (0, 0), (0, 0), 0, rng.uniform(-0.3, 0.3))
        start = rng.uniform(-1, 1, 2)
        start = 3.5 * start / np.linalg.norm(start)
        end = -start * rng.uniform(0.0, 1.2) + rng.uniform(-0.5, 0.5, 2)
        wb = _sweep((0.0, 0.1), start, end, rng.uniform(-1, 1), rng.uniform(-1.5, 1.5))
        state, t = oracle64.time_of_impact(sa[i:i + 1], wa[None], sb[i:i + 1], wb[None])
        target = max(SLOP, sa[i]["radius"] + sb[i]["radius"] - 3 * SLOP)
        tol = 0.25 * SLOP
        PA0, PB0 = _core_points(sa[i]), _core_points(sb[i])

        def dist(tt):
            return brute_distance(_place(PA0, _sweep_xf(wa, tt)), _place(PB0, _sweep_xf(wb, tt)))
        assert state[0] in (3, 4), (i, state[0])
        counts[state[0]] += 1
        if state[0] == 3:
            assert abs(dist(t[0]) - target) < tol + 1e-9, (i, t[0], dist(t[0]), target)
            for tt in np.linspace(0.0, t[0], 60):
                assert dist(tt) > target - tol - 1e-9, (i, tt)
        else:
            assert t[0] == 1.0
    assert counts[3] > 20 and counts[4] > 5


def _wall_scene(speed):
    bd = np.zeros(2, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[0]["type"] = 0
    bd[0]["position"] = (5.0, 0.0)
    bd[1]["type"] = 2
    bd[1]["position"] = (0.0, 0.0)
    bd[1]["linear_velocity"] = (speed, 0.0)
    bd[1]["gravity_scale"] = 0.0
    sd = np.zeros(2, b2g.SHAPE_DEF)
    sd[0]["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (0.05, 2.0)
    sd[1]["type"] = b2g.SHAPE_CIRCLE
    sd[1]["radius"] = 0.1
    fd = np.zeros(2, b2g.FIXTURE_DEF)
    fd["friction"] = 0.2
    fd["restitution_threshold"] = 1.0
    fd["category"] = 1
    fd["mask"] = 0xFFFF
    fd["density"] = 1.0
    return bd, np.arange(2, dtype=np.int32), sd, fd


@pytest.mark.parametrize("continuous", [False, True])
def test_fast_ball_and_thin_wall(oracle, continuous):
    w = oracle.world((0.0, 0.0))
    w.set_flags(warm_starting=True, allow_sleep=False, block_solve=True, continuous=continuous)
    w.add(*_wall_scene(120.0))  # 2 m per step against a 0.1 m wall
    events = 0
    for _ in range(6):
        w.step(np.float32(1 / 60), 8, 3)
        events += w.toi_events
    x = float(w.bodies()[1]["c"][0])
    if continuous:
        assert x < 5.0 - 0.05 - 0.1 + 0.02, x  # stopped in front of the wall
        assert events >= 1
    else:
        assert x > 5.0 + 0.05, x  # tunnelled through
        assert events == 0


def test_fast_boxes_rain_on_thin_ground(oracle):
    """Boxes falling at 60 m/s (1 m per step) onto a 0.1 m thick static ground: with continuous
    physics every box stops on the ground (its centre within the slops of resting height), without
    it most boxes fall through."""
    n = 12
    bd = np.zeros(1 + n, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[1:]["type"] = 2
    bd[1:]["position"] = np.stack([np.arange(n) * 1.5 - 8.0, 3.0 + 0.37 * np.arange(n)], 1)
    bd[1:]["linear_velocity"] = (0.0, -60.0)
    bd[1:]["angle"] = np.linspace(-0.4, 0.4, n)
    sd = np.zeros(1 + n, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (12.0, 0.05)
    sd[1:]["params"][:, :2] = (0.3, 0.2)
    fd = np.zeros(1 + n, b2g.FIXTURE_DEF)
    fd["friction"] = 0.4
    fd["restitution_threshold"] = 1.0
    fd["category"] = 1
    fd["mask"] = 0xFFFF
    fd["density"] = 1.0
    fb = np.arange(1 + n, dtype=np.int32)
    for cont in (False, True):
        w = oracle.world((0.0, -10.0))
        w.set_flags(warm_starting=True, allow_sleep=False, block_solve=True, continuous=cont)
        w.add(bd, fb, sd, fd)
        events = 0
        for _ in range(30):
            w.step(np.float32(1 / 60), 8, 3)
            events += w.toi_events
        y = w.bodies()["c"][1:, 1]
        if cont:
            assert events >= n
            assert (y > 0.05 + 0.2 - 0.02).all() and (y < 0.05 + 0.2 + 0.3).all(), y
        else:
            assert events == 0 and (y < -0.05).sum() >= n // 2, y


def _bullet_scene(bullet):
    """A ball of radius 0.1 at 120 m/s (2 m per step) towards a free dynamic plank 0.1 m thick and
    4 m tall, no gravity: only a bullet gets a time of impact against a dynamic body."""
    bd, fb, sd, fd = _wall_scene(120.0)
    bd[0]["type"] = 2
    bd[0]["gravity_scale"] = 0.0
    bd[1]["bullet"] = int(bullet)
    fd["friction"] = 0.0
    return bd, fb, sd, fd


@pytest.mark.parametrize("bullet", [False, True])
def test_bullet_against_a_dynamic_plank(oracle, bullet):
    """A bullet stops at a free dynamic plank instead of tunnelling through it (TOI events between
    two dynamic bodies, src/private/dynamics/b2_world.rs:616-621), and the TOI solve's impulses
    between the two dynamic bodies conserve linear momentum: m_ball v0 = m_ball v + m_plank V in
    every step. Without the flag the ball passes through (no event between dynamic bodies)."""
    w = oracle.world((0.0, 0.0))
    w.set_flags(warm_starting=True, allow_sleep=False, block_solve=True, continuous=True)
    w.add(*_bullet_scene(bullet))
    m = w.mass()[:, 0]
    p0 = m[1] * 120.0
    events = 0
    for _ in range(6):
        w.step(np.float32(1 / 60), 8, 3)
        events += w.toi_events
        st = w.bodies()
        p = m[0] * st["v"][0, 0] + m[1] * st["v"][1, 0]
        assert abs(p - p0) <= 1e-4 * p0, (p, p0)
    ball, plank = w.bodies()["c"][1, 0], w.bodies()["c"][0, 0]
    if bullet:
        assert events >= 1
        assert ball < plank - 0.05 - 0.1 + 0.02, (ball, plank)
    else:
        assert events == 0 and ball > plank + 0.05, (ball, plank)


def _rain(n=12):
    bd = np.zeros(1 + n, b2g.BODY_DEF)
    bd["gravity_scale"] = 1.0
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    bd[1:]["type"] = 2
    bd[1:]["position"] = np.stack([np.arange(n) * 1.5 - 0.75 * n, 3.0 + 0.37 * np.arange(n)], 1)
    bd[1:]["linear_velocity"] = (0.0, -60.0)
    bd[1:]["angle"] = np.linspace(-0.4, 0.4, n)
    sd = np.zeros(1 + n, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_BOX
    sd[0]["params"][:2] = (1.5 * n, 0.05)
    sd[1:]["params"][:, :2] = (0.3, 0.2)
    fd = np.zeros(1 + n, b2g.FIXTURE_DEF)
    fd["friction"] = 0.4
    fd["restitution_threshold"] = 1.0
    fd["category"] = 1
    fd["mask"] = 0xFFFF
    fd["density"] = 1.0
    return bd, np.arange(1 + n, dtype=np.int32), sd, fd


def test_sub_stepping_resumes_the_pass(oracle):
    """B2world::set_sub_stepping (src/private/dynamics/b2_world.rs:545, 676-679, 896-899, 935): a
    TOI pass stops after each solved event and the following steps resume it -- collide, no solve,
    no reset of alpha0, the cached times and the TOI counts -- until a pass finds no event. Nothing
    between two events of one pass depends on where the steps fall, so once the sub-stepped pass is
    complete the world equals the one stepped without sub-stepping, bit for bit, after the same
    number of passes, with the same events; each sub-step solves at most one event, and a resumed
    step leaves the velocity of a body outside its event unchanged (no solve: no gravity). A reset
    on resuming, a solve in a resumed step or a pass that ends before its events are done fails."""
    DT = np.float32(1 / 60)
    full, sub = oracle.world((0.0, -10.0)), oracle.world((0.0, -10.0))
    for w in (full, sub):
        w.set_flags(warm_starting=True, allow_sleep=False, block_solve=True, continuous=True)
        w.add(*_rain())
    sub.set_sub_stepping(True)
    resumed = 0
    for step in range(30):
        full.step(DT, 8, 3)
        sub.step(DT, 8, 3)
        events = sub.toi_events
        assert events <= 1
        while not sub.step_complete:
            before = sub.bodies()
            sub.step(DT, 8, 3)
            assert sub.toi_events <= 1
            events += sub.toi_events
            resumed += 1
            after = sub.bodies()
            moved = np.any(before["c"] != after["c"], axis=1)
            assert not moved.all()
            np.testing.assert_array_equal(after["v"][~moved], before["v"][~moved])
        assert events == full.toi_events, step
        a, b = full.bodies(), sub.bodies()
        for f in ("c", "a", "v", "w"):
            assert a[f].tobytes() == b[f].tobytes(), (step, f)
    assert resumed >= 10
