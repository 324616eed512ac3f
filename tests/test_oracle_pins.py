"""Pins of the CPU oracle against what the reference itself fixes, without running it.

Reference-held asserts:
  tests/test.rs:99-101            hello_world: box at rest at y = 1.01 +- 0.01 after 60 steps
  tests/world_test.rs:79-80,86-87 circles 100 m apart: no contact; 1 m apart: a contact
  tests/collision_test.rs:47-80   polygon centroid, mass and inertia closed forms
  tests/math_test.rs:37-47        sweep transform at beta 0/1; B2Rot equals (cos a, sin a)
  examples/testbed/tests/dynamic_tree.rs:219-241  tree query == brute-force AABB overlap
Closed forms, brute force and invariants:
  proxy ids 0, 1, 3, 5, ... in a create-only world (src/private/collision/b2_dynamic_tree.rs)
  free fall (symplectic Euler), Pade damping, circle-circle manifold, box/circle mass,
  resting-box impulse = m g h, hull vs an independent monotone-chain hull, SAT depth vs an
  independent f64 separating-axis computation, 2-point block LCP vs exhaustive enumeration,
  linear momentum conservation and non-negative normal impulses; sleep: a box placed at rest
  sleeps in the step whose f32 sum of h reaches the 0.5 s time to sleep, with zero velocity and
  the sleep time reset, never when the world or one body of its island forbids sleeping, and a
  stack sleeps as one island and stays frozen.
Each check is chosen so that a plausible slip (dropped term, wrong sign, swapped operand,
wrong index) fails it.
"""
import ctypes

import numpy as np
import pytest

from helpers import glibc

CIRCLE, EDGE, POLYGON, BOX = 0, 1, 2, 4


def _world(oracle, bodies, fixtures, gravity=(0.0, -10.0)):
    """bodies: list of dicts(type, position, angle, v, w); fixtures: (body, shape_def dict, density, friction)"""
    dt = oracle.dt
    bd = np.zeros(len(bodies), dt["body_def"])
    bd["gravity_scale"] = 1
    bd["awake"] = 1
    bd["allow_sleep"] = 1
    for i, b in enumerate(bodies):
        bd[i]["type"] = b.get("type", 2)
        bd[i]["position"] = b.get("position", (0, 0))
        bd[i]["angle"] = b.get("angle", 0)
        bd[i]["linear_velocity"] = b.get("v", (0, 0))
        bd[i]["angular_velocity"] = b.get("w", 0)
        bd[i]["linear_damping"] = b.get("ld", 0)
        bd[i]["angular_damping"] = b.get("ad", 0)
        bd[i]["allow_sleep"] = b.get("allow_sleep", 1)
    sd = np.zeros(len(fixtures), dt["shape_def"])
    fd = np.zeros(len(fixtures), dt["fixture_def"])
    fb = np.zeros(len(fixtures), np.int32)
    for k, (body, shape, density, friction) in enumerate(fixtures):
        fb[k] = body
        for key, val in shape.items():
            sd[k][key] = val
        fd[k]["density"] = density
        fd[k]["friction"] = friction
        fd[k]["restitution_threshold"] = 1
        fd[k]["category"] = 1
        fd[k]["mask"] = 0xFFFF
    w = oracle.world(gravity)
    proxies = w.add(bd, fb, sd, fd)
    return w, proxies


def _box(hx, hy):
    p = np.zeros(36)
    p[:2] = (hx, hy)
    return {"type": BOX, "params": p}


def _circle(r, p=(0, 0)):
    q = np.zeros(36)
    q[:2] = p
    return {"type": CIRCLE, "radius": r, "params": q}


def test_hello_world_assert(oracle):
    """tests/test.rs:26-102, with B2world::new's defaults: warm starting, sleeping and continuous
    physics on."""
    w, _ = _world(oracle, [{"type": 0, "position": (0, -10)}, {"type": 2, "position": (0, 4)}],
                  [(0, _box(50, 10), 0.0, 0.2), (1, _box(1, 1), 1.0, 0.3)])
    w.set_flags(warm_starting=True, allow_sleep=True, block_solve=True, continuous=True)
    for _ in range(60):
        w.step(1.0 / 60.0, 6, 2)
    b = w.bodies()[1]
    assert abs(b["p"][0]) < 0.01
    assert abs(b["p"][1] - 1.01) < 0.01
    assert abs(b["a"]) < 0.01


def test_begin_contact_circles(oracle):
    """tests/world_test.rs:51-88: r=5 circles 100 m apart have no contact; 1 m apart they do."""
    w, _ = _world(oracle, [{"position": (0, 0)}, {"position": (100, 0)}], [(0, _circle(5), 1, 0.2),
                                                                            (1, _circle(5), 1, 0.2)],
                  gravity=(0, 0))
    w.step(1.0 / 60.0, 8, 3)
    assert w.contact_count == 0
    w2, _ = _world(oracle, [{"position": (0, 0)}, {"position": (1, 0)}], [(0, _circle(5), 1, 0.2),
                                                                           (1, _circle(5), 1, 0.2)],
                   gravity=(0, 0))
    w2.step(1.0 / 60.0, 8, 3)
    fab, flags, man, _ = w2.contacts()
    assert len(fab) == 1 and flags[0] == 1 and man[0]["point_count"] == 1


def test_polygon_mass_data(oracle):
    """tests/collision_test.rs:30-81, for set_as_box_angle and for the hull of the 4 corners."""
    center = np.array([100.0, -50.0], np.float32)
    hx, hy, angle = np.float32(0.5), np.float32(1.5), np.float32(0.25)
    eps = np.float32(np.finfo(np.float32).eps)
    abs_tol, rel_tol = 2 * eps, 2 * eps
    sd = np.zeros(2, oracle.dt["shape_def"])
    sd[0]["type"] = BOX
    sd[0]["flags"] = 1
    sd[0]["params"][:5] = (hx, hy, center[0], center[1], angle)
    sd[1]["type"] = POLYGON
    sd[1]["count"] = 4
    sd[1]["params"][:8] = (center[0] - hx, center[1] - hy, center[0] + hx, center[1] - hy, center[0] - hx,
                           center[1] + hy, center[0] + hx, center[1] + hy)
    shapes, st = oracle.make_shapes(sd)
    assert list(st) == [0, 0]
    mass = np.float32(4.0) * hx * hy
    inertia = (mass / np.float32(3.0)) * (hx * hx + hy * hy) + mass * np.float32(center @ center)
    md = oracle.shape_mass(shapes, 1.0)
    for s, m in zip(shapes, md):
        for k in range(2):
            assert abs(s["centroid"][k] - center[k]) < abs_tol + rel_tol * abs(center[k])
            assert abs(m[1 + k] - center[k]) < abs_tol + rel_tol * abs(center[k])
        assert abs(m[0] - mass) < 20.0 * (abs_tol + rel_tol * mass)
        assert abs(m[3] - inertia) < 40.0 * (abs_tol + rel_tol * inertia)


def test_rotation_is_glibc(oracle):
    """tests/math_test.rs:37-47 requires B2Rot == (cos a, sin a) of the platform libm."""
    libm = glibc()
    a = np.random.default_rng(0).uniform(-50, 50, 2000).astype(np.float32)
    sc = oracle.rot(a)
    ref = np.array([(libm.sinf(float(x)), libm.cosf(float(x))) for x in a], np.float32)
    assert np.array_equal(sc, ref)


def test_sweep_transform_endpoints(oracle):
    """tests/math_test.rs:25-49: a body created at (p, angle) has xf.p == p and the rotation of its
    angle; after a step its transform is the sweep end (synchronize_transform)."""
    w, _ = _world(oracle, [{"position": (1.5, -2.25), "angle": 0.7, "v": (1, 2), "w": 0.3}],
                  [(0, _circle(0.5, (0.25, 0.1)), 1, 0.2)], gravity=(0, 0))
    b = w.bodies()[0]
    assert tuple(b["p"]) == (np.float32(1.5), np.float32(-2.25))
    assert tuple(b["q"]) == tuple(oracle.rot(np.array([0.7], np.float32))[0])
    w.step(1.0 / 60.0, 8, 3)
    b = w.bodies()[0]
    assert tuple(b["q"]) == tuple(oracle.rot(np.array([b["a"]], np.float32))[0])
    # xf.p = c - R(a) * local_center with local centre (0.25, 0.1)
    s, c = b["q"]
    lc = np.array([0.25, 0.1], np.float32)
    px = b["c"][0] - (c * lc[0] - s * lc[1])
    py = b["c"][1] - (s * lc[0] + c * lc[1])
    assert abs(px - b["p"][0]) < 1e-6 and abs(py - b["p"][1]) < 1e-6


def test_proxy_ids_of_a_create_only_world(oracle):
    """Node allocation of the dynamic tree: first leaf 0, then every insert allocates a parent
    before the leaf is reused... leaf k (k >= 1) gets 2k - 1."""
    bodies = [{"position": (3.0 * i, 0)} for i in range(40)]
    fixtures = [(i, _circle(0.5), 1, 0.2) for i in range(40)]
    _, proxies = _world(oracle, bodies, fixtures, gravity=(0, 0))
    assert list(proxies) == [0] + [2 * k - 1 for k in range(1, 40)]


def test_tree_pairs_equal_brute_force(oracle):
    """examples/testbed/tests/dynamic_tree.rs:219-241: tree queries == brute-force overlap. After
    collide the oracle's contact set is exactly the filtered overlap set of the fat AABBs."""
    import sys
    from b2g import scenes
    sc = scenes.heap_world(700, seed=1, half_extent=10.0)
    w = oracle.world()
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    for step in range(30):
        w.step(1.0 / 60.0, 8, 3)
        if step % 10 == 9:
            w.collide()
            fat, _ = w.proxies()
            brute = oracle.brute_pairs(fat)
            body = np.asarray(sc.fixture_body)
            btype = sc.body_defs["type"][body]
            keep = (body[brute[:, 0]] != body[brute[:, 1]]) & ((btype[brute[:, 0]] == 2) | (btype[brute[:, 1]] == 2))
            brute = brute[keep]
            fab, _, _, _ = w.contacts()
            lo, hi = np.minimum(fab[:, 0], fab[:, 1]), np.maximum(fab[:, 0], fab[:, 1])
            a = np.unique(lo.astype(np.int64) << 32 | hi)
            b = np.unique(brute[:, 0].astype(np.int64) << 32 | brute[:, 1])
            assert len(a) == len(fab) and np.array_equal(a, b)


def test_free_fall_closed_form(oracle64):
    """Symplectic Euler: v_n = v0 + n h g, x_n = x0 + h sum_{k=1..n} v_k (f64 oracle, exact to
    round-off)."""
    w, _ = _world(oracle64, [{"position": (0, 100), "v": (1.5, 2.0)}], [(0, _circle(0.5), 1, 0.2)])
    h, g, n = 1.0 / 60.0, -10.0, 90
    for _ in range(n):
        w.step(h, 8, 3)
    b = w.bodies()[0]
    vy = 2.0 + n * h * g
    y = 100.0 + h * sum(2.0 + k * h * g for k in range(1, n + 1))
    assert abs(b["v"][1] - vy) < 1e-10 and abs(b["c"][1] - y) < 1e-9
    assert abs(b["c"][0] - (1.5 * n * h)) < 1e-10


def test_pade_damping(oracle64):
    """v_{n} = v_0 / (1 + h c)^n with no gravity (src/private/dynamics/b2_island_private.rs:159-168)."""
    w, _ = _world(oracle64, [{"position": (0, 0), "v": (3.0, -1.0), "w": 2.0, "ld": 0.7, "ad": 1.3}],
                  [(0, _circle(0.5), 1, 0.2)], gravity=(0, 0))
    h, n = 1.0 / 60.0, 50
    for _ in range(n):
        w.step(h, 8, 3)
    b = w.bodies()[0]
    assert abs(b["v"][0] - 3.0 / (1 + h * 0.7) ** n) < 1e-12
    assert abs(b["w"] - 2.0 / (1 + h * 1.3) ** n) < 1e-12


def test_circle_circle_manifold_closed_form(oracle64):
    """Two circles: normal along the centre line, separation = distance - rA - rB, point midway
    between the surface points."""
    rng = np.random.default_rng(5)
    sd = np.zeros(2, oracle64.dt["shape_def"])
    sd["type"] = CIRCLE
    sd[0]["radius"], sd[1]["radius"] = 0.7, 0.4
    sd[0]["params"][:2] = (0.1, -0.2)
    shapes, _ = oracle64.make_shapes(sd)
    for _ in range(50):
        pa, pb = rng.uniform(-1, 1, 2), rng.uniform(-1, 1, 2)
        aa, ab = rng.uniform(-3, 3, 2)
        xa = np.array([*pa, np.sin(aa), np.cos(aa)])
        xb = np.array([*pb, np.sin(ab), np.cos(ab)])
        m = oracle64.collide(shapes[:1], xa, shapes[1:], xb)
        ca = pa + np.array([np.cos(aa) * 0.1 + np.sin(aa) * 0.2, np.sin(aa) * 0.1 - np.cos(aa) * 0.2])
        cb = pb
        d = np.linalg.norm(cb - ca)
        if d > 1.1:
            assert m[0]["point_count"] == 0
            continue
        assert m[0]["point_count"] == 1
        wm = oracle64.world_manifold(m, xa, [0.7], xb, [0.4])[0]
        n = (cb - ca) / d
        assert np.allclose(wm[:2], n, atol=1e-12)
        assert abs(wm[6] - (d - 1.1)) < 1e-12
        assert np.allclose(wm[2:4], 0.5 * ((ca + 0.7 * n) + (cb - 0.4 * n)), atol=1e-12)


def test_box_and_circle_mass(oracle64):
    sd = np.zeros(2, oracle64.dt["shape_def"])
    sd[0]["type"] = BOX
    sd[0]["params"][:2] = (0.3, 0.8)
    sd[1]["type"] = CIRCLE
    sd[1]["radius"] = 0.6
    sd[1]["params"][:2] = (0.2, -0.1)
    shapes, _ = oracle64.make_shapes(sd)
    md = oracle64.shape_mass(shapes, 2.5)
    m = 2.5 * 4 * 0.3 * 0.8
    assert abs(md[0][0] - m) < 1e-12 and abs(md[0][3] - m / 3 * (0.09 + 0.64)) < 1e-12
    mc = 2.5 * np.pi * 0.36
    assert abs(md[1][0] - mc) < 1e-12
    assert abs(md[1][3] - mc * (0.5 * 0.36 + 0.05)) < 1e-12


def test_resting_box_impulse_and_slop(oracle64):
    """A box resting on the ground: the normal impulses of a step sum to m g h, and the
    penetration settles at the linear slop (the position solver stops at -slop)."""
    w, _ = _world(oracle64, [{"type": 0, "position": (0, 0)}, {"position": (0, 0.51)}],
                  [(0, _box(5, 0.5), 0.0, 0.6), (1, _box(0.5, 0.5), 1.0, 0.6)])
    for _ in range(240):
        w.step(1.0 / 60.0, 8, 3)
    fab, flags, man, _ = w.contacts()
    assert flags[0] == 1 and man[0]["point_count"] == 2
    total = man[0]["points"]["normal_impulse"].sum()
    m = 1.0
    assert abs(total - m * 10.0 / 60.0) < 1e-6
    b = w.bodies()[1]
    # resting height: ground top 0.5 + box half 0.5 + skins 0.02 - slop 0.005
    assert abs(b["c"][1] - (1.0 + 0.02 - 0.005)) < 1e-4


def _monotone_chain(pts):
    pts = sorted(map(tuple, pts))

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])
    lower, upper = [], []
    for p in pts:
        while len(lower) >= 2 and cross(lower[-2], lower[-1], p) <= 0:
            lower.pop()
        lower.append(p)
    for p in reversed(pts):
        while len(upper) >= 2 and cross(upper[-2], upper[-1], p) <= 0:
            upper.pop()
        upper.append(p)
    return lower[:-1] + upper[:-1]


def test_hull_against_monotone_chain(oracle64):
    """The gift-wrap hull of set() equals an independent monotone-chain hull, rotated to start at
    the rightmost (lowest-y on ties) point, counter-clockwise; normals are unit, outward."""
    rng = np.random.default_rng(12)
    for _ in range(200):
        m = int(rng.integers(3, 9))
        pts = rng.uniform(-1, 1, (m, 2))
        sd = np.zeros(1, oracle64.dt["shape_def"])
        sd[0]["type"] = POLYGON
        sd[0]["count"] = m
        sd[0]["params"][:2 * m] = pts.ravel()
        s, st = oracle64.make_shapes(sd)
        ref = _monotone_chain(pts)
        if st[0] != 0 or len(ref) < 3:
            continue
        k = max(range(len(ref)), key=lambda i: (ref[i][0], -ref[i][1]))
        ref = ref[k:] + ref[:k]
        got = s[0]["vertices"][:2 * s[0]["count"]].reshape(-1, 2)
        assert s[0]["count"] == len(ref)
        assert np.allclose(got, np.array(ref), atol=0)
        nrm = s[0]["normals"][:2 * s[0]["count"]].reshape(-1, 2)
        assert np.allclose(np.hypot(nrm[:, 0], nrm[:, 1]), 1.0, atol=1e-12)
        c = got.mean(0)
        for i in range(len(got)):
            assert np.dot(nrm[i], got[i] - c) > 0


def _sat_depth(va, vb):
    """Minimum penetration depth over all edge normals of two convex CCW polygons (f64)."""
    best = -np.inf
    for P, Q in ((va, vb), (vb, va)):
        for i in range(len(P)):
            e = P[(i + 1) % len(P)] - P[i]
            n = np.array([e[1], -e[0]]) / np.hypot(*e)
            best = max(best, np.min((Q - P[i]) @ n))
    return best


def test_polygon_separation_against_sat(oracle64):
    """Deepest manifold separation of two polygons equals the independent SAT separation
    (within the radius skins), away from ties; A's normal points towards B."""
    rng = np.random.default_rng(31)
    checked = 0
    for _ in range(400):
        sd = np.zeros(2, oracle64.dt["shape_def"])
        sd["type"] = BOX
        sd[0]["params"][:2] = rng.uniform(0.3, 1.0, 2)
        sd[1]["params"][:2] = rng.uniform(0.3, 1.0, 2)
        s, _ = oracle64.make_shapes(sd)
        aa, ab = rng.uniform(-np.pi, np.pi, 2)
        pa = np.zeros(2)
        pb = rng.uniform(-1.6, 1.6, 2)
        xa = np.array([*pa, np.sin(aa), np.cos(aa)])
        xb = np.array([*pb, np.sin(ab), np.cos(ab)])
        R = lambda a: np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
        va = s[0]["vertices"][:8].reshape(4, 2) @ R(aa).T + pa
        vb = s[1]["vertices"][:8].reshape(4, 2) @ R(ab).T + pb
        depth = _sat_depth(va, vb)
        m = oracle64.collide(s[:1], xa, s[1:], xb)
        if depth > 0.03:
            assert m[0]["point_count"] == 0
            continue
        if depth < -0.5 or m[0]["point_count"] == 0:
            continue
        wm = oracle64.world_manifold(m, xa, [0.01], xb, [0.01])[0]
        seps = wm[6:6 + m[0]["point_count"]]
        # the deepest clip point's separation is the SAT depth minus the two skins
        assert abs(seps.min() - (depth - 0.02)) < 1e-9
        assert np.dot(wm[:2], (vb.mean(0) - va.mean(0))) > 0
        checked += 1
    assert checked > 50


def _lcp_enumerate(A, b):
    """All solutions of w = A x + b, x >= 0, w >= 0, x.w = 0 for 2x2 A (f64, exhaustive)."""
    sols = []
    for free in ((True, True), (True, False), (False, True), (False, False)):
        x = np.zeros(2)
        idx = [i for i in range(2) if free[i]]
        if idx:
            sub = A[np.ix_(idx, idx)]
            if abs(np.linalg.det(sub)) < 1e-14:
                continue
            x[idx] = np.linalg.solve(sub, -b[idx])
        w = A @ x + b
        if np.all(x >= -1e-12) and np.all(w >= -1e-12) and abs(x @ w) < 1e-9:
            sols.append(x)
    return sols


def test_block_lcp_against_enumeration(oracle64):
    """One velocity iteration of a 2-point contact (box falling flat onto ground, no friction)
    lands on an exact LCP solution: impulses >= 0, post-solve normal velocities >= bias, and
    complementarity (checked against an independent 2x2 enumeration)."""
    w, _ = _world(oracle64, [{"type": 0, "position": (0, 0)}, {"position": (0, 1.0), "v": (0.3, -2.0), "w": 0.8}],
                  [(0, _box(5, 0.5), 0.0, 0.0), (1, _box(0.5, 0.5), 1.0, 0.0)], gravity=(0, 0))
    for _ in range(3):
        w.step(1.0 / 60.0, 1, 0)
    fab, flags, man, _ = w.contacts()
    b = w.bodies()[1]
    assert flags[0] == 1 and man[0]["point_count"] == 2
    x = man[0]["points"]["normal_impulse"]
    assert np.all(x >= 0)
    # reconstruct A and b of the LCP at the solution: the oracle's impulses must be one of the
    # enumerated solutions for the post-velocity-solve state (complementarity)
    xa = np.array([0, 0, 0, 1.0])
    xb = np.concatenate([b["p"], b["q"]])
    wm = oracle64.world_manifold(man, xa, [0.01], xb, [0.01])[0]
    n = wm[:2]
    r = wm[2:6].reshape(2, 2) - b["c"]
    mB, iB = 1.0, 1.0 / (1.0 / 12 * (1 + 1))
    rn = r[:, 0] * n[1] - r[:, 1] * n[0]
    A = mB + iB * np.outer(rn, rn)
    vn = np.array([np.dot(n, b["v"] + b["w"] * np.array([-ri[1], ri[0]])) for ri in r])
    # velocities already include x: vn = A (x - x) + b'  =>  b' = vn - A x at the solution
    sols = _lcp_enumerate(A, vn - A @ x)
    assert any(np.allclose(s, x, atol=1e-9) for s in sols)


def test_momentum_conserved_between_dynamic_bodies(oracle64):
    """Equal and opposite impulses: two free bodies colliding (no gravity) keep total momentum."""
    w, _ = _world(oracle64, [{"position": (0, 0), "v": (2, 0.3)}, {"position": (1.6, 0.2), "v": (-1, 0), "w": 0.5}],
                  [(0, _box(0.5, 0.4), 2.0, 0.6), (1, _circle(0.5), 1.0, 0.6)], gravity=(0, 0))
    mass = w.mass()[:, 0]
    p0 = (mass[:, None] * w.bodies()["v"]).sum(0)
    touched = False
    for _ in range(60):
        w.step(1.0 / 60.0, 8, 3)
        touched |= w.contact_count > 0 and bool(w.contacts()[1].any())
    p1 = (mass[:, None] * w.bodies()["v"]).sum(0)
    assert touched
    assert np.abs(p1 - p0).max() < 1e-12


def test_normal_impulses_non_negative(oracle):
    """After a step every stored normal impulse is >= 0 (the accumulated impulse is clamped at
    zero, src/private/dynamics/b2_contact_solver_private.rs:338). The friction clamp uses the
    normal impulse of the previous iteration (friction is solved first), so |lambda_t| <= mu
    lambda_n is not an invariant of the stored pair and is not asserted."""
    from b2g import scenes
    sc = scenes.heap_world(400, seed=6, half_extent=9.0)
    w = oracle.world()
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    for _ in range(60):
        w.step(1.0 / 60.0, 8, 3)
    fab, flags, man, mat = w.contacts()
    t = flags == 1
    pc = man["point_count"][t]
    ni = man["points"]["normal_impulse"][t]
    valid = np.arange(2)[None, :] < pc[:, None]
    assert valid.sum() > 150
    assert np.all(ni[valid] >= 0)


# ---- sleep (src/private/dynamics/b2_island_private.rs:291-327, src/b2_common.rs:85-91) ----
def _resting_boxes(oracle, n_boxes=1, allow=None, world_sleep=True):
    """Boxes stacked at their resting heights on a static ground box (top at y = 0.5): box i
    at y = 1.015 + 1.015 i (half 0.5, two polygon skins 0.02, minus the 0.005 slop)."""
    allow = allow or [1] * n_boxes
    bodies = [{"type": 0, "position": (0, 0)}] + [{"position": (0, 1.015 * (i + 1)), "allow_sleep": allow[i]}
                                                  for i in range(n_boxes)]
    fixtures = [(0, _box(5, 0.5), 0.0, 0.6)] + [(i + 1, _box(0.5, 0.5), 1.0, 0.6) for i in range(n_boxes)]
    w, _ = _world(oracle, bodies, fixtures)
    w.set_flags(warm_starting=True, allow_sleep=world_sleep, block_solve=True)
    return w


def _steps_to_time_to_sleep(h=np.float32(1.0 / 60.0)):
    t, n = np.float32(0.0), 0
    while not t >= np.float32(0.5):
        t = np.float32(t + h)
        n += 1
    return n


def test_resting_box_sleeps_after_time_to_sleep(oracle):
    """A box placed at rest stays below the sleep tolerances from the first step, so its sleep
    time grows by h each step (f32 sum) and it sleeps in the step that brings it to 0.5 s,
    with zero velocity and the sleep time reset."""
    w = _resting_boxes(oracle)
    n = _steps_to_time_to_sleep()
    for k in range(1, n + 1):
        w.step(1.0 / 60.0, 8, 3)
        if k < n:
            assert w.bodies()[1]["awake"] == 1
    b = w.bodies()[1]
    assert b["awake"] == 0 and b["sleep_time"] == 0 and b["w"] == 0 and np.all(b["v"] == 0)


def test_sleep_time_sums_h_in_float(oracle):
    w = _resting_boxes(oracle)
    h = np.float32(1.0 / 60.0)
    t = np.float32(0.0)
    for _ in range(10):
        w.step(1.0 / 60.0, 8, 3)
        t = np.float32(t + h)
        assert w.bodies()[1]["sleep_time"] == t


def test_no_sleep_when_the_world_forbids_it(oracle):
    w = _resting_boxes(oracle, world_sleep=False)
    for _ in range(3 * _steps_to_time_to_sleep()):
        w.step(1.0 / 60.0, 8, 3)
    b = w.bodies()[1]
    assert b["awake"] == 1 and b["sleep_time"] == 0


def test_stack_sleeps_as_one_island_and_stays_frozen(oracle):
    """Two stacked boxes form one island: both fall asleep in the same step; asleep, nothing
    moves."""
    w = _resting_boxes(oracle, n_boxes=2)
    n = _steps_to_time_to_sleep()
    for _ in range(n - 1):
        w.step(1.0 / 60.0, 8, 3)
    assert np.all(w.bodies()["awake"][1:] == 1)
    w.step(1.0 / 60.0, 8, 3)
    assert np.all(w.bodies()["awake"][1:] == 0)
    before = w.bodies()
    for _ in range(60):
        w.step(1.0 / 60.0, 8, 3)
    after = w.bodies()
    for f in ("c", "a", "v", "w", "p", "q"):
        np.testing.assert_array_equal(before[f], after[f])


def test_a_body_that_may_not_sleep_keeps_its_island_awake(oracle):
    """The island sleeps only when every body reached the time to sleep; a body whose def
    forbids sleeping resets its sleep time every step, so its whole stack stays awake."""
    w = _resting_boxes(oracle, n_boxes=2, allow=[1, 0])
    for _ in range(3 * _steps_to_time_to_sleep()):
        w.step(1.0 / 60.0, 8, 3)
    b = w.bodies()
    assert np.all(b["awake"][1:] == 1)
    assert b["sleep_time"][2] == 0 and b["sleep_time"][1] > 0.5
