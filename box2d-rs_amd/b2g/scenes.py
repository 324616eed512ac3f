"""Seeded scene generators for the benchmark and test workloads.

Every generator returns a :class:`Scene`: numpy structured arrays of body definitions,
fixture->body indices, shape definitions and fixture definitions (layouts of ``include/b2g.h``)
that both the GPU world and the CPU oracle consume, in creation order.

Randomness is splitmix64 used as a counter-based generator: draw ``k`` of a stream with seed
``s`` is ``mix(s + (k + 1) * 0x9E3779B97F4A7C15)``; body ``i`` of a scene owns draws
``[i * D, (i + 1) * D)``. Worlds of a batch use seed ``(config << 32) | world``.

Workloads (sizes as named by the benchmark configuration list):
  pyramid       20-row box pyramid on a ground edge (examples/testbed/tests/pyramid.rs:58-87,
                with density 1 and friction 0.6)
  hello_world   the falling box of tests/test.rs:26-102
  large_world   65,536 mixed circles/boxes/polygons on a 1.2 m grid in a 400 x 400 m box
  heap_world    the large-world shape mix in a small container (batched worlds, adjoint)
  stack_world   columns of unit boxes 1.5 m apart (batched worlds)
  stress_world  1,048,576 large-world bodies plus 16,384 towers of 8 boxes in a 1600 x 1600 m box
"""
from collections import namedtuple

import numpy as np

from . import BODY_DEF, FIXTURE_DEF, SHAPE_DEF, B2_DYNAMIC_BODY, B2_STATIC_BODY, SHAPE_BOX, SHAPE_CIRCLE, \
    SHAPE_EDGE, SHAPE_POLYGON, JOINT_DEF, JOINT_REVOLUTE

Scene = namedtuple("Scene", "body_defs fixture_body shape_defs fixture_defs meta")
BatchScene = namedtuple("BatchScene", "body_defs body_world fixture_body shape_defs fixture_defs meta")

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, index):
    """Output number ``index`` (array) of the splitmix64 stream started at ``seed``."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(index, np.uint64) + np.uint64(1)) * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, index):
    """U[0, 1) doubles from the top 53 bits."""
    return (splitmix64(seed, index) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _fixture_defs(n, friction=0.6, density=1.0, restitution=0.0):
    fd = np.zeros(n, FIXTURE_DEF)
    fd["friction"] = friction
    fd["restitution"] = restitution
    fd["restitution_threshold"] = 1.0
    fd["density"] = density
    fd["category"] = 1
    fd["mask"] = 0xFFFF
    return fd


def _body_defs(n, body_type=B2_DYNAMIC_BODY):
    bd = np.zeros(n, BODY_DEF)
    bd["type"] = body_type
    bd["allow_sleep"] = 1
    bd["awake"] = 1
    bd["gravity_scale"] = 1.0
    return bd


def _container(half_w, half_h, cy=0.0):
    """One static body with four two-sided edges around [-half_w, half_w] x [cy-half_h, cy+half_h]."""
    bd = _body_defs(1, B2_STATIC_BODY)
    lo, hi = cy - half_h, cy + half_h
    corners = [(-half_w, lo), (half_w, lo), (half_w, hi), (-half_w, hi)]
    sd = np.zeros(4, SHAPE_DEF)
    sd["type"] = SHAPE_EDGE
    for k in range(4):
        a, b = corners[k], corners[(k + 1) % 4]
        sd[k]["params"][:4] = (a[0], a[1], b[0], b[1])
    fd = _fixture_defs(4, friction=0.6, density=0.0)
    return bd, np.zeros(4, np.int32), sd, fd


def concat(*parts):
    """Concatenate scene parts (body_defs, fixture_body, shape_defs, fixture_defs), re-basing
    fixture->body indices."""
    bds, fbs, sds, fds = [], [], [], []
    base = 0
    for bd, fb, sd, fd in parts:
        bds.append(bd)
        fbs.append(np.asarray(fb, np.int32) + base)
        sds.append(sd)
        fds.append(fd)
        base += len(bd)
    return np.concatenate(bds), np.concatenate(fbs), np.concatenate(sds), np.concatenate(fds)


def hello_world():
    """tests/test.rs:26-102: a 1x1 box (half-extent 1) falls from y=4 onto a ground box."""
    bd = _body_defs(2)
    bd[0]["type"] = B2_STATIC_BODY
    bd[0]["position"] = (0.0, -10.0)
    bd[1]["position"] = (0.0, 4.0)
    sd = np.zeros(2, SHAPE_DEF)
    sd["type"] = SHAPE_BOX
    sd[0]["params"][:2] = (50.0, 10.0)
    sd[1]["params"][:2] = (1.0, 1.0)
    fd = _fixture_defs(2, friction=0.2, density=0.0)
    fd[1]["density"] = 1.0
    fd[1]["friction"] = 0.3
    return Scene(bd, np.array([0, 1], np.int32), sd, fd, {"name": "hello_world", "dynamic": 1})


def pyramid(rows=20, density=1.0, friction=0.6):
    """Box pyramid of examples/testbed/tests/pyramid.rs:58-87 on a ground edge (-40,0)-(40,0)."""
    ground = _body_defs(1, B2_STATIC_BODY)
    gsd = np.zeros(1, SHAPE_DEF)
    gsd["type"] = SHAPE_EDGE
    gsd[0]["params"][:4] = (-40.0, 0.0, 40.0, 0.0)
    gfd = _fixture_defs(1, friction=friction, density=0.0)
    xs = []
    x = np.array([-7.0, 0.75], np.float32)
    dx = np.array([0.5625, 1.25], np.float32)
    dy = np.array([1.125, 0.0], np.float32)
    for i in range(rows):
        y = x.copy()
        for _ in range(i, rows):
            xs.append(y.copy())
            y = (y + dy).astype(np.float32)
        x = (x + dx).astype(np.float32)
    n = len(xs)
    bd = _body_defs(n)
    bd["position"] = np.array(xs, np.float32)
    sd = np.zeros(n, SHAPE_DEF)
    sd["type"] = SHAPE_BOX
    sd["params"][:, 0] = 0.5
    sd["params"][:, 1] = 0.5
    fd = _fixture_defs(n, friction=friction, density=density)
    parts = concat((ground, [0], gsd, gfd), (bd, np.arange(n), sd, fd))
    return Scene(*parts, {"name": "pyramid", "rows": rows, "dynamic": n})


_DRAWS = 24  # draws reserved per body


def _mixed_shapes(u):
    """Shape definitions of the large-world mix from per-body uniforms u[:, 0..]:
    50% circles r~U[0.25,0.5], 30% boxes hx,hy~U[0.25,0.5], 20% convex polygons of 3..8
    vertices at angles 2*pi*k/n + U[-0.25,0.25]*2*pi/n and radius U[0.25,0.5] (passed to set())."""
    n = len(u)
    sd = np.zeros(n, SHAPE_DEF)
    kind = u[:, 0]
    circ = kind < 0.5
    box = (kind >= 0.5) & (kind < 0.8)
    poly = kind >= 0.8
    sd["type"][circ] = SHAPE_CIRCLE
    sd["radius"][circ] = 0.25 + 0.25 * u[circ, 1]
    sd["type"][box] = SHAPE_BOX
    sd["params"][box, 0] = 0.25 + 0.25 * u[box, 1]
    sd["params"][box, 1] = 0.25 + 0.25 * u[box, 2]
    idx = np.nonzero(poly)[0]
    nv = 3 + np.minimum((u[idx, 1] * 6).astype(np.int64), 5)
    rad = 0.25 + 0.25 * u[idx, 2]
    sd["type"][idx] = SHAPE_POLYGON
    sd["count"][idx] = nv
    for j in range(8):
        ang = 2.0 * np.pi * j / nv + (u[idx, 3 + j] - 0.5) * 0.5 * 2.0 * np.pi / nv
        valid = j < nv
        px = (rad * np.cos(ang)).astype(np.float32)
        py = (rad * np.sin(ang)).astype(np.float32)
        sd["params"][idx[valid], 2 * j] = px[valid]
        sd["params"][idx[valid], 2 * j + 1] = py[valid]
    return sd


def _normal(u1, u2):
    return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)


def mixed_bodies(n, seed, centers, speed_sigma=1.0, friction=None):
    """n dynamic bodies of the large-world mix at the given centres (n x 2 float64), with
    velocity N(0, speed_sigma^2) per component, angle U[0, 2 pi), w = 0, density 1."""
    draws = uniform(seed, np.arange(n * _DRAWS, dtype=np.uint64)).reshape(n, _DRAWS)
    sd = _mixed_shapes(draws)
    bd = _body_defs(n)
    bd["position"] = centers.astype(np.float32)
    bd["angle"] = (2.0 * np.pi * draws[:, 11]).astype(np.float32)
    bd["linear_velocity"][:, 0] = (speed_sigma * _normal(draws[:, 12], draws[:, 13])).astype(np.float32)
    bd["linear_velocity"][:, 1] = (speed_sigma * _normal(draws[:, 14], draws[:, 15])).astype(np.float32)
    fd = _fixture_defs(n, friction=0.6, density=1.0)
    if friction is not None:
        fd["friction"] = friction
    return bd, np.arange(n, dtype=np.int32), sd, fd, draws


def large_world(n=65536, seed=1, spacing=1.2, jitter=0.05, half_extent=200.0):
    """65,536 dynamic bodies on a jittered sqrt(n) x sqrt(n) grid (1.2 m spacing) in a static
    400 x 400 m box of four two-sided edges."""
    side = int(np.ceil(np.sqrt(n)))
    i = np.arange(n)
    gx = (i % side - (side - 1) / 2.0) * spacing
    gy = (i // side - (side - 1) / 2.0) * spacing
    ju = uniform(seed ^ 0x5EED, np.arange(2 * n, dtype=np.uint64)).reshape(n, 2)
    centers = np.stack([gx + (ju[:, 0] - 0.5) * 2 * jitter, gy + (ju[:, 1] - 0.5) * 2 * jitter], axis=1)
    bd, fb, sd, fd, _ = mixed_bodies(n, seed, centers)
    parts = concat(_container(half_extent, half_extent), (bd, fb, sd, fd))
    return Scene(*parts, {"name": "large_world", "dynamic": n, "seed": seed})


# Grid spacing of the stress world, chosen so that its 200-step run (dt 1/60, 8/3 iterations,
# sleeping and continuous physics off) has about 3M candidate pairs per step on average: the
# fat-AABB overlaps grow from 0.19M at the first step to 5.4M at the last as the bodies fall and
# pile, a mean of 3.21M at 1.3 m (measured on MI355X over whole runs at 1.0, 1.15, 1.3 and 1.5 m:
# means 5.19M, 4.02M, 3.21M and 3.67M).
STRESS_SPACING = 1.3


def stress_world(n=1 << 20, towers=16384, tower_height=8, seed=5, spacing=STRESS_SPACING, half_extent=800.0):
    """The large-world mix on a sqrt(n) x sqrt(n) jittered grid in the lower part of a
    1600 x 1600 m box, and `towers` columns of `tower_height` unit boxes in rows above it."""
    side = int(np.ceil(np.sqrt(n)))
    i = np.arange(n)
    gx = (i % side - (side - 1) / 2.0) * spacing
    gy = -half_extent + 1.0 + (i // side) * spacing
    ju = uniform(seed ^ 0x5EED, np.arange(2 * n, dtype=np.uint64)).reshape(n, 2)
    centers = np.stack([gx + (ju[:, 0] - 0.5) * 0.1, gy + (ju[:, 1] - 0.5) * 0.1], axis=1)
    bd, fb, sd, fd, _ = mixed_bodies(n, seed, centers)
    top = float(gy.max()) + 3.0
    rows = max(1, int((half_extent - 2.0 - top) // (tower_height + 1.5)))
    per_row = int(np.ceil(towers / rows))
    pitch = (2 * half_extent - 4.0) / per_row
    k = np.arange(towers)
    tx = -half_extent + 2.0 + (k % per_row + 0.5) * pitch
    ty = top + (k // per_row) * (tower_height + 1.5)
    m = towers * tower_height
    tb = _body_defs(m)
    tb["position"][:, 0] = np.repeat(tx, tower_height)
    tb["position"][:, 1] = (np.repeat(ty, tower_height) + 0.5 + np.tile(np.arange(tower_height), towers)).astype(np.float32)
    tsd = np.zeros(m, SHAPE_DEF)
    tsd["type"] = SHAPE_BOX
    tsd["params"][:, :2] = 0.5
    tfd = _fixture_defs(m)
    parts = concat(_container(half_extent, half_extent), (bd, fb, sd, fd), (tb, np.arange(m, dtype=np.int32), tsd, tfd))
    return Scene(*parts, {"name": "stress_world", "dynamic": n + m, "seed": seed})


def heap_world(n, seed, half_extent=20.0, spacing=1.2, boxes_and_circles_only=False, friction=None):
    """n bodies of the large-world mix on a grid inside a small static box container."""
    cols = max(1, int((2 * half_extent - 2.0) // spacing))
    i = np.arange(n)
    gx = (i % cols - (cols - 1) / 2.0) * spacing
    gy = -half_extent + 1.0 + (i // cols) * spacing
    centers = np.stack([gx, gy], axis=1)
    bd, fb, sd, fd, draws = mixed_bodies(n, seed, centers, friction=friction)
    if boxes_and_circles_only:
        poly = sd["type"] == SHAPE_POLYGON
        sd["type"][poly] = SHAPE_BOX
        sd["count"][poly] = 0
        sd["params"][poly] = 0
        sd["params"][poly, 0] = 0.25 + 0.25 * draws[poly, 1]
        sd["params"][poly, 1] = 0.25 + 0.25 * draws[poly, 2]
    top = max(half_extent, -half_extent + 2.0 + (n // cols + 1) * spacing)
    parts = concat(_container(half_extent, (top + half_extent) / 2.0, cy=(top - half_extent) / 2.0),
                   (bd, fb, sd, fd))
    return Scene(*parts, {"name": "heap_world", "dynamic": n, "seed": seed})


def stack_world(n, seed, half_extent=20.0):
    """Columns of unit boxes (half-extent 0.5) of height U{5..30}, 1.5 m apart, until n bodies."""
    u = uniform(seed, np.arange(256, dtype=np.uint64))
    pos = []
    col = 0
    while len(pos) < n:
        h = 5 + int(u[col % 256] * 26)
        x = -half_extent + 1.5 + 1.5 * col
        for r in range(h):
            if len(pos) == n:
                break
            pos.append((x, -half_extent + 0.5 + r * 1.0))
        col += 1
    pos = np.array(pos, np.float64)
    width = max(half_extent, float(pos[:, 0].max()) + 1.5)
    bd = _body_defs(n)
    bd["position"] = pos.astype(np.float32)
    sd = np.zeros(n, SHAPE_DEF)
    sd["type"] = SHAPE_BOX
    sd["params"][:, :2] = 0.5
    fd = _fixture_defs(n)
    parts = concat(_container(width, half_extent), (bd, np.arange(n, dtype=np.int32), sd, fd))
    return Scene(*parts, {"name": "stack_world", "dynamic": n, "seed": seed})


def world_seed(config, world):
    return (int(config) << 32) | int(world)


def batched_worlds(n_worlds=4096, config=3, lo=16, hi=1024):
    """n_worlds independent worlds, body count per world round(exp(U[ln lo, ln hi])); even
    worlds are heaps of the large-world mix in a 40 x 40 m box, odd worlds are box stacks of
    height U{5..30}. World w uses seed (config << 32) | w."""
    parts, worlds, sizes, kinds = [], [], [], []
    for wi in range(n_worlds):
        seed = world_seed(config, wi)
        u = uniform(seed ^ 0xC0FFEE, np.arange(1, dtype=np.uint64))[0]
        nw = int(round(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))))
        sc = heap_world(nw, seed, half_extent=20.0) if wi % 2 == 0 else stack_world(nw, seed, half_extent=20.0)
        parts.append(sc[:4])
        worlds.append(np.full(len(sc.body_defs), wi, np.int32))
        sizes.append(nw)
        kinds.append("heap" if wi % 2 == 0 else "stack")
    bd, fb, sd, fd = concat(*parts)
    meta = {"name": "batched_worlds", "worlds": n_worlds, "dynamic": int(sum(sizes)), "sizes": sizes,
            "kinds": kinds, "config": config}
    return BatchScene(bd, np.concatenate(worlds), fb, sd, fd, meta)


def differentiable_worlds(n_worlds=1024, n=64, config=4):
    """n_worlds worlds of n boxes and circles in a 20 x 20 m box, per-body friction U[0.2, 1.0],
    and one goal point per world U([-8, 8]^2) for the loss |c_body0 - goal|^2 (body 0 = the
    world's first dynamic body)."""
    parts, worlds, goals, first = [], [], [], []
    nb = 0
    for wi in range(n_worlds):
        seed = world_seed(config, wi)
        u = uniform(seed ^ 0xF00D, np.arange(n + 2, dtype=np.uint64))
        sc = heap_world(n, seed, half_extent=10.0, boxes_and_circles_only=True,
                        friction=(0.2 + 0.8 * u[:n]).astype(np.float32))
        parts.append(sc[:4])
        worlds.append(np.full(len(sc.body_defs), wi, np.int32))
        first.append(nb + int(np.nonzero(sc.body_defs["type"] == B2_DYNAMIC_BODY)[0][0]))
        nb += len(sc.body_defs)
        goals.append(-8.0 + 16.0 * u[n:n + 2])
    bd, fb, sd, fd = concat(*parts)
    meta = {"name": "differentiable_worlds", "worlds": n_worlds, "dynamic": n_worlds * n, "config": config,
            "sizes": [n] * n_worlds, "goal": np.array(goals, np.float64), "body0": np.array(first, np.int64)}
    return BatchScene(bd, np.concatenate(worlds), fb, sd, fd, meta)


def world_of(batch, wi):
    """The single-world Scene of world wi of a batch (for per-world references)."""
    bsel = np.nonzero(batch.body_world == wi)[0]
    fsel = np.nonzero(np.isin(batch.fixture_body, bsel))[0]
    remap = np.full(len(batch.body_world), -1, np.int64)
    remap[bsel] = np.arange(len(bsel))
    return Scene(batch.body_defs[bsel], remap[batch.fixture_body[fsel]].astype(np.int32), batch.shape_defs[fsel],
                 batch.fixture_defs[fsel], {"name": "world %d of batch" % wi, "dynamic": int(batch.meta["sizes"][wi])})


def chains(n_chains=4096, links=16, spacing=2.0, rows=32):
    """Hanging chains of the testbed's chain test (examples/testbed/tests/chain.rs:66-104): each
    chain is `links` boxes of half extents 0.6 x 0.125 (density 20, friction 0.2) 1 m apart,
    hinged end to end at integer x by revolute joints with collide_connected off (neighbouring
    links overlap), the first hinged to a static pin. Chains stand in a grid (rows of chains
    `spacing` apart, each row links + 1 m below the last) over a ground box and start horizontal,
    so they swing down. Returns (Scene, joint_defs)."""
    per_row = (n_chains + rows - 1) // rows
    row_h = links + 1.0
    ground = _body_defs(1, B2_STATIC_BODY)
    gsd = np.zeros(1, SHAPE_DEF)
    gsd["type"] = SHAPE_BOX
    gsd["flags"] = 1
    gsd[0]["params"][:5] = (per_row * spacing + links + 10.0, 0.5, 0.0, -0.5, 0.0)
    gfd = _fixture_defs(1, friction=0.6, density=0.0)
    pins = _body_defs(n_chains, B2_STATIC_BODY)
    n = n_chains * links
    bd = _body_defs(n)
    sd = np.zeros(n, SHAPE_DEF)
    sd["type"] = SHAPE_BOX
    sd["params"][:, 0] = 0.6
    sd["params"][:, 1] = 0.125
    fd = _fixture_defs(n, friction=0.2, density=20.0)
    jd = np.zeros(n, JOINT_DEF)
    jd["type"] = JOINT_REVOLUTE
    base_pin = 1
    base_link = 1 + n_chains
    for c in range(n_chains):
        x0 = (c % per_row) * spacing - 0.5 * per_row * spacing
        y = 2.0 + (rows - 1 - c // per_row) * row_h + links
        pins[c]["position"] = (x0, y)
        for k in range(links):
            i = c * links + k
            bd[i]["position"] = (x0 + 0.5 + k, y)
            jd[i]["body_a"] = base_pin + c if k == 0 else base_link + i - 1
            jd[i]["body_b"] = base_link + i
            jd[i]["local_anchor_a"] = (0.0, 0.0) if k == 0 else (0.5, 0.0)
            jd[i]["local_anchor_b"] = (-0.5, 0.0)
    parts = concat((ground, [0], gsd, gfd), (pins, np.zeros(0, np.int32), np.zeros(0, SHAPE_DEF),
                                             np.zeros(0, FIXTURE_DEF)),
                   (bd, np.arange(n), sd, fd))
    return Scene(*parts, {"name": "chains", "dynamic": n, "chains": n_chains, "links": links}), jd
