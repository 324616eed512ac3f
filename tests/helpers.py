"""Shared helpers of the test-suite: random shapes and poses, record conversion, comparisons."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CIRCLE, EDGE, POLYGON, BOX = 0, 1, 2, 4
# registry-ordered (A, B) shape kinds of the five contact classes
PAIR_KINDS = [("circle", "circle"), ("polygon", "circle"), ("polygon", "polygon"), ("edge", "circle"),
              ("edge", "polygon")]


def random_shape_defs(rng, kind, n, dtype):
    """n shape definitions of one kind: circle, polygon (boxes, rotated boxes and hulls) or edge."""
    sd = np.zeros(n, dtype)
    for i in range(n):
        if kind == "circle":
            sd[i]["type"] = CIRCLE
            sd[i]["radius"] = rng.uniform(0.1, 1.0)
            sd[i]["params"][:2] = rng.uniform(-0.3, 0.3, 2) if rng.random() < 0.3 else 0.0
        elif kind == "polygon":
            r = rng.random()
            if r < 0.3:
                sd[i]["type"] = BOX
                sd[i]["params"][:2] = rng.uniform(0.1, 1.0, 2)
            elif r < 0.45:
                sd[i]["type"] = BOX
                sd[i]["flags"] = 1
                sd[i]["params"][:5] = [*rng.uniform(0.1, 1.0, 2), *rng.uniform(-0.3, 0.3, 2), rng.uniform(-3, 3)]
            else:
                m = int(rng.integers(3, 9))
                rad = rng.uniform(0.2, 1.0)
                ang = 2 * np.pi * np.arange(m) / m + rng.uniform(-0.25, 0.25, m) * 2 * np.pi / m
                pts = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1) + rng.uniform(-0.2, 0.2, 2)
                rng.shuffle(pts)
                sd[i]["type"] = POLYGON
                sd[i]["count"] = m
                sd[i]["params"][:2 * m] = pts.ravel()
        else:
            sd[i]["type"] = EDGE
            t = rng.uniform(0, 2 * np.pi)
            half = 0.5 * rng.uniform(0.5, 3.0) * np.array([np.cos(t), np.sin(t)])
            c = rng.uniform(-0.3, 0.3, 2)
            a, b = c - half, c + half
            if rng.random() < 0.5:
                sd[i]["params"][:4] = [*a, *b]
            else:
                sd[i]["flags"] = 1
                v0 = a + rng.uniform(-1, 1, 2) - 0.5 * (b - a)
                v3 = b + rng.uniform(-1, 1, 2) + 0.5 * (b - a)
                sd[i]["params"][:8] = [*a, *b, *v0, *v3]
    return sd


def extent(sd):
    """Rough radius of each shape definition about its local origin."""
    out = np.zeros(len(sd))
    for i, d in enumerate(sd):
        if d["type"] == CIRCLE:
            out[i] = d["radius"] + np.hypot(*d["params"][:2])
        elif d["type"] == BOX:
            out[i] = np.hypot(d["params"][0], d["params"][1]) + (np.hypot(*d["params"][2:4]) if d["flags"] else 0)
        elif d["type"] == POLYGON:
            pts = d["params"][:2 * d["count"]].reshape(-1, 2)
            out[i] = np.hypot(pts[:, 0], pts[:, 1]).max()
        else:
            pts = d["params"][:4].reshape(-1, 2)
            out[i] = np.hypot(pts[:, 0], pts[:, 1]).max()
    return out


def near_poses(rng, ea, eb, n, oracle):
    """Transforms (p.x, p.y, sin, cos) of n shape pairs placed near or at contact; (sin, cos) are
    computed by the oracle (glibc) so both sides read identical inputs."""
    pa = rng.uniform(-5, 5, (n, 2))
    ang_a = rng.uniform(-np.pi, np.pi, n)
    ang_b = rng.uniform(-np.pi, np.pi, n)
    d = rng.uniform(0.0, 1.05, n) * (ea + eb)
    t = rng.uniform(0, 2 * np.pi, n)
    pb = pa + np.stack([d * np.cos(t), d * np.sin(t)], 1)
    sa = oracle.rot(ang_a.astype(np.float32))
    sb = oracle.rot(ang_b.astype(np.float32))
    xa = np.concatenate([pa.astype(np.float32), sa], 1)
    xb = np.concatenate([pb.astype(np.float32), sb], 1)
    return xa, xb


def compare_manifolds(got, ref, rtol=1e-5, atol=1e-6):
    """Point counts, types and ids must agree exactly; geometry within tolerance. Returns a list
    of (index, reason) mismatches."""
    bad = []
    for i in range(len(ref)):
        g, r = got[i], ref[i]
        if int(g["point_count"]) != int(r["point_count"]):
            bad.append((i, "point_count %d != %d" % (g["point_count"], r["point_count"])))
            continue
        pc = int(r["point_count"])
        if pc == 0:
            continue
        if int(g["type"]) != int(r["type"]):
            bad.append((i, "type"))
            continue
        for p in range(pc):
            if int(g["points"][p]["id"]) != int(r["points"][p]["id"]):
                bad.append((i, "id of point %d" % p))
        geo_g = np.concatenate([g["local_normal"], g["local_point"]] +
                               [g["points"][p]["local_point"] for p in range(pc)]).astype(np.float64)
        geo_r = np.concatenate([r["local_normal"], r["local_point"]] +
                               [r["points"][p]["local_point"] for p in range(pc)]).astype(np.float64)
        if not np.allclose(geo_g, geo_r, rtol=rtol, atol=atol):
            bad.append((i, "geometry %s vs %s" % (geo_g, geo_r)))
    return bad


def glibc():
    libm = ctypes.CDLL("libm.so.6")
    libm.sinf.argtypes = [ctypes.c_float]
    libm.sinf.restype = ctypes.c_float
    libm.cosf.argtypes = [ctypes.c_float]
    libm.cosf.restype = ctypes.c_float
    return libm


def to_numpy_f32(a):
    return np.asarray(a, np.float32)


def overlap_pairs(fat, fixture_body, body_type, big=8.0):
    """Exact set of fixture pairs (lower, higher index) whose fat AABBs overlap as closed boxes
    (b2_test_overlap, src/b2_collision.rs:355-368), on different bodies with at least one dynamic
    (add_pair's filters for the default fixture filter), found with a uniform grid instead of
    all N^2 tests. Boxes with a side over `big` metres (container walls) are tested against every
    box directly. Returns an int64 array of keys lower << 32 | higher, sorted."""
    fat = np.asarray(fat, np.float64).reshape(-1, 4)
    n = len(fat)
    size = np.maximum(fat[:, 2] - fat[:, 0], fat[:, 3] - fat[:, 1])
    large = np.nonzero(size > big)[0]
    small = np.nonzero(size <= big)[0]
    keys = []

    def overlap(a, b):
        return ~((fat[b, 0] > fat[a, 2]) | (fat[b, 1] > fat[a, 3]) | (fat[a, 0] > fat[b, 2]) | (fat[a, 1] > fat[b, 3]))

    if len(small):
        cs = float(size[small].max()) * 1.0001
        f = fat[small]
        x0, y0 = np.floor(f[:, 0] / cs).astype(np.int64), np.floor(f[:, 1] / cs).astype(np.int64)
        x1, y1 = np.floor(f[:, 2] / cs).astype(np.int64), np.floor(f[:, 3] / cs).astype(np.int64)
        ox, oy = x0.min(), y0.min()
        span = int(max(x1.max() - ox, y1.max() - oy)) + 2
        cell, box = [], []
        for dx in (0, 1):
            for dy in (0, 1):  # a box covers at most 2 x 2 cells of side >= its size
                m = (x0 + dx <= x1) & (y0 + dy <= y1)
                cell.append((x0[m] + dx - ox) * span + (y0[m] + dy - oy))
                box.append(small[m])
        cell, box = np.concatenate(cell), np.concatenate(box)
        order = np.lexsort((box, cell))
        cell, box = cell[order], box[order]
        d = 1
        while True:
            same = cell[d:] == cell[:-d]
            if not same.any():
                break
            a, b = box[:-d][same], box[d:][same]
            c = cell[:-d][same]
            ok = overlap(a, b)
            a, b, c = a[ok], b[ok], c[ok]
            # count a pair only in the cell holding the lower corner of the two boxes' overlap
            ref = ((np.floor(np.maximum(fat[a, 0], fat[b, 0]) / cs).astype(np.int64) - ox) * span +
                   (np.floor(np.maximum(fat[a, 1], fat[b, 1]) / cs).astype(np.int64) - oy))
            a, b = a[c == ref], b[c == ref]
            keys.append(np.minimum(a, b).astype(np.int64) << 32 | np.maximum(a, b))
            d += 1
    for g in large:
        other = np.arange(n)
        other = other[other != g]
        ok = overlap(np.full(len(other), g), other)
        o = other[ok]
        keys.append(np.minimum(o, g).astype(np.int64) << 32 | np.maximum(o, g))
    keys = np.unique(np.concatenate(keys)) if keys else np.zeros(0, np.int64)
    body = np.asarray(fixture_body)
    lo, hi = keys >> 32, keys & 0xFFFFFFFF
    bt = np.asarray(body_type)[body]
    keep = (body[lo] != body[hi]) & ((bt[lo] == 2) | (bt[hi] == 2))
    return keys[keep]


def pair_keys(pairs):
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    return np.unique(np.minimum(pairs[:, 0], pairs[:, 1]) << 32 | np.maximum(pairs[:, 0], pairs[:, 1]))
