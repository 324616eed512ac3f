"""Pins of the oracle's polygon-circle, edge-circle, edge-polygon and general-hull polygon-polygon
collide functions against geometry computed independently in f64, without the reference's
algorithm.

  b2_collide_polygon_and_circle  src/private/collision/b2_collide_circle.rs:36-133
  b2_collide_edge_and_circle     src/private/collision/b2_collide_edge.rs:11-123
  b2_collide_edge_and_polygon    src/private/collision/b2_collide_edge.rs:230-475
  b2_collide_polygons, b2_find_max_separation
                                 src/private/collision/b2_collide_polygon.rs:8-46, 102-225

What fixes each expected value:
  * polygon-circle: the closest feature of the polygon to the circle centre by brute force over
    clamped edge projections (inside: the face of largest signed distance). The manifold's local
    normal and point are that face's normal and midpoint, or the closest vertex and the unit vector
    from it; the world separation is the signed distance minus both radii; a contact exists iff the
    signed distance is at most the radii's sum.
  * edge-circle: the clamped projection of the centre on the segment; region A or B give a
    circles manifold at that vertex with feature id (vertex 0 or 1, vertex), region AB a face
    manifold whose normal points to the centre's side; the separation is the point-segment
    distance minus the radii; one-sided edges admit only the right side and leave a centre that
    lies in a neighbouring edge's interior region to that edge.
  * edge-polygon (two-sided): the separating-axis depths over the edge normal (both signs) and the
    polygon's normals computed directly from vertex projections; the clip point count is the number
    of incident-face end points whose depth is within the skins, the ids follow the feature that
    made each point, and the deepest separation equals the chosen axis' depth minus the skins.
  * polygon-polygon on random 3-8-gon hulls: the same separating-axis depths, the flip rule's
    choice, the incident face as the most anti-parallel face, and ids and depths as above.
Cases within a small distance of a tie (two features equally near, a depth at a threshold, an end
point on a side plane) are skipped; each test asserts how many cases it checked. Each check fails
on a plausible slip: a flipped normal or sign, swapped vertex ids or indices, a wrong region.
"""
import numpy as np
import pytest

CIRCLE, EDGE, POLYGON, BOX = 0, 1, 2, 4
E_CIRCLES, E_FACE_A, E_FACE_B = 0, 1, 2
VERTEX, FACE = 0, 1
SKIN = 0.01  # polygon and edge radius: 2 * linear slop (src/b2_common.rs:39,48)


def _key(index_a, index_b, type_a, type_b):
    return index_a | index_b << 8 | type_a << 16 | type_b << 24


def _xf(rng, spread=1.0):
    a = rng.uniform(-np.pi, np.pi)
    return np.array([*rng.uniform(-spread, spread, 2), np.sin(a), np.cos(a)])


def _rot(xf):
    s, c = xf[2], xf[3]
    return np.array([[c, -s], [s, c]])


def _apply(xf, v):
    return v @ _rot(xf).T + xf[:2]


def _inv_apply(xf, v):
    return (v - xf[:2]) @ _rot(xf)


def _hull_shapes(oracle, rng, n):
    """n random polygons through the oracle's hull (pinned against a monotone chain elsewhere)."""
    out = []
    while len(out) < n:
        sd = np.zeros(1, oracle.dt["shape_def"])
        m = int(rng.integers(3, 9))
        rad = rng.uniform(0.3, 1.0)
        ang = 2 * np.pi * np.arange(m) / m + rng.uniform(-0.25, 0.25, m) * 2 * np.pi / m
        pts = np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1) + rng.uniform(-0.2, 0.2, 2)
        rng.shuffle(pts)
        sd[0]["type"] = POLYGON
        sd[0]["count"] = m
        sd[0]["params"][:2 * m] = pts.ravel()
        s, st = oracle.make_shapes(sd)
        if st[0] == 0:
            out.append(s[0])
    return np.array(out)


def _verts(shape):
    k = int(shape["count"])
    return shape["vertices"][:2 * k].reshape(k, 2).astype(np.float64), \
        shape["normals"][:2 * k].reshape(k, 2).astype(np.float64)


def _circle_shape(oracle, r, p):
    sd = np.zeros(1, oracle.dt["shape_def"])
    sd[0]["type"] = CIRCLE
    sd[0]["radius"] = r
    sd[0]["params"][:2] = p
    return oracle.make_shapes(sd)[0]


def _edge_shape(oracle, a, b, v0=None, v3=None):
    sd = np.zeros(1, oracle.dt["shape_def"])
    sd[0]["type"] = EDGE
    if v0 is None:
        sd[0]["params"][:4] = [*a, *b]
    else:
        sd[0]["flags"] = 1
        sd[0]["params"][:8] = [*a, *b, *v0, *v3]
    return oracle.make_shapes(sd)[0]


def _closest_feature(verts, c):
    """Signed distance from point c to a convex CCW polygon and the feature that realises it:
    ('face', i) or ('vertex', k); ties and near-boundary points return None."""
    m = len(verts)
    seps = []
    for i in range(m):
        e = verts[(i + 1) % m] - verts[i]
        n = np.array([e[1], -e[0]]) / np.hypot(*e)
        seps.append(np.dot(n, c - verts[i]))
    seps = np.array(seps)
    if seps.max() < -1e-7:
        order = np.sort(seps)
        if order[-1] - order[-2] < 1e-7:
            return None
        return seps.max(), ("face", int(np.argmax(seps)))
    if seps.max() < 1e-6:
        return None
    best = None
    for i in range(m):
        a, b = verts[i], verts[(i + 1) % m]
        e = b - a
        t = np.dot(c - a, e) / np.dot(e, e)
        tc = min(max(t, 0.0), 1.0)
        d = np.linalg.norm(c - (a + tc * e))
        if tc == 0.0:
            feat = ("vertex", i)
        elif tc == 1.0:
            feat = ("vertex", (i + 1) % m)
        else:
            if min(t, 1 - t) < 1e-7:
                return None
            feat = ("face", i)
        if best is None or d < best[0] - 1e-9:
            best = (d, feat)
        elif abs(d - best[0]) <= 1e-9 and feat != best[1]:
            # the same vertex reached from both of its faces is not a tie
            if not (feat[0] == "vertex" and best[1] == feat):
                return None
    return best


def test_polygon_circle_against_closest_feature(oracle64):
    rng = np.random.default_rng(401)
    polys = _hull_shapes(oracle64, rng, 300)
    checked = {"face": 0, "vertex": 0, "inside": 0, "none": 0}
    for poly in polys:
        verts, normals = _verts(poly)
        for _ in range(6):
            r = rng.uniform(0.1, 0.6)
            cp = rng.uniform(-0.3, 0.3, 2) if rng.random() < 0.3 else np.zeros(2)
            circ = _circle_shape(oracle64, r, cp)
            xa = _xf(rng)
            # circle centre near the polygon: inside, across faces and around corners
            c_local = verts[rng.integers(len(verts))] * rng.uniform(0.3, 1.6) + rng.uniform(-0.3, 0.3, 2)
            c_world = _apply(xa, c_local[None])[0]
            ab = rng.uniform(-np.pi, np.pi)
            xb = np.array([0, 0, np.sin(ab), np.cos(ab)])
            xb[:2] = c_world - _rot(xb) @ cp
            feat = _closest_feature(verts, _inv_apply(xa, (_apply(xb, cp[None]))[0][None])[0])
            if feat is None:
                continue
            sd, (kind, idx) = feat
            radius = SKIN + r
            if abs(sd - radius) < 1e-9:
                continue
            m = oracle64.collide(np.array([poly]), xa, circ, xb)[0]
            if sd > radius:
                assert m["point_count"] == 0, (sd, radius)
                checked["none"] += 1
                continue
            assert m["point_count"] == 1 and m["type"] == E_FACE_A
            assert m["points"][0]["id"] == 0
            assert np.allclose(m["points"][0]["local_point"], cp, atol=0)
            q = _inv_apply(xa, _apply(xb, cp[None]))[0]
            if kind == "face":
                i = idx
                assert np.allclose(m["local_normal"], normals[i], atol=1e-12)
                assert np.allclose(m["local_point"], 0.5 * (verts[i] + verts[(i + 1) % len(verts)]), atol=1e-12)
                checked["inside" if sd < 0 else "face"] += 1
            else:
                v = verts[idx]
                assert np.allclose(m["local_point"], v, atol=0)
                assert np.allclose(m["local_normal"], (q - v) / np.linalg.norm(q - v), atol=1e-9)
                checked["vertex"] += 1
            wm = oracle64.world_manifold(m[None], xa, [SKIN], xb, [r])[0]
            assert abs(wm[6] - (sd - radius)) < 1e-9
            # the normal points from the polygon towards the circle
            assert np.dot(wm[:2], c_world - _apply(xa, verts.mean(0)[None])[0]) > 0 or sd < 0
    assert min(checked.values()) > 40, checked


def test_edge_circle_regions_two_sided(oracle64):
    rng = np.random.default_rng(402)
    checked = {"A": 0, "B": 0, "AB": 0, "none": 0}
    for _ in range(1500):
        a = rng.uniform(-1, 1, 2)
        b = a + rng.uniform(0.3, 2.0) * np.array([np.cos(t := rng.uniform(0, 2 * np.pi)), np.sin(t)])
        edge = _edge_shape(oracle64, a, b)
        r = rng.uniform(0.1, 0.6)
        circ = _circle_shape(oracle64, r, np.zeros(2))
        xa = _xf(rng)
        q = a + rng.uniform(-0.5, 1.5) * (b - a) + rng.uniform(-0.9, 0.9, 2)
        xb = np.array([*_apply(xa, q[None])[0], 0.0, 1.0])
        e = b - a
        t = np.dot(q - a, e) / np.dot(e, e)
        if abs(t) < 1e-7 or abs(t - 1) < 1e-7:
            continue
        tc = min(max(t, 0.0), 1.0)
        d = np.linalg.norm(q - (a + tc * e))
        radius = SKIN + r
        if abs(d - radius) < 1e-9 or d < 1e-6:
            continue
        m = oracle64.collide(edge, xa, circ, xb)[0]
        if d > radius:
            assert m["point_count"] == 0
            checked["none"] += 1
            continue
        assert m["point_count"] == 1
        assert np.allclose(m["points"][0]["local_point"], 0, atol=0)
        if t < 0 or t > 1:
            v, idx = (a, 0) if t < 0 else (b, 1)
            assert m["type"] == E_CIRCLES
            assert np.allclose(m["local_point"], v, atol=0)
            assert m["points"][0]["id"] == _key(idx, 0, VERTEX, VERTEX)
            checked["A" if idx == 0 else "B"] += 1
        else:
            assert m["type"] == E_FACE_A
            n = np.array([e[1], -e[0]]) / np.hypot(*e)
            if np.dot(n, q - a) < 0:
                n = -n
            assert np.allclose(m["local_normal"], n, atol=1e-12)
            assert np.allclose(m["local_point"], a, atol=0)
            assert m["points"][0]["id"] == _key(0, 0, FACE, VERTEX)
            checked["AB"] += 1
        wm = oracle64.world_manifold(m[None], xa, [SKIN], xb, [r])[0]
        assert abs(wm[6] - (d - radius)) < 1e-9
    assert min(checked.values()) > 60, checked


def test_edge_circle_one_sided_neighbours(oracle64):
    """A one-sided edge collides only on its right side (normal (e.y, -e.x)), and a centre beyond
    vertex A that lies in the interior region of the previous edge belongs to that edge; beyond a
    convex corner (the previous edge turns away) the vertex region is admitted."""
    r = 0.3
    circ = _circle_shape(oracle64, r, np.zeros(2))
    a, b = np.array([0.0, 0.0]), np.array([2.0, 0.0])
    xa = np.array([0, 0, 0.0, 1.0])
    straight = _edge_shape(oracle64, a, b, v0=np.array([-2.0, 0.0]), v3=np.array([4.0, 0.0]))
    # normal (e.y, -e.x) = (0, -1): the right side of a -> b is below the edge
    below = oracle64.collide(straight, xa, circ, np.array([1.0, -0.2, 0, 1.0]))[0]
    above = oracle64.collide(straight, xa, circ, np.array([1.0, 0.2, 0, 1.0]))[0]
    assert below["point_count"] == 1 and above["point_count"] == 0
    assert np.allclose(below["local_normal"], [0, -1], atol=0)
    # beyond A, on the right side: the collinear previous edge's interior region owns the centre
    beyond_a = np.array([-0.1, -0.2, 0, 1.0])
    assert oracle64.collide(straight, xa, circ, beyond_a)[0]["point_count"] == 0
    assert oracle64.collide(_edge_shape(oracle64, a, b), xa, circ, beyond_a)[0]["point_count"] == 1
    # previous edge turning up and away (v0 above the line, convex corner at A): vertex region admitted
    corner = _edge_shape(oracle64, a, b, v0=np.array([-1.0, 1.5]), v3=np.array([4.0, 0.0]))
    m = oracle64.collide(corner, xa, circ, beyond_a)[0]
    assert m["point_count"] == 1 and m["type"] == E_CIRCLES and m["points"][0]["id"] == _key(0, 0, VERTEX, VERTEX)
    # the same beyond B with v3
    beyond_b = np.array([2.1, -0.2, 0, 1.0])
    assert oracle64.collide(straight, xa, circ, beyond_b)[0]["point_count"] == 0
    corner_b = _edge_shape(oracle64, a, b, v0=np.array([-2.0, 0.0]), v3=np.array([3.0, 1.5]))
    m = oracle64.collide(corner_b, xa, circ, beyond_b)[0]
    assert m["point_count"] == 1 and m["points"][0]["id"] == _key(1, 0, VERTEX, VERTEX)


def _axis_depths_edge(verts, a, b):
    """Separating-axis depths of polygon verts (edge frame) against segment a-b: the edge normal in
    both signs, then each polygon face (segment end points against the face)."""
    e = (b - a) / np.linalg.norm(b - a)
    n = np.array([e[1], -e[0]])
    edge_axes = [(np.min((verts - a) @ (s * n)), s * n) for s in (1.0, -1.0)]
    m = len(verts)
    poly = []
    for i in range(m):
        f = verts[(i + 1) % m] - verts[i]
        ni = np.array([f[1], -f[0]]) / np.hypot(*f)
        poly.append(min(np.dot(ni, a - verts[i]), np.dot(ni, b - verts[i])))
    return edge_axes, np.array(poly)


def test_edge_polygon_edge_axis_long_edge(oracle64):
    """A polygon over the middle of a long two-sided edge: the edge axis is chosen, the incident
    face is the most anti-parallel one, its end points are the clip points, no side plane cuts."""
    rng = np.random.default_rng(403)
    polys = _hull_shapes(oracle64, rng, 400)
    checked = {1: 0, 2: 0, "none": 0}
    for poly in polys:
        verts_l, normals_l = _verts(poly)
        L = 6.0
        a, b = np.array([-L, 0.0]), np.array([L, 0.0])
        edge = _edge_shape(oracle64, a, b)
        # polygon pose in the edge frame: random angle, resting near the line on either side
        side = 1.0 if rng.random() < 0.5 else -1.0
        if rng.random() < 0.5:
            ang = rng.uniform(-np.pi, np.pi)
        else:  # a face lying nearly flat on the edge (two clip points)
            f = int(rng.integers(len(verts_l)))
            ang = -side * np.pi / 2 - np.arctan2(normals_l[f][1], normals_l[f][0]) + rng.uniform(-0.02, 0.02)
        R = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
        v = verts_l @ R.T
        low = np.min(v[:, 1] * side)
        shift = np.array([rng.uniform(-2, 2), side * (rng.uniform(-0.03, 0.04) - low)])
        v = v + shift
        xa = _xf(rng)
        xrel = np.array([*shift, np.sin(ang), np.cos(ang)])
        xb = np.array([*_apply(xa, xrel[None, :2])[0], 0, 0])
        ab = np.arctan2(xa[2], xa[3]) + ang
        xb[2], xb[3] = np.sin(ab), np.cos(ab)
        v = _inv_apply(xa, _apply(xb, verts_l))  # recompute through the transforms actually passed
        edge_axes, poly_seps = _axis_depths_edge(v, a, b)
        E = max(edge_axes[0][0], edge_axes[1][0])
        P = poly_seps.max()
        radius = 2 * SKIN
        if abs(edge_axes[0][0] - edge_axes[1][0]) < 1e-6:
            continue
        if not (P - radius < 0.98 * (E - radius) + 0.001 - 1e-6):
            continue  # hysteresis picks the polygon axis or is near its threshold
        if min(abs(E - radius), abs(P - radius)) < 1e-9:
            continue
        m = oracle64.collide(edge, xa, np.array([poly]), xb)[0]
        if max(E, P) > radius:
            assert m["point_count"] == 0
            checked["none"] += 1
            continue
        nrm = edge_axes[0][1] if edge_axes[0][0] > edge_axes[1][0] else edge_axes[1][1]
        nv = normals_l @ _rot(np.array([0, 0, np.sin(ang), np.cos(ang)])).T
        dots = nv @ nrm
        order = np.sort(dots)
        if order[1] - order[0] < 1e-9:
            continue
        i1 = int(np.argmin(dots))
        i2 = (i1 + 1) % len(v)
        d1, d2 = np.dot(nrm, v[i1] - a), np.dot(nrm, v[i2] - a)
        if min(abs(d1 - radius), abs(d2 - radius)) < 1e-9:
            continue
        assert m["type"] == E_FACE_A
        assert np.allclose(m["local_normal"], nrm, atol=1e-12)
        kept = [(i, d) for i, d in ((i1, d1), (i2, d2)) if d <= radius]
        assert m["point_count"] == len(kept)
        ids = [int(m["points"][k]["id"]) for k in range(len(kept))]
        assert ids == [_key(0, i, FACE, VERTEX) for i, _ in kept]
        wm = oracle64.world_manifold(m[None], xa, [SKIN], xb, [SKIN])[0]
        assert abs(wm[6:6 + len(kept)].min() - (E - radius)) < 1e-9
        checked[len(kept)] += 1
    assert min(checked.values()) > 30, checked


def test_edge_polygon_polygon_axis_short_edge(oracle64):
    """A short edge lying along one face of a polygon, inside that face's extent: the polygon axis
    is chosen (face B), the edge's end points are the clip points (v2 first), ids are swapped."""
    rng = np.random.default_rng(404)
    polys = _hull_shapes(oracle64, rng, 300)
    checked = {1: 0, 2: 0}
    for poly in polys:
        verts_l, normals_l = _verts(poly)
        m_ = len(verts_l)
        face = int(rng.integers(m_))
        p1, p2 = verts_l[face], verts_l[(face + 1) % m_]
        if np.linalg.norm(p2 - p1) < 0.3:
            continue
        nf = normals_l[face]
        # segment end points in the polygon frame: inside the face's middle, slightly off it
        u1, u2 = rng.uniform(0.25, 0.75, 2)
        e1 = p1 + u1 * (p2 - p1) + nf * rng.uniform(-0.015, 0.025)
        e2 = p1 + u2 * (p2 - p1) + nf * rng.uniform(-0.015, 0.025)
        if np.linalg.norm(e2 - e1) < 0.05:
            continue
        xb = _xf(rng)
        wa, wb = _apply(xb, e1[None])[0], _apply(xb, e2[None])[0]
        xa = _xf(rng)
        a, b = _inv_apply(xa, wa[None])[0], _inv_apply(xa, wb[None])[0]
        edge = _edge_shape(oracle64, a, b)
        v = _inv_apply(xa, _apply(xb, verts_l))
        edge_axes, poly_seps = _axis_depths_edge(v, a, b)
        E = max(edge_axes[0][0], edge_axes[1][0])
        radius = 2 * SKIN
        srt = np.sort(poly_seps)
        if srt[-1] - srt[-2] < 1e-6 or int(np.argmax(poly_seps)) != face:
            continue
        P = srt[-1]
        if not (P - radius > 0.98 * (E - radius) + 0.001 + 1e-6) or P > radius:
            continue
        res = oracle64.collide(edge, xa, np.array([poly]), xb)[0]
        # end point depths against the reference face, in the polygon frame
        dv2 = np.dot(nf, _inv_apply(xb, _apply(xa, b[None]))[0] - p1)
        dv1 = np.dot(nf, _inv_apply(xb, _apply(xa, a[None]))[0] - p1)
        if min(abs(dv1 - radius), abs(dv2 - radius)) < 1e-9:
            continue
        assert res["type"] == E_FACE_B
        assert np.allclose(res["local_normal"], nf, atol=0) and np.allclose(res["local_point"], p1, atol=0)
        kept = [(k, d) for k, d in ((1, dv2), (0, dv1)) if d <= radius]
        assert res["point_count"] == len(kept)
        assert [int(res["points"][j]["id"]) for j in range(len(kept))] == \
            [_key(face, k, FACE, VERTEX) for k, _ in kept]
        if kept:
            wm = oracle64.world_manifold(res[None], xa, [SKIN], xb, [SKIN])[0]
            # face B: world separations are measured along the polygon face normal
            assert abs(wm[6:6 + len(kept)].min() - (min(d for _, d in kept) - radius)) < 1e-9
            checked[len(kept)] += 1
    assert min(checked.values()) > 30, checked


def _face_depths(P, Q):
    """For each face of convex CCW polygon P: the depth of Q's deepest vertex along its normal."""
    out = []
    for i in range(len(P)):
        f = P[(i + 1) % len(P)] - P[i]
        n = np.array([f[1], -f[0]]) / np.hypot(*f)
        out.append(np.min((Q - P[i]) @ n))
    return np.array(out)


def test_polygon_polygon_general_hulls(oracle64):
    """Random 3-8-gon hulls: reference face from the separating-axis depths and the flip rule,
    incident face as the most anti-parallel, clip points = its end points when both lie inside
    the reference face's skin-extended side planes; counts, ids and depths checked."""
    rng = np.random.default_rng(405)
    polys = _hull_shapes(oracle64, rng, 3000)
    checked = {"A": 0, "B": 0, "none": 0, 1: 0, 2: 0}
    tol = 0.1 * 0.005
    radius = 2 * SKIN
    for k in range(0, len(polys), 2):
        sa, sb = polys[k], polys[k + 1]
        la, na = _verts(sa)
        lb, nb = _verts(sb)
        xa = _xf(rng, 0.2)
        xb = _xf(rng, 0.2)
        u = rng.uniform(-np.pi, np.pi)
        reach = np.abs(la).max() + np.abs(lb).max()
        xb[:2] = xa[:2] + reach * rng.uniform(0.4, 0.9) * np.array([np.cos(u), np.sin(u)])
        va, vb = _apply(xa, la), _apply(xb, lb)
        da, db = _face_depths(va, vb), _face_depths(vb, va)
        sepA, sepB = da.max(), db.max()
        if min(abs(sepA - radius), abs(sepB - radius)) < 1e-9 or abs(sepB - (sepA + tol)) < 1e-9:
            continue
        res = oracle64.collide(np.array([sa]), xa, np.array([sb]), xb)[0]
        if sepA > radius or sepB > radius:
            assert res["point_count"] == 0
            checked["none"] += 1
            continue
        flip = sepB > sepA + tol
        ref_v, ref_d, inc_v, inc_l = (vb, db, va, la) if flip else (va, da, vb, lb)
        srt = np.sort(ref_d)
        if srt[-1] - srt[-2] < 1e-9:
            continue
        e1 = int(np.argmax(ref_d))
        p1, p2 = ref_v[e1], ref_v[(e1 + 1) % len(ref_v)]
        tng = (p2 - p1) / np.linalg.norm(p2 - p1)
        nrm = np.array([tng[1], -tng[0]])
        inc_n = np.array([[f[1], -f[0]] for f in (np.roll(inc_v, -1, 0) - inc_v)])
        inc_n /= np.linalg.norm(inc_n, axis=1)[:, None]
        dots = inc_n @ nrm
        ds = np.sort(dots)
        if ds[1] - ds[0] < 1e-9:
            continue
        i1 = int(np.argmin(dots))
        i2 = (i1 + 1) % len(inc_v)
        s1 = np.dot(tng, inc_v[i1] - p1)
        s2 = np.dot(tng, inc_v[i2] - p1)
        ext = np.linalg.norm(p2 - p1)
        inside = all(-radius + 1e-7 < s < ext + radius - 1e-7 for s in (s1, s2))
        if not inside:
            continue
        d1, d2 = np.dot(nrm, inc_v[i1] - p1), np.dot(nrm, inc_v[i2] - p1)
        if min(abs(d1 - radius), abs(d2 - radius)) < 1e-9:
            continue
        assert res["type"] == (E_FACE_B if flip else E_FACE_A)
        ref_n = (nb if flip else na)[e1]
        assert np.allclose(res["local_normal"], ref_n, atol=1e-12)
        kept = [(i, d) for i, d in ((i1, d1), (i2, d2)) if d <= radius]
        assert res["point_count"] == len(kept)
        want = [_key(i, e1, VERTEX, FACE) if flip else _key(e1, i, FACE, VERTEX) for i, _ in kept]
        assert [int(res["points"][j]["id"]) for j in range(len(kept))] == want
        for j, (i, _) in enumerate(kept):
            assert np.allclose(res["points"][j]["local_point"], inc_l[i], atol=1e-12)
        wm = oracle64.world_manifold(res[None], xa, [SKIN], xb, [SKIN])[0]
        assert abs(wm[6:6 + len(kept)].min() - (min(d for _, d in kept) - radius)) < 1e-9
        # the incident face's deeper end point realises the separating-axis depth of the reference face
        assert abs(min(d1, d2) - ref_d[e1]) < 1e-9
        checked["B" if flip else "A"] += 1
        checked[len(kept)] += 1
    assert min(checked.values()) > 25, checked
