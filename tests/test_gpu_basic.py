"""First GPU checks: the reference's hello_world assertion, stateless narrowphase and
broadphase parity against the oracle, and shape/mass setup parity."""
import zlib

import numpy as np
import pytest

from helpers import PAIR_KINDS, compare_manifolds, extent, near_poses, random_shape_defs

pytestmark = pytest.mark.gpu


def test_hello_world_gpu():
    """tests/test.rs:99-101 on the GPU path, through the reference-style API with B2world::new's
    defaults (warm starting, sleeping and continuous physics on)."""
    import b2g
    world = b2g.B2world((0.0, -10.0))
    ground = world.create_body(b2g.B2bodyDef(position=(0.0, -10.0)))
    box = b2g.B2polygonShape()
    box.set_as_box(50.0, 10.0)
    ground.create_fixture_by_shape(box, 0.0)
    body = world.create_body(b2g.B2bodyDef(body_type=b2g.B2_DYNAMIC_BODY, position=(0.0, 4.0)))
    dyn = b2g.B2polygonShape()
    dyn.set_as_box(1.0, 1.0)
    body.create_fixture(b2g.B2fixtureDef(shape=dyn, density=1.0, friction=0.3))
    for _ in range(60):
        world.step(1.0 / 60.0, 6, 2)
    p = body.get_position()
    assert abs(p.x) < 0.01
    assert abs(p.y - 1.01) < 0.01
    assert abs(body.get_angle()) < 0.01


def _shape_world(sd_list):
    """A world whose fixtures are the given shapes, each on its own static body."""
    import b2g
    sd = np.concatenate(sd_list)
    n = len(sd)
    bd = np.zeros(n, b2g.BODY_DEF)
    bd["gravity_scale"] = 1
    fd = np.zeros(n, b2g.FIXTURE_DEF)
    fd["friction"] = 0.2
    fd["category"] = 1
    fd["mask"] = 0xFFFF
    w = b2g.World()
    w.add(bd, np.arange(n, dtype=np.int32), sd, fd)
    return w


def test_hello_world_c_abi_defaults_match_oracle(oracle):
    """tests/test.rs:26-102 through b2g_world_create's defaults (no set_flags call), which are
    B2world::new's (src/private/dynamics/b2_world.rs:37-43): each of the 60 steps agrees with the
    oracle world run with warm starting, sleeping and continuous physics on, and the box ends
    at rest at y = 1.01."""
    import b2g
    from b2g import scenes
    sc = scenes.hello_world()
    w = b2g.World()
    assert w.flags == {"warm_starting": True, "allow_sleep": True, "continuous": True, "block_solve": True}
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    ow = oracle.world()
    ow.set_flags(warm_starting=True, allow_sleep=True, block_solve=True, continuous=True)
    ow.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    for k in range(60):
        w.step(1.0 / 60.0, 6, 2)
        ow.step(1.0 / 60.0, 6, 2)
        g, o = w.body_states()[1], ow.bodies()[1]
        assert w.profile()["toi_events"] == ow.toi_events, "step %d" % k
        for f in ("c", "a", "v", "w"):
            np.testing.assert_allclose(g[f], o[f], rtol=1e-5, atol=1e-5, err_msg="step %d %s" % (k, f))
        assert g["awake"] == o["awake"], "step %d" % k
    assert abs(g["p"][1] - 1.01) < 0.01 and abs(g["p"][0]) < 0.01


@pytest.mark.parametrize("kinds", PAIR_KINDS, ids=["-".join(k) for k in PAIR_KINDS])
def test_gpu_collide_matches_oracle(oracle, kinds):
    import b2g
    rng = np.random.default_rng(1000 + (zlib.crc32("-".join(kinds).encode()) & 0xff))
    n = 3000
    sda = random_shape_defs(rng, kinds[0], n, b2g.SHAPE_DEF)
    sdb = random_shape_defs(rng, kinds[1], n, b2g.SHAPE_DEF)
    sa, st_a = oracle.make_shapes(sda)
    sb, st_b = oracle.make_shapes(sdb)
    keep = (st_a == 0) & (st_b == 0)
    sa, sb, sda, sdb = sa[keep], sb[keep], sda[keep], sdb[keep]
    m = len(sa)
    w = _shape_world([sda, sdb])
    xa, xb = near_poses(rng, extent(sda), extent(sdb), m, oracle)
    ref = oracle.collide(sa, xa, sb, xb)
    pairs = np.stack([np.arange(m), m + np.arange(m)], 1).astype(np.int32)
    got = w.collide_pairs(pairs, xa, xb)
    assert int((ref["point_count"] > 0).sum()) > m // 10
    bad = compare_manifolds(got, ref, rtol=1e-5, atol=1e-6)
    assert not bad, bad[:5]
    # the contract is 1e-5; the device code is expected to be bit-exact
    exact = compare_manifolds(got, ref, rtol=0.0, atol=0.0)
    assert len(exact) == 0, exact[:5]


def test_gpu_mass_matches_oracle(oracle):
    import b2g
    from b2g import scenes
    sc = scenes.large_world(n=4096, seed=3)
    w = b2g.World()
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    ow = oracle.world()
    ow.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    assert np.array_equal(w.mass(), ow.mass())
    gs, os_ = w.body_states(), ow.bodies()
    for f in ("p", "q", "c", "a", "v", "w"):
        assert np.array_equal(gs[f], os_[f]), f
    fat_o, _ = ow.proxies()
    assert np.array_equal(w.fat_aabbs(), fat_o)


def test_gpu_broadphase_matches_brute_force(oracle):
    import b2g
    from b2g import scenes
    sc = scenes.large_world(n=20000, seed=5)
    ow = oracle.world()
    ow.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    for _ in range(3):
        ow.step(1.0 / 60.0, 8, 3)
    fat, _ = ow.proxies()
    w = b2g.World()
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    got = w.broadphase_pairs(fat)
    # independent check: all fixture pairs whose fat AABBs overlap, minus what add_pair filters
    # (same body; both static)
    brute = oracle.brute_pairs(fat)
    body = np.asarray(sc.fixture_body)
    btype = sc.body_defs["type"][body]
    keep = (body[brute[:, 0]] != body[brute[:, 1]]) & ((btype[brute[:, 0]] == 2) | (btype[brute[:, 1]] == 2))
    brute = brute[keep]
    a = np.unique(got[:, 0].astype(np.int64) << 32 | got[:, 1])
    b = np.unique(brute[:, 0].astype(np.int64) << 32 | brute[:, 1])
    assert len(a) == len(got), "duplicate pairs"
    assert np.array_equal(a, b)
    # and the oracle world's own contact set (built by the reference's tree) is the same set
    ow.collide()
    fab, _, _, _ = ow.contacts()
    lo, hi = np.minimum(fab[:, 0], fab[:, 1]), np.maximum(fab[:, 0], fab[:, 1])
    c = np.unique(lo.astype(np.int64) << 32 | hi)
    assert np.array_equal(a, c)


def test_gpu_pyramid_tracks_oracle(oracle):
    """120 steps of the 20-row pyramid: the rows (started 0.25 m apart) settle about 5 m; the GPU
    trajectory stays within 10 cm of the oracle's (the two solve the same constraints in
    different Gauss-Seidel orders, and device sin/cos may differ by an ulp)."""
    import b2g
    from b2g import scenes
    sc = scenes.pyramid()
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)  # the oracle world's flags
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    ow = oracle.world()
    ow.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    for _ in range(120):
        w.step(1.0 / 60.0, 8, 3)
        ow.step(1.0 / 60.0, 8, 3)
    st, so = w.body_states(), ow.bodies()
    assert np.all(np.isfinite(st["c"]))
    assert np.max(np.abs(st["c"] - so["c"])) < 0.1
    assert abs(float(st[-1]["c"][1]) - float(so[-1]["c"][1])) < 0.05
    prof = w.profile()
    assert prof["touching"] > 300
