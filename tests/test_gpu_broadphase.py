"""Broadphase parity on the GPU.

* synchronize_fixtures + move_proxy (src/private/dynamics/b2_fixture.rs:147-173,
  src/private/collision/b2_dynamic_tree.rs:109-168): from the oracle's state after a step and
  its fat AABBs before that step's synchronisation, the GPU's new fat AABBs equal the oracle's
  bit for bit.
* The LBVH pair set equals brute-force fat-AABB overlap under adversarial inputs (duplicate
  boxes, identical centres, clusters, huge boxes, a single proxy).
* The device rotation (B2Rot::set) agrees with glibc sinf/cosf to one ulp (and exactly on
  most angles).
"""
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def test_synchronize_matches_oracle(oracle):
    import b2g
    from b2g import scenes
    sc = scenes.heap_world(1200, seed=21, half_extent=14.0)
    ow = oracle.world()
    ow.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)  # the oracle's
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    moved_total = 0
    for k in range(1, 41):
        fat_prev, _ = ow.proxies()
        ow.step(DT, 8, 3)
        if k % 8:
            continue
        st = ow.bodies()
        fat_next, _ = ow.proxies()
        w.set_body_states(st)
        w.set_fat_aabbs(fat_prev)
        w.step_phase(b2g.PHASE_FIND, DT, 8, 3)
        got = w.fat_aabbs()
        moved_total += int(np.any(fat_next != fat_prev, axis=1).sum())
        assert np.array_equal(got, fat_next), "step %d: %d proxies differ" % (
            k, int(np.any(got != fat_next, axis=1).sum()))
    assert moved_total > 100  # the rule was exercised both ways


def _brute(fat):
    n = len(fat)
    lo, hi = fat[:, :2], fat[:, 2:]
    out = []
    for i in range(n):
        ok = ~((lo[i + 1:, 0] - hi[i, 0] > 0) | (lo[i + 1:, 1] - hi[i, 1] > 0) |
               (lo[i, 0] - hi[i + 1:, 0] > 0) | (lo[i, 1] - hi[i + 1:, 1] > 0))
        for j in np.nonzero(ok)[0]:
            out.append((i, i + 1 + j))
    return np.array(out, np.int64).reshape(-1, 2)


@pytest.mark.parametrize("case", ["clusters", "duplicates", "huge", "single", "line"])
def test_lbvh_pairs_adversarial(case):
    import b2g
    rng = np.random.default_rng(zlib.crc32(case.encode()) % 1000)
    n = {"single": 1, "huge": 3000}.get(case, 5000)
    sd = np.zeros(n, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_CIRCLE
    sd["radius"] = 0.5
    bd = np.zeros(n, b2g.BODY_DEF)
    bd["type"] = b2g.B2_DYNAMIC_BODY
    bd["gravity_scale"] = 1
    fd = np.zeros(n, b2g.FIXTURE_DEF)
    fd["density"] = 1
    fd["category"], fd["mask"] = 1, 0xFFFF
    w = b2g.World()
    w.add(bd, np.arange(n, dtype=np.int32), sd, fd)
    if case == "clusters":
        c = rng.normal(0, 50, (20, 2))[rng.integers(0, 20, n)] + rng.normal(0, 1.0, (n, 2))
        half = rng.uniform(0.05, 0.6, (n, 2))
    elif case == "duplicates":
        c = np.repeat(rng.uniform(-20, 20, (n // 10, 2)), 10, axis=0)
        half = np.full((n, 2), 0.3)
    elif case == "huge":
        c = rng.uniform(-100, 100, (n, 2))
        half = rng.uniform(0.1, 0.5, (n, 2))
        half[:5] = 80.0
    elif case == "single":
        c = np.zeros((1, 2))
        half = np.ones((1, 2))
    else:  # all centres on one line with exactly touching boxes
        c = np.stack([np.arange(n) * 1.0, np.zeros(n)], 1)
        half = np.full((n, 2), 0.5)
    fat = np.concatenate([c - half, c + half], 1).astype(np.float32)
    got = w.broadphase_pairs(fat)
    ref = _brute(fat)
    g = np.unique(got[:, 0].astype(np.int64) << 32 | got[:, 1])
    r = np.unique(ref[:, 0] << 32 | ref[:, 1]) if len(ref) else np.zeros(0, np.int64)
    assert len(g) == len(got)
    assert np.array_equal(g, r)


def test_device_rotation_matches_glibc(oracle):
    """The transform synchronisation after a step recomputes xf.q from the angle on the device;
    with no contacts, velocity or gravity it must equal glibc's (sinf, cosf) of the angle."""
    import b2g
    rng = np.random.default_rng(3)
    n = 100000
    angles = np.concatenate([rng.uniform(-np.pi, np.pi, n // 2), rng.uniform(-200, 200, n // 2)]).astype(np.float32)
    bd = np.zeros(n, b2g.BODY_DEF)
    bd["type"] = b2g.B2_DYNAMIC_BODY
    bd["position"][:, 0] = np.arange(n) * 3.0
    bd["angle"] = angles
    sd = np.zeros(n, b2g.SHAPE_DEF)
    sd["type"] = b2g.SHAPE_CIRCLE
    sd["radius"] = 0.5
    fd = np.zeros(n, b2g.FIXTURE_DEF)
    fd["density"] = 1
    fd["category"], fd["mask"] = 1, 0xFFFF
    w = b2g.World(gravity=(0.0, 0.0))
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    w.add(bd, np.arange(n, dtype=np.int32), sd, fd)
    w.step(DT, 8, 3)
    q = w.body_states()["q"]
    ref = oracle.rot(angles)
    mismatch = int(np.any(q != ref, axis=1).sum())
    # glibc sinf/cosf are not correctly rounded in every case; the device's double-precision
    # sincos rounded once differs from them by one ulp on a few percent of angles
    ulp = np.spacing(np.abs(ref).astype(np.float32))
    assert np.all(np.abs(q - ref) <= ulp)
    assert mismatch <= n // 20, "%d of %d rotations differ from glibc" % (mismatch, n)


def _filtered_scene(seed=31):
    """A heap of the large-world mix in a box with non-default fixture filters: categories of one
    of four bits, masks of random bit subsets, groups 0 (most), +-1 and +-2; a tenth of the bodies
    carry a second fixture (a small circle beside them, so same-body pairs arise) and a few are
    kinematic."""
    import b2g
    from b2g import scenes
    sc = scenes.heap_world(700, seed=seed, half_extent=14.0)
    rng = np.random.default_rng(seed)
    bd = sc.body_defs.copy()
    dyn = np.nonzero(bd["type"] == 2)[0]
    kin = rng.choice(dyn, 12, replace=False)
    bd["type"][kin] = b2g.B2_KINEMATIC_BODY
    bd["linear_velocity"][kin] = rng.normal(0, 1, (len(kin), 2)).astype(np.float32)
    extra = rng.choice(dyn, len(dyn) // 10, replace=False)
    esd = np.zeros(len(extra), b2g.SHAPE_DEF)
    esd["type"] = b2g.SHAPE_CIRCLE
    esd["radius"] = 0.2
    esd["params"][:, :2] = np.array([0.35, 0.0], np.float32)
    efd = np.zeros(len(extra), b2g.FIXTURE_DEF)
    efd["density"], efd["friction"] = 1.0, 0.6
    sd = np.concatenate([sc.shape_defs, esd])
    fd = np.concatenate([sc.fixture_defs, efd])
    fb = np.concatenate([np.asarray(sc.fixture_body), extra.astype(np.int32)])
    nf = len(fd)
    fd["category"] = (1 << rng.integers(0, 4, nf)).astype(np.uint16)
    fd["mask"] = np.where(rng.random(nf) < 0.3, 0xFFFF, rng.integers(0, 16, nf)).astype(np.uint16)
    fd["group"] = np.where(rng.random(nf) < 0.8, 0, rng.choice([-2, -1, 1, 2], nf)).astype(np.int16)
    return bd, fb, sd, fd


def test_filters_and_registry_orientation_match_oracle(oracle):
    """add_pair's filters (src/private/dynamics/b2_contact_manager.rs:178-300 with the default
    filter of src/private/dynamics/b2_world_callbacks.rs:6-18 and should_collide of
    src/private/dynamics/b2_body.rs:391-416: same body, no dynamic body, group, category and mask)
    and the registry's fixture A/B orientation (src/private/dynamics/b2_contact_registers.rs:67-103):
    after the pair search of a step from the oracle's state and fat AABBs, the GPU's contacts are
    the oracle's, as ordered (fixture A, fixture B) pairs."""
    import b2g
    bd, fb, sd, fd = _filtered_scene()
    ow = oracle.world()
    ow.add(bd, fb, sd, fd)
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)  # the oracle's
    w.add(bd, fb, sd, fd)
    checked = 0
    for k in range(1, 31):
        fat_prev, _ = ow.proxies()
        ow.step(DT, 8, 3)
        if k % 10:
            continue
        ow.collide()  # the contact list of the next collide: overlapping and filter-passing pairs
        ofab, _, _, _ = ow.contacts()
        w.set_body_states(ow.bodies())
        w.set_fat_aabbs(fat_prev)
        w.step_phase(b2g.PHASE_FIND, DT, 8, 3)
        gfab, _, _ = w.contacts()
        og = set(map(tuple, ofab.tolist()))
        gg = set(map(tuple, gfab.tolist()))
        assert len(gg) == len(gfab)
        assert gg == og, "step %d: %d GPU-only, %d oracle-only pairs" % (k, len(gg - og), len(og - gg))
        checked += len(gg)
    # the filters and the registry were exercised: some overlapping pairs were filtered out, and
    # some pairs are in swapped (higher, lower) order
    fat, _ = ow.proxies()
    body = np.asarray(fb)
    overlap = oracle.brute_pairs(fat)
    keep = (body[overlap[:, 0]] != body[overlap[:, 1]]) & (
        (bd["type"][body[overlap[:, 0]]] == 2) | (bd["type"][body[overlap[:, 1]]] == 2))
    assert len(og) < int(keep.sum())
    assert any(a > b for a, b in og) and any(a < b for a, b in og)
    assert checked > 1000
