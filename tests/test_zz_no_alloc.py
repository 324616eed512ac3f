"""b2g_collide_pairs_noalloc (the narrowphase over given pairs with its device staging in the
library's fixed scratch area) returns byte-identical manifolds to b2g_collide_pairs, on two
calls in a row (the second smaller, so stale scratch contents from the first must not leak),
rejects more than B2G_NOALLOC_MAX_PAIRS pairs and refuses a world that is not on the device yet.
"""
import numpy as np
import pytest

from helpers import PAIR_KINDS, compare_manifolds, extent, near_poses, random_shape_defs


def _world_and_pairs(oracle, rng, per_kind=400, n_pairs=None):
    import b2g
    sds, blocks = [], []
    base = 0
    for ka, kb in PAIR_KINDS:
        sda = random_shape_defs(rng, ka, per_kind, b2g.SHAPE_DEF)
        sdb = random_shape_defs(rng, kb, per_kind, b2g.SHAPE_DEF)
        _, st_a = oracle.make_shapes(sda)
        _, st_b = oracle.make_shapes(sdb)
        sda, sdb = sda[st_a == 0], sdb[st_b == 0]
        blocks.append((base, len(sda), base + len(sda), len(sdb)))
        sds += [sda, sdb]
        base += len(sda) + len(sdb)
    sd = np.concatenate(sds)
    n = len(sd)
    bd = np.zeros(n, b2g.BODY_DEF)
    fd = np.zeros(n, b2g.FIXTURE_DEF)
    fd["category"], fd["mask"] = 1, 0xFFFF
    w = b2g.World()
    w.add(bd, np.arange(n, dtype=np.int32), sd, fd)
    ext = extent(sd)
    per = -(-n_pairs // len(blocks))
    pairs = []
    for a0, na, b0, nb in blocks:
        pairs.append(np.stack([a0 + rng.integers(0, na, per), b0 + rng.integers(0, nb, per)], 1))
    pairs = np.concatenate(pairs)[:n_pairs].astype(np.int32)
    xa, xb = near_poses(rng, ext[pairs[:, 0]], ext[pairs[:, 1]], len(pairs), oracle)
    w.shape_defs = sd
    return w, pairs, xa, xb


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point(oracle):
    import b2g
    rng = np.random.default_rng(2024)
    nmax = b2g.NOALLOC_MAX_PAIRS
    w, pairs, xa, xb = _world_and_pairs(oracle, rng, n_pairs=nmax)
    print("noalloc: %d fixtures, %d pairs" % (w.fixture_count, len(pairs)), flush=True)
    # the world's bodies are not on the device until an entry point uploads them
    with pytest.raises(b2g.B2gError) as e:
        w.collide_pairs_noalloc(pairs[:10], xa[:10], xb[:10])
    assert e.value.status == b2g.ERR_INVALID_ARGUMENT
    shapes, _ = oracle.make_shapes(w.shape_defs)
    for n in (nmax, 5000):
        ref = w.collide_pairs(pairs[:n], xa[:n], xb[:n])
        got = w.collide_pairs_noalloc(pairs[:n], xa[:n], xb[:n])
        assert got.tobytes() == ref.tobytes(), "call with %d pairs differs" % n
        # and both are the oracle's manifolds (counts, types and ids exact)
        orc = oracle.collide(shapes[pairs[:n, 0]], xa[:n], shapes[pairs[:n, 1]], xb[:n])
        t = orc["point_count"] > 0
        print("n=%d: touching %d (oracle %d), types %s (oracle %s)" % (
            n, (ref["point_count"] > 0).sum(), t.sum(), np.unique(ref["type"][ref["point_count"] > 0]),
            np.unique(orc["type"][t])), flush=True)
        bad = compare_manifolds(ref, orc)
        assert not bad, [(i, r, pairs[i].tolist()) for i, r in bad[:5]]
    big = np.zeros((nmax + 1, 2), np.int32)
    xf = np.zeros((nmax + 1, 4), np.float32)
    with pytest.raises(b2g.B2gError) as e:
        w.collide_pairs_noalloc(big, xf, xf)
    assert e.value.status == b2g.ERR_CAPACITY
    # an empty call is a no-op
    assert len(w.collide_pairs_noalloc(pairs[:0], xa[:0], xb[:0])) == 0
    w.close()


@pytest.mark.gpu
def test_noalloc_on_a_second_world(oracle):
    """A second world on the same device shares the scratch area: results still match."""
    import b2g
    rng = np.random.default_rng(2025)
    w1, p1, a1, b1 = _world_and_pairs(oracle, rng, per_kind=100, n_pairs=3000)
    w2, p2, a2, b2 = _world_and_pairs(oracle, rng, per_kind=120, n_pairs=2000)
    r1 = w1.collide_pairs(p1, a1, b1)
    r2 = w2.collide_pairs(p2, a2, b2)
    assert w2.collide_pairs_noalloc(p2, a2, b2).tobytes() == r2.tobytes()
    assert w1.collide_pairs_noalloc(p1, a1, b1).tobytes() == r1.tobytes()
    w1.close()
    w2.close()
