"""Parity at the full sizes of the benchmark workloads.

65,536-body large world (the step's headline workload):
  * the GPU pair search over the oracle's fat AABBs after a few oracle steps equals the exact
    overlap set (grid pair finder) and the oracle world's own contact set (built by the
    reference's dynamic tree); over the GPU's own fat AABBs after 300 steps (the pile in
    contact, about 300k pairs) it equals the exact overlap set;
  * the narrowphase over every candidate pair, at the oracle's transforms after a few steps and
    at the GPU's after 300 steps, is bit-exact against the oracle (counts, types, feature ids and
    geometry);
  * after 300 GPU steps, one solve of the whole touching set replayed by the oracle in the GPU's
    colour order agrees within 1e-5, and every colour class is conflict-free.
1.18M-body stress world (1,048,576 mixed bodies + 16,384 towers of 8 boxes):
  * over 40 steps states stay finite and inside the container;
  * the contact set after a step equals the exact overlap set of the GPU's fat AABBs;
  * the solve's colour classes share no dynamic body.
"""
import numpy as np
import pytest

from helpers import compare_manifolds, overlap_pairs, pair_keys
from test_gpu_solver import DT, solver_replay_check

pytestmark = pytest.mark.gpu

PARITY_FLAGS = dict(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)


@pytest.fixture(scope="module")
def large_after_steps(oracle):
    """The 65,536-body world after 4 oracle steps (about 0.6 s of oracle time per step)."""
    from b2g import scenes
    sc = scenes.large_world(n=65536, seed=1)
    ow = oracle.world()
    ow.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    for _ in range(4):
        ow.step(DT, 8, 3)
    ow.collide()  # drop the contacts whose fat AABBs stopped overlapping, as the next step would
    return sc, ow


def test_large_world_pairs_equal_overlap_and_oracle_sets(large_after_steps):
    import b2g
    sc, ow = large_after_steps
    fat, _ = ow.proxies()
    w = b2g.World()
    w.set_flags(**PARITY_FLAGS)
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    got = w.broadphase_pairs(fat)
    keys = pair_keys(got)
    assert len(keys) == len(got), "duplicate pairs"
    exact = overlap_pairs(fat, sc.fixture_body, sc.body_defs["type"])
    fab, _, _, _ = ow.contacts()
    print("large world after 4 oracle steps: %d candidate pairs" % len(keys), flush=True)
    assert len(keys) > 30000
    assert np.array_equal(keys, exact)
    assert np.array_equal(keys, pair_keys(fab))
    w.close()


def test_large_world_narrowphase_bit_exact(oracle, large_after_steps):
    """Every candidate pair of the oracle's contact list (A/B in registry order) collided at the
    oracle's transforms."""
    import b2g
    sc, ow = large_after_steps
    fab, _, _, _ = ow.contacts()
    st = ow.bodies()
    body = np.asarray(sc.fixture_body)
    xf = np.concatenate([st["p"], st["q"]], 1).astype(np.float32)
    xa, xb = xf[body[fab[:, 0]]], xf[body[fab[:, 1]]]
    w = b2g.World()
    w.set_flags(**PARITY_FLAGS)
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    got = w.collide_pairs(fab, xa, xb)
    shapes, _ = oracle.make_shapes(sc.shape_defs)
    ref = oracle.collide(shapes[fab[:, 0]], xa, shapes[fab[:, 1]], xb)
    bad = compare_manifolds(got, ref, rtol=0.0, atol=0.0)
    assert not bad, bad[:5]
    touching = int((ref["point_count"] > 0).sum())
    kinds = set(map(int, ref["type"][ref["point_count"] > 0]))
    print("large world narrowphase after 4 oracle steps: %d pairs, %d touching, types %s" % (len(fab), touching, kinds))
    assert touching > 500 and len(fab) > 30000
    w.close()


@pytest.fixture(scope="module")
def large_gpu_300():
    """The 65,536-body world after 300 GPU steps (the pile in contact)."""
    import b2g
    from b2g import scenes
    sc = scenes.large_world(n=65536, seed=1)
    w = b2g.World()
    w.set_flags(**PARITY_FLAGS)
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    for _ in range(300):
        w.step(DT, 8, 3)
    yield sc, w
    w.close()


def test_large_world_pile_pairs_and_narrowphase(oracle, large_gpu_300):
    """After 300 GPU steps: the pair search over the world's fat AABBs is the exact overlap set, and
    the narrowphase of every candidate pair at the world's transforms is bit-exact against the
    oracle's."""
    sc, w = large_gpu_300
    fat = w.fat_aabbs()
    keys = pair_keys(w.broadphase_pairs(fat))
    exact = overlap_pairs(fat, sc.fixture_body, sc.body_defs["type"])
    assert np.array_equal(keys, exact)
    fab, _, _ = w.contacts()
    assert np.array_equal(pair_keys(fab), exact)
    st = w.body_states()
    body = np.asarray(sc.fixture_body)
    xf = np.concatenate([st["p"], st["q"]], 1).astype(np.float32)
    xa, xb = xf[body[fab[:, 0]]], xf[body[fab[:, 1]]]
    got = w.collide_pairs(fab, xa, xb)
    shapes, _ = oracle.make_shapes(sc.shape_defs)
    ref = oracle.collide(shapes[fab[:, 0]], xa, shapes[fab[:, 1]], xb)
    bad = compare_manifolds(got, ref, rtol=0.0, atol=0.0)
    assert not bad, bad[:5]
    touching = int((ref["point_count"] > 0).sum())
    kinds = set(map(int, ref["type"][ref["point_count"] > 0]))
    print("large world pile: %d candidate pairs, %d touching, types %s" % (len(fab), touching, kinds), flush=True)
    assert len(fab) > 200000 and touching > 60000 and kinds == {0, 1, 2}


def _colours_conflict_free(order, offs, contact_bodies, body_type):
    """No dynamic body appears in two constraints of one colour."""
    for c in range(len(offs) - 1):
        cb = contact_bodies[order[offs[c]:offs[c + 1]]]
        b = cb.ravel()
        b = b[body_type[b] == 2]
        if len(np.unique(b)) != len(b):
            return False
    return True


def test_large_world_solver_replay_after_300_steps(oracle):
    from b2g import scenes
    sc = scenes.large_world(n=65536, seed=1)
    info = solver_replay_check(oracle, sc, 300)
    print("large world replay: %d constraints, %d colours, %d islands" % (info["constraints"], info["colours"],
                                                                          info["islands"]), flush=True)
    assert info["constraints"] > 60000
    pos = {int(k): i for i, k in enumerate(info["touch"])}
    order = np.array([pos[int(k)] for k in info["order"]])
    assert _colours_conflict_free(order, info["offsets"], info["contact_bodies"], sc.body_defs["type"])


def test_stress_world_40_steps():
    import b2g
    from b2g import scenes
    sc = scenes.stress_world()
    n_bodies = len(sc.body_defs)
    w = b2g.World()
    w.set_flags(**PARITY_FLAGS)
    w.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
    half = 800.0
    pairs = []
    for k in range(1, 41):
        if k == 40:
            w.step_phase(b2g.PHASE_COLLIDE, DT, 8, 3)
            w.step_phase(b2g.PHASE_SOLVE, DT, 8, 3)
            order_c, offs = w.solver_order()
            fab, flags, _ = w.contacts()
            w.step_phase(b2g.PHASE_FIND, DT, 8, 3)
        else:
            w.step(DT, 8, 3)
        pairs.append(w.profile()["pairs"])
        if k % 10 == 0:
            st = w.body_states()
            assert np.all(np.isfinite(st["c"])) and np.all(np.isfinite(st["v"])), "step %d" % k
            dyn = sc.body_defs["type"] == 2
            assert np.all(np.abs(st["c"][dyn]) < half), "step %d: a body left the container" % k
            print("stress step %d: %d candidate pairs" % (k, pairs[-1]), flush=True)
    # colour classes of the last solve share no dynamic body
    body = np.asarray(sc.fixture_body)
    touch = np.nonzero(flags & 1)[0]
    cb = np.stack([body[fab[touch, 0]], body[fab[touch, 1]]], 1)
    pos = np.full(len(fab), -1, np.int64)
    pos[touch] = np.arange(len(touch))
    order = pos[order_c]
    assert np.all(order >= 0) and len(order) == len(touch) > 100000
    assert _colours_conflict_free(order, offs, cb, sc.body_defs["type"])
    # the contact set after the step's pair search is the overlap set of the current fat AABBs
    fab2, _, _ = w.contacts()
    exact = overlap_pairs(w.fat_aabbs(), sc.fixture_body, sc.body_defs["type"])
    assert np.array_equal(pair_keys(fab2), exact)
    assert n_bodies == 1 + (1 << 20) + 16384 * 8
    w.close()
