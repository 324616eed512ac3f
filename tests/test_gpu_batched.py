"""Batched worlds on the GPU: many independent worlds advance through the same kernels.
Bodies of different worlds never form pairs; the pair set equals the per-world brute force;
the coloured solve of the whole batch replays on the CPU within 1e-5."""
import numpy as np
import pytest

from test_gpu_solver import DT, _world, solver_replay_check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    from b2g import scenes
    return scenes.batched_worlds(n_worlds=48, config=3, lo=16, hi=300)


def test_no_cross_world_pairs_and_brute_force(oracle, batch):
    w = _world(batch)
    assert w.world_count == 48
    for _ in range(20):
        w.step(DT, 8, 3)
    fab, flags, man = w.contacts()
    body = np.asarray(batch.fixture_body)
    bw = batch.body_world
    assert np.all(bw[body[fab[:, 0]]] == bw[body[fab[:, 1]]])
    fat = w.fat_aabbs()
    got = w.broadphase_pairs(fat)
    brute = oracle.brute_pairs(fat)
    btype = batch.body_defs["type"][body]
    keep = ((body[brute[:, 0]] != body[brute[:, 1]]) & ((btype[brute[:, 0]] == 2) | (btype[brute[:, 1]] == 2)) &
            (bw[body[brute[:, 0]]] == bw[body[brute[:, 1]]]))
    brute = brute[keep]
    a = np.unique(got[:, 0].astype(np.int64) << 32 | got[:, 1])
    b = np.unique(brute[:, 0].astype(np.int64) << 32 | brute[:, 1])
    assert len(a) == len(got) and np.array_equal(a, b)
    # the current contact set is the same set
    lo, hi = np.minimum(fab[:, 0], fab[:, 1]), np.maximum(fab[:, 0], fab[:, 1])
    assert np.array_equal(np.unique(lo.astype(np.int64) << 32 | hi), a)


def test_batched_solver_replay(oracle, batch):
    info = solver_replay_check(oracle, batch, 15)
    assert info["islands"] >= 48
