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


def _batched_run(batch, steps, sleep=False, **tune):
    import b2g
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=sleep, continuous=False, block_solve=True)
    for k, v in tune.items():
        w.tune(getattr(b2g, k), v)
    w.add_batched(batch.body_defs, batch.body_world, batch.fixture_body, batch.shape_defs, batch.fixture_defs)
    for _ in range(steps):
        w.step(DT, 8, 3)
    out = w.body_states().tobytes()
    _, _, man = w.contacts()
    w.close()
    return out, man.tobytes()


@pytest.mark.parametrize("sleep", [False, True])
def test_world_solve_equals_colour_launches(batch, sleep):
    """The per-world solve (one workgroup runs all iterations of its world, colour after colour)
    and the per-colour launches solve the same constraints in the same colour order: states and
    manifolds (impulses included) agree bit for bit after 90 steps, for every threads-per-world
    setting, with sleeping off and on."""
    ref = _batched_run(batch, 90, sleep, TUNE_WORLD_SOLVE=0, TUNE_ISLAND_SOLVE=0)
    for threads in (64, 128, 256, 512):
        got = _batched_run(batch, 90, sleep, TUNE_WORLD_SOLVE=1, TUNE_WORLD_THREADS_SMALL=threads,
                           TUNE_WORLD_THREADS_MEDIUM=threads, TUNE_WORLD_THREADS_LARGE=threads)
        assert got[0] == ref[0], "states differ with %d threads per world" % threads
        assert got[1] == ref[1], "manifolds differ with %d threads per world" % threads
    # the island solve (bins of whole islands, which never cross worlds) gives the same bits
    for threads in (64, 256):
        assert _batched_run(batch, 90, sleep, TUNE_WORLD_SOLVE=0, TUNE_ISLAND_SOLVE=1,
                            TUNE_ISLAND_THREADS=threads) == ref, "island solve, %d threads per bin" % threads


def test_world_solve_is_the_default_for_batches(oracle, batch):
    """With the default settings a batch of 48 worlds takes the per-world solve, and its solve
    replays on the CPU within 1e-5 (test_batched_solver_replay runs the default path); the
    per-colour launches give the same bits."""
    assert _batched_run(batch, 30) == _batched_run(batch, 30, TUNE_WORLD_SOLVE=1)
