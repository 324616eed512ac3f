"""The island solve (TUNE_ISLAND_SOLVE) against the per-colour launches.

Islands share no dynamic body, so solving whole islands in bins -- one workgroup per bin runs all
velocity and all position iterations of its islands, colour after colour, with the constraint in
registers and the bodies in an LDS table -- is the same Gauss-Seidel sequence as one launch per
colour and iteration. States and manifolds (impulses included) must agree bit for bit:
  * a heap and a scatter of boxes and circles (one large island beside many small ones; with 64
    threads per bin the large island keeps the per-colour launches in the same step as the
    bins), for every threads-per-bin setting, sleeping off and on;
  * the 65,536-body world over its first 40 steps (thousands of small islands, all binned);
  * batched worlds (the island solve forced over the per-world solve);
  * a taped run: the tape records the island solve's order, its frozen replay reproduces the
    forward run, and its gradients equal those of the per-colour run's tape.
The solve's colour-order replay by the oracle (test_gpu_solver.py) runs the default path, which
is this one for single worlds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)


def _run(scene, steps, sleep=False, **tune):
    import b2g
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=sleep, continuous=False, block_solve=True)
    for k, v in tune.items():
        w.tune(getattr(b2g, k), v)
    if hasattr(scene, "body_world"):
        w.add_batched(scene.body_defs, scene.body_world, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
    else:
        w.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
    touching = 0
    for _ in range(steps):
        w.step(DT, 8, 3)
        touching = max(touching, w.profile()["touching"])
    out = (w.body_states().tobytes(), w.contacts()[2].tobytes())
    w.close()
    return out, touching


def _heap_and_scatter():
    """A heap of 600 bodies (one large island once it settles) beside 400 bodies dropped apart on
    the ground (islands of one to a few constraints)."""
    from b2g import scenes
    heap = scenes.heap_world(600, seed=11, half_extent=10.0)
    spread = scenes.large_world(n=400, seed=3, half_extent=30.0)
    shifted = spread.body_defs.copy()
    shifted["position"][:, 0] += 45.0
    parts = scenes.concat(tuple(heap[:4]), (shifted, spread.fixture_body, spread.shape_defs, spread.fixture_defs))
    return scenes.Scene(*parts, {"name": "heap_and_scatter"})


@pytest.mark.parametrize("sleep", [False, True])
def test_island_solve_equals_colour_launches(sleep):
    sc = _heap_and_scatter()
    (ref, touching) = _run(sc, 120, sleep, TUNE_ISLAND_SOLVE=0)
    assert touching > 500
    for threads in (64, 128, 256, 512):
        got, _ = _run(sc, 120, sleep, TUNE_ISLAND_SOLVE=1, TUNE_ISLAND_THREADS=threads)
        assert got[0] == ref[0], "states differ with %d threads per bin" % threads
        assert got[1] == ref[1], "manifolds differ with %d threads per bin" % threads


def test_island_solve_is_the_default_for_single_worlds():
    sc = _heap_and_scatter()
    assert _run(sc, 30)[0] == _run(sc, 30, TUNE_ISLAND_SOLVE=1)[0]


def test_island_solve_large_world_40_steps():
    from b2g import scenes
    sc = scenes.large_world(n=65536, seed=1)
    ref, touching = _run(sc, 40, TUNE_ISLAND_SOLVE=0)
    got, _ = _run(sc, 40)
    print("large world, 40 steps: up to %d touching contacts" % touching, flush=True)
    assert touching > 10000
    assert got == ref


def test_island_solve_batched_worlds():
    from b2g import scenes
    batch = scenes.batched_worlds(n_worlds=48, config=3, lo=16, hi=300)
    ref, _ = _run(batch, 60, TUNE_WORLD_SOLVE=0, TUNE_ISLAND_SOLVE=0)
    assert _run(batch, 60, TUNE_WORLD_SOLVE=0, TUNE_ISLAND_SOLVE=1)[0] == ref
    assert _run(batch, 60, TUNE_WORLD_SOLVE=1, TUNE_ISLAND_SOLVE=1, TUNE_ISLAND_THREADS=128)[0] == ref


def test_island_solve_tape():
    import b2g
    from b2g import scenes
    batch = scenes.differentiable_worlds(8, 24)
    res = []
    for island in (0, 1):
        w = b2g.World()
        w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
        w.tune(b2g.TUNE_WORLD_SOLVE, 0)
        w.tune(b2g.TUNE_ISLAND_SOLVE, island)
        w.add_batched(batch.body_defs, batch.body_world, batch.fixture_body, batch.shape_defs, batch.fixture_defs)
        w.tape_begin(16)
        for _ in range(40):
            w.step(DT, 8, 3)
        w.tape_end()
        st = w.body_states()
        fwd = np.concatenate([st["c"], st["a"][:, None], st["v"], st["w"][:, None]], 1)
        rep = w.tape_replay()
        dyn = batch.body_defs["type"] == 2
        assert np.abs(rep[dyn] - fwd[dyn]).max() <= 1e-4
        cot = np.zeros((len(fwd), 6), np.float32)
        cot[dyn, 0] = 1.0
        cot[dyn, 4] = 0.5
        g, gf = w.tape_backward(cot)
        res.append((st.tobytes(), g, gf))
        w.close()
    assert res[0][0] == res[1][0]
    for a, b in ((res[0][1], res[1][1]), (res[0][2], res[1][2])):
        scale = max(1e-6, float(np.abs(a).max()))
        assert np.abs(a - b).max() <= 1e-5 * scale
