"""Autotuner of the launch settings of a b2g world (no counterpart in the reference).

The settings it picks change how the step is launched, not what it computes: threads per world of
the per-world solve for each body-count bucket of a batch (B2G_TUNE_WORLD_THREADS_SMALL / MEDIUM /
LARGE: 64, 128, 256 or 512) and the number of colouring rounds launched before the first read-back
(B2G_TUNE_COLOR_ROUNDS: the colouring is the same for any batching of its rounds). Every setting
gives the same states bit for bit, so a world can be tuned while it steps, or a copy tuned and the
result applied to the original.

Method: coordinate descent, one setting at a time, each candidate timed over a few steps of the
world (device time of the solve's kernels for the threads per world, wall time of whole steps for
the colouring rounds), candidates interleaved over several passes so that the slow drift of the
workload (a pile settling) does not favour the first or the last.

    python -m b2g.autotune [--workload batched|adjoint] [--worlds 4096]

prints the chosen settings and the timing table as one JSON line.
"""
import argparse
import json
import sys
import time

from . import (TUNE_COLOR_ROUNDS, TUNE_WORLD_SOLVE, TUNE_WORLD_THREADS_LARGE, TUNE_WORLD_THREADS_MEDIUM,
               TUNE_WORLD_THREADS_SMALL)

THREADS = (64, 128, 256, 512)
ROUNDS = (4, 8, 16, 32)
BUCKETS = (("small", TUNE_WORLD_THREADS_SMALL, 64), ("medium", TUNE_WORLD_THREADS_MEDIUM, 128),
           ("large", TUNE_WORLD_THREADS_LARGE, 256))


def _solve_kernel_ms(world, steps, dt, vi, pi):
    """Device milliseconds per step of the per-world velocity and position solves."""
    total = 0.0
    for kernel in ("solve_velocity", "solve_position"):
        world.time_kernel(kernel)
        for _ in range(steps):
            world.step(dt, vi, pi)
        ms, _, _, _ = world.kernel_time()
        world.time_kernel("none")
        total += ms / steps
    return total


def _step_ms(world, steps, dt, vi, pi):
    world.body_states()
    t0 = time.perf_counter()
    for _ in range(steps):
        world.step(dt, vi, pi)
    world.body_states()
    return 1000.0 * (time.perf_counter() - t0) / steps


def tune(world, steps=6, passes=2, dt=1.0 / 60.0, velocity_iterations=8, position_iterations=3):
    """Tunes `world` in place (it steps while being measured) and returns
    {"settings": {...}, "timings": {...}}; the chosen settings are left applied."""
    settings = {"world_solve": 1, "color_rounds": 16}
    timings = {}
    world.tune(TUNE_WORLD_SOLVE, 1)
    for name, param, default in BUCKETS:
        world.tune(param, default)
        settings["threads_" + name] = default
    for name, param, _ in BUCKETS:
        acc = {t: 0.0 for t in THREADS}
        for _ in range(passes):
            for t in THREADS:
                world.tune(param, t)
                acc[t] += _solve_kernel_ms(world, steps, dt, velocity_iterations, position_iterations) / passes
        best = min(acc, key=acc.get)
        world.tune(param, best)
        settings["threads_" + name] = best
        timings["threads_" + name] = {str(t): round(v, 4) for t, v in acc.items()}
    acc = {r: 0.0 for r in ROUNDS}
    for _ in range(passes):
        for r in ROUNDS:
            world.tune(TUNE_COLOR_ROUNDS, r)
            acc[r] += _step_ms(world, steps, dt, velocity_iterations, position_iterations) / passes
    best = min(acc, key=acc.get)
    world.tune(TUNE_COLOR_ROUNDS, best)
    settings["color_rounds"] = best
    timings["color_rounds"] = {str(r): round(v, 4) for r, v in acc.items()}
    return {"settings": settings, "timings": timings, "unit": "ms per step"}


def apply(world, settings):
    world.tune(TUNE_WORLD_SOLVE, settings.get("world_solve", -1))
    for name, param, _ in BUCKETS:
        if "threads_" + name in settings:
            world.tune(param, settings["threads_" + name])
    if "color_rounds" in settings:
        world.tune(TUNE_COLOR_ROUNDS, settings["color_rounds"])


def main(argv=None):
    import b2g
    from b2g import scenes
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="batched", choices=("batched", "adjoint"))
    ap.add_argument("--worlds", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=60)
    args = ap.parse_args(argv)
    if args.workload == "batched":
        batch = scenes.batched_worlds(args.worlds or 4096)
    else:
        batch = scenes.differentiable_worlds(args.worlds or 1024, 64)
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    w.add_batched(batch.body_defs, batch.body_world, batch.fixture_body, batch.shape_defs, batch.fixture_defs)
    for _ in range(args.warmup):
        w.step(1.0 / 60.0, 8, 3)
    out = tune(w)
    out["workload"] = "%s: %d worlds, %d bodies, tuned after %d steps" % (
        args.workload, batch.meta["worlds"], int(batch.meta["dynamic"]), args.warmup)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
