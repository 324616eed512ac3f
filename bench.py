"""Benchmark: body-steps/s of the 65,536-body world on one MI355X.

One step is one B2world::step (collide, coloured island solve with 8 velocity / 3 position
iterations, synchronize + LBVH pair search) of the large-world scene: 65,536 dynamic bodies
(50% circles, 30% boxes, 20% convex polygons) in a static 400 x 400 m box, seeded. Inputs
are resident in HBM before timing. Prints one JSON line.

  python bench.py [--gpus 1] [--steps 600] [--warmup 10] [--bodies 65536] [--no-cpu]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "box2d-rs_amd"))

HBM_PEAK_GBS = 8000.0

# Roofline kernel: k_traverse, the LBVH pair search, the longest launch of a step (one launch,
# about a quarter millisecond on this workload). The solver's per-colour kernels take more time
# in total but each launch lasts 4-11 us, so an event pair around one of them mostly measures
# its dispatch; they run from a captured graph in the timed region.
# Algorithmic HBM bytes of one traversal launch (f32/i32 sizes, each read or written once): per
# query proxy its record -- fat AABB 16 B, body 4, body type 4, world 4, filter 8, shape header
# 16 -- and its pair count 4 B: 56 B; per pair found the partner's record for the filters (body,
# type, world, filter, shape header: 36 B), its sorted leaf box and proxy id (20 B) and the slot
# write (4 B): 60 B. Internal-node reads of the walk are not counted: they are the search's
# overhead above the algorithmic minimum.
QUERY_BYTES, PAIR_BYTES = 56, 60
UNIT_BYTES = {"traverse": QUERY_BYTES, "solve_position": 136, "solve_velocity": 224, "warm_start": 128,
              "init_constraints": 404}
UNIT_NAME = {"traverse": "proxy query", "solve_velocity": "constraint-iteration",
             "solve_position": "constraint-iteration", "warm_start": "constraint", "init_constraints": "constraint"}


def cpu_baseline(scene, seconds=15.0, max_steps=200):
    """The oracle (sequential C++ restatement of the reference step, 1 thread) on the same
    scene, timed on this host for a bounded number of steps from the initial state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import b2o
    oracle = b2o.Oracle("f32")
    w = oracle.world()
    w.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
    steps, elapsed = 0, 0.0
    while elapsed < seconds and steps < max_steps:
        elapsed += w.step(1.0 / 60.0, 8, 3, steps=2)
        steps += 2
    n = int(scene.meta["dynamic"])
    return {"value": n * steps / elapsed, "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": "first %d steps of the same %d-body scene, oracle f32 on %s" % (
                steps, n, platform.processor() or platform.machine())}


def measured_traffic(kernel):
    """HBM bytes per launch of the kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary for this workload (calibrated against a copy kernel), or None when there is none."""
    path = os.path.join(ROOT, "profiles", "traffic_k_%s_large_world_65536.json" % kernel)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--bodies", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--roofline-kernel", default="traverse", choices=sorted(UNIT_BYTES))
    ap.add_argument("--solver-graph", type=int, default=1, choices=(0, 1),
                    help="replay the solve's per-colour launches from a captured graph (default 1; profilers "
                         "that intercept graph capture need 0)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; a single world or a batch of worlds is stepped on one "
              "MI355X (--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import b2g
    from b2g import scenes

    scene = scenes.large_world(n=args.bodies, seed=1)
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)
    w.tune(b2g.TUNE_SOLVER_GRAPH, args.solver_graph)
    w.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
    dt = 1.0 / 60.0
    for _ in range(args.warmup):
        w.step(dt, 8, 3)
    w.body_states()  # synchronise
    # HIP events around every launch of the dominant kernel on the library's stream, over the
    # timed region
    w.time_kernel(args.roofline_kernel)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w.step(dt, 8, 3)
    w.body_states()
    t1 = time.perf_counter()
    k_ms, k_launches, k_items, k_aux = w.kernel_time()
    w.time_kernel("none")
    p = w.profile()
    elapsed = t1 - t0
    n = int(scene.meta["dynamic"])
    value = n * args.steps / elapsed
    out = {
        "metric": "body-steps/s (64k-body world)",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded large-world scene)",
        "config": {"workload": "large_world: %d dynamic bodies, 400x400 m box, dt 1/60, 8/3 iterations, "
                               "warm start on, sleep and continuous off" % n,
                   "bodies": n, "last_step": {k: p[k] for k in ("pairs", "touching", "colours")}},
    }
    bytes_per_launch = UNIT_BYTES[args.roofline_kernel] * k_items / max(k_launches, 1)
    if args.roofline_kernel == "traverse":
        bytes_per_launch += PAIR_BYTES * k_aux / max(k_launches, 1)
    avg_launch_s = k_ms / 1000.0 / max(k_launches, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9 if k_launches else 0.0
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.roofline_kernel),
                       "kernel": "k_" + args.roofline_kernel, "launches": k_launches,
                       "avg_launch_us": avg_launch_s * 1e6, "bytes_per_launch": bytes_per_launch,
                       "unit_bytes": UNIT_BYTES[args.roofline_kernel],
                       "items_per_launch": k_items / max(k_launches, 1), "item": UNIT_NAME[args.roofline_kernel],
                       "pairs_per_launch": k_aux / max(k_launches, 1)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
