"""Benchmark: body-steps/s of the 65,536-body world on one MI355X.

One step is one B2world::step (collide, coloured island solve with 8 velocity / 3 position
iterations, synchronize + LBVH pair search) of the large-world scene: 65,536 dynamic bodies
(50% circles, 30% boxes, 20% convex polygons) in a static 400 x 400 m box, seeded. Inputs
are resident in HBM before timing. Prints one JSON line.

  python bench.py [--gpus 1] [--steps 600] [--warmup 10] [--bodies 65536] [--no-cpu] [--no-extras]

The line's value covers the timed window (steps warmup+1 .. warmup+steps, named in config).
Beside it the line carries:
  full_run     all 600 steps of the same world from its initial state, timed as one block;
  roofline     k_solve_velocity (the kernel with the largest share of a step) over the window;
  cpu_baseline the oracle (one core) on the first steps of the same scene;
  batched      4096 worlds of 16-1024 bodies, 300 steps, against the same worlds one by one;
  adjoint      1024 worlds x 64 bodies: 120 taped steps and the reverse pass;
  stress       1,048,576 mixed bodies + 16,384 towers of 8 boxes, 200 steps.
--no-extras leaves out the last three.

Other workloads as the main line: --workload batched | stress | adjoint | chains (4096 hanging
chains of 16 links on revolute joints). --continuous runs the large world with continuous
physics (TOI) on.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "box2d-rs_amd"))

HBM_PEAK_GBS = 8000.0

# Roofline kernel (default): k_solve_velocity, one velocity iteration over the constraints of one
# colour: per constraint-iteration 224 B (the constraint's rows: bodies, point count, normal and
# tangent masses, anchors, bias, block matrix and its inverse, impulses; both bodies' velocities
# read and written). It is the kernel with the largest share of a step's time; its launches are
# short (a few microseconds), so it is bound by launch and dependent-load latency more than by
# bandwidth, which the fraction shows. In steps where small islands hold most constraints the
# velocity iterations run as one launch of k_solve_bins_velocity (all 8 iterations of every bin of
# islands); its algorithmic bytes are counted the same way (224 B per constraint-iteration), while
# its HBM traffic is far lower: it reads each constraint and body once and keeps them on chip
# (registers and LDS) across the iterations.
# k_collide (--roofline-kernel collide): algorithmic HBM bytes of one narrowphase launch (f32/i32 sizes, each read or written once): per
# candidate pair 120 B (pair, fixture and body records, the two transforms, the previous
# manifold's ids and impulses, the new manifold) plus both shapes: circle 12 B, polygon 12 + 16 n
# (n vertices and normals), edge 36 B. The shape bytes per pair are averaged over the contact
# list read back before and after the timed steps.
# k_traverse (--roofline-kernel traverse), the LBVH pair search: per query its packed leaf record
# (16 B), fat AABB (16 B), world (4 B) and pair count (4 B): 40 B; per pair found the partner's
# box (16 B, held in its parent's node record) and leaf record (16 B) and the slot write (4 B):
# 36 B. Internal-node reads of the walk are not counted: they are the search's overhead above
# the algorithmic minimum.
QUERY_BYTES, PAIR_BYTES = 40, 36
COLLIDE_PAIR_BYTES = 120
UNIT_BYTES = {"traverse": QUERY_BYTES, "collide": COLLIDE_PAIR_BYTES, "solve_position": 136, "solve_velocity": 224,
              "warm_start": 128, "init_constraints": 404}
UNIT_NAME = {"traverse": "proxy query", "collide": "candidate pair", "solve_velocity": "constraint-iteration",
             "solve_position": "constraint-iteration", "warm_start": "constraint", "init_constraints": "constraint"}


def shape_bytes(shape_defs):
    """Per fixture: the narrowphase's algorithmic bytes of its shape (circle 12, polygon 12 + 16 n,
    edge 36; a box is a 4-gon, a polygon's hull keeps all of the scene's points)."""
    t, n = shape_defs["type"], shape_defs["count"]
    out = np.full(len(shape_defs), 36, np.int64)
    out[t == 0] = 12
    out[t == 2] = 12 + 16 * n[t == 2]
    out[t == 4] = 12 + 16 * 4
    return out


def pair_shape_bytes(w, sbytes):
    """Mean shape bytes per candidate pair of the world's current contact list."""
    fab, _, _ = w.contacts()
    return float((sbytes[fab[:, 0]] + sbytes[fab[:, 1]]).mean()) if len(fab) else 0.0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count()


def cpu_baseline(scene, seconds=15.0, max_steps=200, joints=None, continuous=False):
    """The oracle (sequential C++ restatement of the reference step, 1 thread) on the same
    scene, timed on this host for a bounded number of steps from the initial state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import b2o
    oracle = b2o.Oracle("f32")
    w = oracle.world()
    w.set_flags(warm_starting=True, allow_sleep=False, block_solve=True,  # the GPU run's flags
                continuous=continuous)
    w.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
    if joints is not None:
        w.add_joints(joints)
    steps, elapsed = 0, 0.0
    while elapsed < seconds and steps < max_steps:
        elapsed += w.step(1.0 / 60.0, 8, 3, steps=2)
        steps += 2
    n = int(scene.meta["dynamic"])
    return {"value": n * steps / elapsed, "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": "first %d steps of the same %d-body scene, oracle f32, one thread" % (steps, n),
            "cpu_model": cpu_model(), "nproc": host_cores()}


def measured_traffic(name, window):
    """HBM bytes per launch of the named kernel (k_...) from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (calibrated against a copy kernel), when that summary covers
    the same step window (first, last); otherwise None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_large_world_65536.json" % name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if list(rec.get("window", [])) != list(window):
        return None
    return rec.get("bytes_per_launch")


def roofline(kernel, k_ms, k_launches, k_items, k_aux, window=None, pair_shape=None):
    bytes_per_launch = UNIT_BYTES[kernel] * k_items / max(k_launches, 1)
    if kernel == "traverse":
        bytes_per_launch += PAIR_BYTES * k_aux / max(k_launches, 1)
    if kernel == "collide":
        bytes_per_launch += sum(pair_shape) / len(pair_shape) * k_items / max(k_launches, 1)
    avg_launch_s = k_ms / 1000.0 / max(k_launches, 1)
    achieved = bytes_per_launch / avg_launch_s / 1e9 if k_launches else 0.0
    name = solver_kernel_name(kernel)
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(name, window) if window else None,
            "kernel": name, "launches": k_launches, "avg_launch_us": avg_launch_s * 1e6,
            "bytes_per_launch": bytes_per_launch, "unit_bytes": UNIT_BYTES[kernel],
            "items_per_launch": k_items / max(k_launches, 1), "item": UNIT_NAME[kernel]}
    if kernel == "traverse":
        out["pairs_per_launch"] = k_aux / max(k_launches, 1)
    if kernel == "collide":
        out["shape_bytes_per_pair"] = list(pair_shape)
    return out


def timed_steps(w, steps, kernel, dt=1.0 / 60.0):
    """K steps bracketed by a stream synchronisation on both sides, with the given kernel's
    launches event-timed ("none": no events); returns (seconds, kernel totals)."""
    w.body_states()
    w.time_kernel(kernel)
    t0 = time.perf_counter()
    for _ in range(steps):
        w.step(dt, 8, 3)
    w.body_states()
    t1 = time.perf_counter()
    totals = w.kernel_time()
    w.time_kernel("none")
    return t1 - t0, totals


def replay_kernel(make_world, warmup, steps, kernel):
    """The kernel's launch totals over a replay of the same window on a fresh copy of the world
    (the same warm-up steps, then the same steps with the kernel's launches event-timed). The
    headline window itself runs without events: an event pair around each of the solver's short
    launches serialises the stream and slows the whole step by about a fifth. The world, its
    scene and the steps are deterministic, so the replay launches the same kernels on the same
    data."""
    r = make_world()
    for _ in range(warmup):
        r.step(1.0 / 60.0, 8, 3)
    r.body_states()
    r.time_kernel(kernel)
    bin_steps = 0
    for _ in range(steps):
        r.step(1.0 / 60.0, 8, 3)
        bin_steps += r.profile()["island_bins"] > 0
    totals = r.kernel_time()
    r.time_kernel("none")
    r.close()
    LAST_REPLAY["bin_steps"], LAST_REPLAY["steps"] = bin_steps, steps
    return totals


# steps of the last replay whose solve ran the island bins (names the roofline's kernel)
LAST_REPLAY = {"bin_steps": 0, "steps": 0}


def solver_kernel_name(kernel):
    if kernel not in ("solve_velocity", "solve_position"):
        return "k_" + kernel
    b, n = LAST_REPLAY["bin_steps"], LAST_REPLAY["steps"]
    if b == 0:
        return "k_" + kernel
    if b == n:
        return "k_solve_bins_" + kernel.split("_")[1]
    return "k_%s + k_solve_bins_%s (bins in %d of %d steps)" % (kernel, kernel.split("_")[1], b, n)


def awake_fraction(w, body_defs):
    dyn = body_defs["type"] == 2
    return float((w.body_states()["awake"][dyn] != 0).mean())


def base_record(args, metric, value, unit, elapsed, steps, workload, extra):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": args.warmup,
            "ms_per_step": 1000.0 * elapsed / max(steps, 1), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scenes)",
            "config": dict({"workload": workload}, **extra)}


def large_world_flags(args):
    return dict(warm_starting=True, allow_sleep=args.sleep, continuous=args.continuous, block_solve=True)


def large_world_maker(args, b2g, scene):
    def make():
        w = b2g.World()
        w.set_flags(**large_world_flags(args))
        w.tune(b2g.TUNE_SOLVER_GRAPH, args.solver_graph)
        w.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
        return w
    return make


def full_run(args, b2g, scene, steps=600):
    """All `steps` steps of a fresh copy of the world from its initial state, timed as one block
    (inputs uploaded before the clock starts), with the roofline kernel over a replay of it."""
    make = large_world_maker(args, b2g, scene)
    w = make()
    elapsed, _ = timed_steps(w, steps, "none")
    k_ms, k_launches, k_items, k_aux = replay_kernel(make, 0, steps, args.roofline_kernel)
    p = w.profile()
    n = int(scene.meta["dynamic"])
    out = {"steps": steps, "value": n * steps / elapsed, "unit": "body-steps/s", "ms_per_step": 1000.0 * elapsed / steps,
           "last_step": {k: p[k] for k in ("pairs", "touching", "colours")}}
    if args.roofline_kernel != "collide":
        r = roofline(args.roofline_kernel, k_ms, k_launches, k_items, k_aux)
        out["roofline"] = {k: r[k] for k in ("kernel", "achieved", "frac", "launches", "avg_launch_us",
                                             "bytes_per_launch", "items_per_launch")}
    w.close()
    return out


def run_large_world(args, b2g, scenes):
    scene = scenes.large_world(n=args.bodies, seed=1)
    make = large_world_maker(args, b2g, scene)
    w = make()
    for _ in range(args.warmup):
        w.step(1.0 / 60.0, 8, 3)
    p0 = w.profile()
    steps = args.steps if args.steps is not None else 600
    events = 0
    if args.continuous:
        # TOI events over the timed steps, counted in an untimed replay of the same steps
        probe = b2g.World()
        probe.set_flags(**large_world_flags(args))
        probe.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
        for i in range(args.warmup + steps):
            probe.step(1.0 / 60.0, 8, 3)
            if i >= args.warmup:
                events += probe.profile()["toi_events"]
        del probe
    sbytes = shape_bytes(scene.shape_defs)
    pair_shape = [pair_shape_bytes(w, sbytes)] if args.roofline_kernel == "collide" else None
    elapsed, _ = timed_steps(w, steps, "none")
    if pair_shape:
        pair_shape.append(pair_shape_bytes(w, sbytes))
    p = w.profile()
    n = int(scene.meta["dynamic"])
    window = (args.warmup + 1, args.warmup + steps)
    k_ms, k_launches, k_items, k_aux = replay_kernel(make, args.warmup, steps, args.roofline_kernel)
    out = base_record(args, "body-steps/s (64k-body world)", n * steps / elapsed, "body-steps/s", elapsed, steps,
                      "large_world: %d dynamic bodies, 400x400 m box, dt 1/60, 8/3 iterations, warm start on, "
                      "%s, continuous %s" % (n, "sleep on" if args.sleep else "sleep off",
                                             "on" if args.continuous else "off"),
                      {"bodies": n, "window": {"first_step": window[0], "last_step": window[1]},
                       "touching": {"before_window": p0["touching"], "last_step": p["touching"]},
                       "last_step": {k: p[k] for k in ("pairs", "touching", "colours")}})
    if args.roofline_kernel in ("solve_velocity", "solve_position"):
        iters = 8 if args.roofline_kernel == "solve_velocity" else 3
        out["config"]["touching"]["window_mean"] = k_items / (iters * steps)
    if args.continuous:
        out["config"]["continuous"] = True
        out["config"]["toi_events"] = events
    if args.sleep:
        out["config"]["sleep"] = True
        out["config"]["awake_after"] = awake_fraction(w, scene.body_defs)
    out["roofline"] = roofline(args.roofline_kernel, k_ms, k_launches, k_items, k_aux,
                               window=window if args.bodies == 65536 else None, pair_shape=pair_shape)
    w.close()
    if not args.no_full_run and args.bodies == 65536 and not args.continuous:
        out["full_run"] = full_run(args, b2g, scene)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, continuous=args.continuous)
    return out


def run_batched(args, b2g, scenes):
    """4096 independent worlds (16-1024 bodies each, heaps and stacks) stepped together, against
    the same worlds stepped one at a time as separate worlds (a sample of them)."""
    from b2g import autotune
    batch = scenes.batched_worlds(4096)
    settings = {}

    def make():
        w = b2g.World()
        w.set_flags(warm_starting=True, allow_sleep=args.sleep, continuous=False, block_solve=True)
        w.tune(b2g.TUNE_SOLVER_GRAPH, args.solver_graph)
        w.add_batched(batch.body_defs, batch.body_world, batch.fixture_body, batch.shape_defs, batch.fixture_defs)
        if settings:
            autotune.apply(w, settings)
        return w
    # launch settings (threads per world by world size, colouring rounds) tuned on a warmed-up copy;
    # they do not change the results
    tuned = None
    if not args.no_autotune:
        t = make()
        for _ in range(args.warmup):
            t.step(1.0 / 60.0, 8, 3)
        tuned = autotune.tune(t)
        settings.update(tuned["settings"])
        t.close()
    w = make()
    for _ in range(args.warmup):
        w.step(1.0 / 60.0, 8, 3)
    steps = args.steps if args.steps is not None else 300
    sbytes = shape_bytes(batch.shape_defs)
    pair_shape = [pair_shape_bytes(w, sbytes)] if args.roofline_kernel == "collide" else None
    elapsed, _ = timed_steps(w, steps, "none")
    if pair_shape:
        pair_shape.append(pair_shape_bytes(w, sbytes))
    nw = batch.meta["worlds"]
    p = w.profile()
    out = base_record(args, "world-steps/s (4096 batched worlds)", nw * steps / elapsed, "world-steps/s", elapsed,
                      steps, "batched_worlds: 4096 worlds of 16-1024 bodies (heaps and box stacks), %d bodies"
                      % batch.meta["dynamic"], {"worlds": nw, "bodies": int(batch.meta["dynamic"]),
                                                "last_step": {k: p[k] for k in ("pairs", "touching", "colours")}})
    out["body_steps_per_s"] = batch.meta["dynamic"] * steps / elapsed
    if tuned:
        out["config"]["autotuned"] = tuned["settings"]
    if args.sleep:
        out["config"]["sleep"] = True
        out["config"]["awake_after"] = awake_fraction(w, batch.body_defs)
    # plain per-world stepping: a sample of the same worlds, each its own world, stepped in turn
    sample = list(range(0, nw, nw // 32)) if not args.no_plain else []
    per_world = []
    for wi in sample:
        sc = scenes.world_of(batch, wi)
        one = b2g.World()
        one.set_flags(warm_starting=True, allow_sleep=args.sleep, continuous=False, block_solve=True)
        one.add(sc.body_defs, sc.fixture_body, sc.shape_defs, sc.fixture_defs)
        for _ in range(args.warmup):
            one.step(1.0 / 60.0, 8, 3)
        t, _ = timed_steps(one, 20, "none")
        per_world.append(t / 20)
    if per_world:
        plain = 1.0 / (sum(per_world) / len(per_world))
        out["plain_per_world"] = {"value": plain, "unit": "world-steps/s", "sample_worlds": len(sample),
                                  "ratio": out["value"] / plain}
    w.close()
    k_ms, k_launches, k_items, k_aux = replay_kernel(make, args.warmup, steps, args.roofline_kernel)
    out["roofline"] = roofline(args.roofline_kernel, k_ms, k_launches, k_items, k_aux, pair_shape=pair_shape)
    return out


def run_chains(args, b2g, scenes):
    """4096 hanging chains of 16 links (the testbed's chain: revolute joints, neighbouring links
    overlapping with collide_connected off) swinging down side by side: 65,536 bodies, 65,536
    joints."""
    scene, joints = scenes.chains(4096, 16)

    def make():
        w = b2g.World()
        w.set_flags(warm_starting=True, allow_sleep=args.sleep, continuous=False, block_solve=True)
        w.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
        w.add_joints(joints)
        return w
    w = make()
    for _ in range(args.warmup):
        w.step(1.0 / 60.0, 8, 3)
    steps = args.steps if args.steps is not None else 300
    sbytes = shape_bytes(scene.shape_defs)
    pair_shape = [pair_shape_bytes(w, sbytes)] if args.roofline_kernel == "collide" else None
    elapsed, _ = timed_steps(w, steps, "none")
    k_ms, k_launches, k_items, k_aux = replay_kernel(make, args.warmup, steps, args.roofline_kernel)
    if pair_shape:
        pair_shape.append(pair_shape_bytes(w, sbytes))
    p = w.profile()
    n = int(scene.meta["dynamic"])
    out = base_record(args, "body-steps/s (4096 chains x 16 links)", n * steps / elapsed, "body-steps/s", elapsed,
                      steps, "chains: 4096 chains x 16 links (revolute joints), %d bodies, %d joints, dt 1/60, 8/3 "
                      "iterations, warm start on, %s" % (n, len(joints), "sleep on" if args.sleep else "sleep off"),
                      {"bodies": n, "joints": len(joints),
                       "last_step": {k: p[k] for k in ("pairs", "touching", "colours")}})
    out["roofline"] = roofline(args.roofline_kernel, k_ms, k_launches, k_items, k_aux, pair_shape=pair_shape)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, joints=joints)
    return out


def run_stress(args, b2g, scenes):
    """1,048,576 large-world bodies and 16,384 towers of 8 boxes in a 1600 x 1600 m box: body-steps/s,
    plus the pair search's pairs/s (traversal time), the narrowphase's manifolds/s (collide time)
    and the solver's body-iterations/s (solve phase time)."""
    scene = scenes.stress_world()
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=args.sleep, continuous=False, block_solve=True)
    w.tune(b2g.TUNE_SOLVER_GRAPH, args.solver_graph)
    w.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
    for _ in range(args.warmup):
        w.step(1.0 / 60.0, 8, 3)
    steps = args.steps if args.steps is not None else 200
    elapsed, _ = timed_steps(w, steps, "none")
    n = int(scene.meta["dynamic"])
    p = w.profile()
    # the phase rates from steps after the timed window, each phase's launches event-timed
    _, (k_ms, k_launches, k_items, k_aux) = timed_steps(w, 10, "traverse")
    _, (c_ms, c_launches, c_items, _) = timed_steps(w, 5, "collide")
    out = base_record(args, "body-steps/s (1.18M-body stress world)", n * steps / elapsed, "body-steps/s", elapsed,
                      steps, "stress_world: 1,048,576 mixed bodies + 16,384 towers x 8 boxes, 1600x1600 m box",
                      {"bodies": n, "last_step": {k: p[k] for k in ("pairs", "touching", "colours")}})
    if args.sleep:
        out["config"]["sleep"] = True
        out["config"]["awake_after"] = awake_fraction(w, scene.body_defs)
    out["pairs_per_step"] = k_aux / max(k_launches, 1)  # over the 10 steps after the window
    out["pairs_per_s"] = k_aux / (k_ms / 1000.0) if k_ms else None
    out["manifolds_per_s"] = c_items / (c_ms / 1000.0) if c_ms else None
    out["solver_body_iterations_per_s"] = n * (8 + 3) / (p["solve"] / 1000.0) if p["solve"] else None
    out["roofline"] = roofline("traverse", k_ms, k_launches, k_items, k_aux)
    return out


def run_adjoint(args, b2g, scenes):
    """1024 worlds x 64 bodies: 120 steps recorded on a tape (body-state checkpoints every 16
    steps), then the reverse pass for the loss sum_w |c_body0(w) - goal_w|^2."""
    import numpy as np
    batch = scenes.differentiable_worlds(1024, 64)
    w = b2g.World()
    w.set_flags(warm_starting=True, allow_sleep=False, continuous=False, block_solve=True)  # a tape needs these
    w.add_batched(batch.body_defs, batch.body_world, batch.fixture_body, batch.shape_defs, batch.fixture_defs)
    w.body_states()
    steps = args.steps if args.steps is not None else 120
    w.tape_begin(16)
    t0 = time.perf_counter()
    for _ in range(steps):
        w.step(1.0 / 60.0, 8, 3)
    st = w.body_states()
    t1 = time.perf_counter()
    w.tape_end()
    cot = np.zeros((len(st), 6), np.float32)
    b0 = batch.meta["body0"]
    cot[b0, 0:2] = 2.0 * (st["c"][b0] - batch.meta["goal"])
    t2 = time.perf_counter()
    g, gf = w.tape_backward(cot)
    t3 = time.perf_counter()
    nw = batch.meta["worlds"]
    out = base_record(args, "adjoint world-steps/s (1024 worlds x 64 bodies)", nw * steps / ((t1 - t0) + (t3 - t2)),
                      "world-steps/s", (t1 - t0) + (t3 - t2), steps,
                      "differentiable_worlds: 1024 worlds x 64 boxes/circles, 20x20 m boxes, taped forward + "
                      "reverse pass, checkpoints every 16 steps", {"worlds": nw, "slots": w.tape_info()[1]})
    out["forward_ms"] = 1000.0 * (t1 - t0)
    out["backward_ms"] = 1000.0 * (t3 - t2)
    out["grad_norm"] = float(np.linalg.norm(g[:, 3:6]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 600; batched 300, stress 40, "
                                                            "adjoint 120)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--bodies", type=int, default=65536)
    ap.add_argument("--sleep", action="store_true",
                    help="allow_sleep on (large_world, batched): islands at rest sleep and are skipped")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="large_world: leave out the batched, adjoint and stress summaries")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched: leave out the one-world-at-a-time comparison")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched: launch settings at their defaults instead of tuned on a copy of the batch")
    ap.add_argument("--no-full-run", action="store_true",
                    help="large_world: leave out the 600-step run from the initial state")
    ap.add_argument("--continuous", action="store_true",
                    help="large_world with continuous physics (the TOI pass) on, the reference's default")
    ap.add_argument("--workload", default="large_world", choices=("large_world", "batched", "stress", "adjoint", "chains"))
    ap.add_argument("--roofline-kernel", default="solve_velocity", choices=sorted(UNIT_BYTES))
    ap.add_argument("--solver-graph", type=int, default=0, choices=(0, 1),
                    help="replay the solve's per-colour launches from a captured graph (default 0: direct "
                         "launches are faster on MI355X; timing a solve kernel launches the solve directly)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; a single world or a batch of worlds is stepped on one "
              "MI355X (--gpus %d is not supported)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import b2g
    from b2g import scenes
    run = {"large_world": run_large_world, "batched": run_batched, "stress": run_stress,
           "adjoint": run_adjoint, "chains": run_chains}[args.workload]
    out = run(args, b2g, scenes)
    if args.workload == "large_world" and not args.no_extras:
        out.update(extras(args, b2g, scenes))
    print(json.dumps(out))


def extras(args, b2g, scenes):
    """Summaries of the batched, adjoint and stress workloads at their own default sizes and
    step counts (each timed as its own block, after its own warm-up)."""
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup, sub.roofline_kernel = None, 10, "solve_velocity"
    out = {}
    b = run_batched(sub, b2g, scenes)
    out["batched"] = {"metric": b["metric"], "value": b["value"], "unit": b["unit"], "steps": b["steps"],
                      "ms_per_step": b["ms_per_step"], "worlds": b["config"]["worlds"],
                      "autotuned": b["config"].get("autotuned"),
                      "bodies": b["config"]["bodies"], "body_steps_per_s": b["body_steps_per_s"],
                      "plain_per_world": b.get("plain_per_world"),
                      "roofline": {k: b["roofline"][k] for k in ("kernel", "achieved", "frac", "avg_launch_us")}}
    a = run_adjoint(sub, b2g, scenes)
    out["adjoint"] = {"metric": a["metric"], "value": a["value"], "unit": a["unit"], "steps": a["steps"],
                      "forward_ms": a["forward_ms"], "backward_ms": a["backward_ms"],
                      "worlds": a["config"]["worlds"], "recorded_contacts": a["config"]["slots"]}
    sub.roofline_kernel = "traverse"
    st = run_stress(sub, b2g, scenes)
    out["stress"] = {"metric": st["metric"], "value": st["value"], "unit": st["unit"], "steps": st["steps"],
                     "ms_per_step": st["ms_per_step"], "bodies": st["config"]["bodies"],
                     "pairs_per_step": st["pairs_per_step"], "pairs_per_s": st["pairs_per_s"],
                     "manifolds_per_s": st["manifolds_per_s"],
                     "solver_body_iterations_per_s": st["solver_body_iterations_per_s"],
                     "last_step": st["config"]["last_step"]}
    return out


if __name__ == "__main__":
    main()
