"""b2g -- MI355X-native ``B2world::step`` for box2d-rs worlds (Python binding).

Two layers over the C ABI of ``libb2g.so`` (declared in ``include/b2g.h``):

* :class:`World` -- bulk, array-oriented access: add bodies/fixtures from numpy structured
  arrays, step, step one phase, read body states, contacts, fat AABBs, solver order;
* the reference-style API -- :class:`B2world`, :class:`B2bodyDef`, :class:`B2fixtureDef`,
  :class:`B2polygonShape`, :class:`B2circleShape`, :class:`B2edgeShape`, :class:`B2body` --
  with the names and argument meaning of box2d-rs (src/b2_world.rs, src/b2_body.rs,
  src/b2_fixture.rs, src/shapes/*), so that code written against the reference ports by
  renaming the import.

The library is loaded eagerly; importing this package raises if ``libb2g.so`` has not been
built (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libb2g.so")

OK, ERR_INVALID_ARGUMENT, ERR_CAPACITY, ERR_DEVICE, ERR_DEGENERATE, ERR_UNSUPPORTED = range(6)
B2_STATIC_BODY, B2_KINEMATIC_BODY, B2_DYNAMIC_BODY = 0, 1, 2
SHAPE_CIRCLE, SHAPE_EDGE, SHAPE_POLYGON, SHAPE_BOX, SHAPE_POLYGON_EXACT = 0, 1, 2, 4, 5
NOALLOC_MAX_PAIRS = 65536  # B2G_NOALLOC_MAX_PAIRS
PHASE_COLLIDE, PHASE_SOLVE, PHASE_FIND = 0, 1, 2
# kernels b2g_world_time_kernel can time (B2G_KERNEL_*)
KERNELS = {"none": -1, "solve_velocity": 0, "solve_position": 1, "warm_start": 2, "init_constraints": 3,
           "traverse": 4, "collide": 5, "color_round": 6}
# tuning knobs (B2G_TUNE_*)
TUNE_COLOR_ROUNDS, TUNE_SOLVER_GRAPH, TUNE_COLOR_REUSE = 2, 3, 4
TUNE_WORLD_SOLVE, TUNE_WORLD_THREADS_SMALL, TUNE_WORLD_THREADS_MEDIUM, TUNE_WORLD_THREADS_LARGE = 5, 6, 7, 8
TUNE_ISLAND_SOLVE, TUNE_ISLAND_THREADS = 9, 10
TUNE_TOI_GLOBAL_ISLAND = 11

# Joint definition (B2G_JOINT_*: B2jointType values, src/b2_joint.rs:46-58)
JOINT_DISTANCE, JOINT_PRISMATIC, JOINT_REVOLUTE, JOINT_WELD = 1, 6, 8, 9
JOINT_FRICTION, JOINT_GEAR, JOINT_MOTOR, JOINT_MOUSE, JOINT_PULLEY, JOINT_WHEEL = 2, 3, 4, 5, 7, 10
JOINT_DEF = np.dtype([("type", "<i4"), ("body_a", "<i4"), ("body_b", "<i4"), ("collide_connected", "<i4"),
                      ("local_anchor_a", "<f4", 2), ("local_anchor_b", "<f4", 2), ("length", "<f4"),
                      ("min_length", "<f4"), ("max_length", "<f4"), ("stiffness", "<f4"), ("damping", "<f4"),
                      ("reference_angle", "<f4"), ("lower_angle", "<f4"), ("upper_angle", "<f4"),
                      ("max_motor_torque", "<f4"), ("motor_speed", "<f4"), ("enable_limit", "<i4"),
                      ("enable_motor", "<i4"), ("local_axis_a", "<f4", 2), ("max_force", "<f4"),
                      ("max_torque", "<f4"), ("linear_offset", "<f4", 2), ("angular_offset", "<f4"),
                      ("correction_factor", "<f4"), ("target", "<f4", 2), ("ground_anchor_a", "<f4", 2),
                      ("ground_anchor_b", "<f4", 2), ("length_a", "<f4"), ("length_b", "<f4"), ("ratio", "<f4"),
                      ("joint1", "<i4"), ("joint2", "<i4")])

BODY_DEF = np.dtype([("type", "<i4"), ("allow_sleep", "<i4"), ("awake", "<i4"), ("fixed_rotation", "<i4"),
                     ("position", "<f4", 2), ("angle", "<f4"), ("linear_velocity", "<f4", 2),
                     ("angular_velocity", "<f4"), ("linear_damping", "<f4"), ("angular_damping", "<f4"),
                     ("gravity_scale", "<f4"), ("bullet", "<i4")])
SHAPE_DEF = np.dtype([("type", "<i4"), ("count", "<i4"), ("flags", "<i4"), ("radius", "<f4"), ("params", "<f4", 36)])
FIXTURE_DEF = np.dtype([("friction", "<f4"), ("restitution", "<f4"), ("restitution_threshold", "<f4"),
                        ("density", "<f4"), ("is_sensor", "<i4"), ("category", "<u2"), ("mask", "<u2"),
                        ("group", "<i2"), ("pad0", "<i2"), ("pad1", "<i4")])
BODY_STATE = np.dtype([("p", "<f4", 2), ("q", "<f4", 2), ("c", "<f4", 2), ("a", "<f4"), ("v", "<f4", 2),
                       ("w", "<f4"), ("c0", "<f4", 2), ("a0", "<f4"), ("sleep_time", "<f4"), ("awake", "<f4"),
                       ("pad", "<f4")])
POINT = np.dtype([("local_point", "<f4", 2), ("normal_impulse", "<f4"), ("tangent_impulse", "<f4"), ("id", "<u4"),
                  ("pad", "<i4")])
MANIFOLD = np.dtype([("type", "<i4"), ("point_count", "<i4"), ("local_normal", "<f4", 2),
                     ("local_point", "<f4", 2), ("points", POINT, 2)])
assert BODY_DEF.itemsize == 56 and SHAPE_DEF.itemsize == 160 and FIXTURE_DEF.itemsize == 32
assert BODY_STATE.itemsize == 64 and MANIFOLD.itemsize == 72


class B2gError(RuntimeError):
    """A b2g C-ABI call returned a non-zero status."""

    def __init__(self, status, message):
        super().__init__("b2g error %d: %s" % (status, message))
        self.status = status


class _SolverState(ctypes.Structure):
    _fields_ = [("inv_dt0", ctypes.c_float), ("n_colours", ctypes.c_int32), ("scratch_colours", ctypes.c_int32),
                ("steps_since_scratch", ctypes.c_int32)]


class _Profile(ctypes.Structure):
    _fields_ = [("step", ctypes.c_float), ("collide", ctypes.c_float), ("solve", ctypes.c_float),
                ("broadphase", ctypes.c_float), ("pairs", ctypes.c_int32), ("touching", ctypes.c_int32),
                ("colours", ctypes.c_int32), ("islands", ctypes.c_int32), ("position_iterations", ctypes.c_int32),
                ("toi_events", ctypes.c_int32), ("island_bins", ctypes.c_int32), ("pad", ctypes.c_int32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libb2g.so is not built (%s); run `make -C box2d-rs_amd`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    P, I, F = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float
    I64 = ctypes.c_int64
    sig = {
        "b2g_last_error": ([], ctypes.c_char_p),
        "b2g_version": ([], ctypes.c_char_p),
        "b2g_world_create": ([F, F, ctypes.POINTER(P)], I),
        "b2g_world_destroy": ([P], I),
        "b2g_world_set_flags": ([P, I, I, I, I], I),
        "b2g_world_add": ([P, I, P, I, P, P, P, P], I),
        "b2g_world_add_batched": ([P, I, P, P, I, P, P, P, P], I),
        "b2g_world_world_count": ([P, P], I),
        "b2g_world_step": ([P, F, I, I], I),
        "b2g_world_step_phase": ([P, I, F, I, I], I),
        "b2g_world_body_count": ([P, P], I),
        "b2g_world_fixture_count": ([P, P], I),
        "b2g_world_contact_count": ([P, P], I),
        "b2g_world_get_body_states": ([P, P, I], I),
        "b2g_world_set_body_states": ([P, P, I], I),
        "b2g_world_get_mass": ([P, P, I], I),
        "b2g_world_get_fat_aabbs": ([P, P, I], I),
        "b2g_world_set_fat_aabbs": ([P, P, I], I),
        "b2g_world_get_contacts": ([P, I, P, P, P, P], I),
        "b2g_world_get_solver_order": ([P, I, P, P, I, P, P], I),
        "b2g_world_get_profile": ([P, P], I),
        "b2g_collide_pairs": ([P, I, P, P, P, P], I),
        "b2g_world_set_fixture_frictions": ([P, P, I], I),
        "b2g_world_set_bullet": ([P, I, I], I),
        "b2g_world_set_sub_stepping": ([P, I], I),
        "b2g_world_get_sub_stepping": ([P, P, P], I),
        "b2g_collide_pairs_noalloc": ([P, I, P, P, P, P], I),
        "b2g_broadphase_pairs": ([P, P, I64, P, P], I),
        "b2g_world_time_kernel": ([P, I], I),
        "b2g_world_tune": ([P, I, I], I),
        "b2g_tape_begin": ([P, I], I),
        "b2g_tape_end": ([P], I),
        "b2g_tape_free": ([P], I),
        "b2g_tape_info": ([P, P, P], I),
        "b2g_tape_replay": ([P, P], I),
        "b2g_tape_backward": ([P, P, P, P], I),
        "b2g_tape_export": ([P, I, I, P, P, P, I, P, P, P], I),
        "b2g_world_kernel_time": ([P, P, P, P, P], I),
        "b2g_world_get_shapes": ([P, P, I], I),
        "b2g_world_get_contact_colours": ([P, I, P, P], I),
        "b2g_world_set_contacts": ([P, I, P, P, P], I),
        "b2g_world_get_destroyed_contacts": ([P, I, P, P, P], I),
        "b2g_world_add_joints": ([P, I, P], I),
        "b2g_world_get_joint_impulses": ([P, P, I], I),
        "b2g_world_set_joint_target": ([P, I, ctypes.c_float, ctypes.c_float], I),
        "b2g_world_set_joint_impulses": ([P, P, I], I),
        "b2g_world_get_solver_state": ([P, P], I),
        "b2g_world_set_solver_state": ([P, P], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_lib = _load()
EXPORTED = tuple(sorted(n for n in dir(_lib) if n.startswith("b2g_")))


def version():
    return _lib.b2g_version().decode()


def _check(status):
    if status != OK:
        raise B2gError(status, _lib.b2g_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _as(arr, dtype):
    """Structured array in this library's layout (fields matched by name)."""
    if arr.dtype == dtype:
        return np.ascontiguousarray(arr)
    out = np.zeros(arr.shape, dtype)
    for name in dtype.names:
        if name in arr.dtype.names:
            out[name] = arr[name]
    return out


class World:
    """Array-oriented handle on one GPU world (one HIP stream)."""

    def __init__(self, gravity=(0.0, -10.0)):
        h = ctypes.c_void_p()
        _check(_lib.b2g_world_create(float(gravity[0]), float(gravity[1]), ctypes.byref(h)))
        self._h = h
        self.gravity = (float(gravity[0]), float(gravity[1]))
        # the library's defaults, those of B2world::new
        self.flags = {"warm_starting": True, "allow_sleep": True, "continuous": True, "block_solve": True}
        # host copies of the definitions handed over (the serde world writer reads them)
        self._body_defs, self._fixture_body, self._fixture_defs = [], [], []
        self._n_bodies, self._batched = 0, False

    def close(self):
        if getattr(self, "_h", None):
            _lib.b2g_world_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_flags(self, warm_starting=True, allow_sleep=True, continuous=True, block_solve=True):
        _check(_lib.b2g_world_set_flags(self._h, int(warm_starting), int(allow_sleep), int(continuous),
                                        int(block_solve)))
        self.flags = {"warm_starting": bool(warm_starting), "allow_sleep": bool(allow_sleep),
                      "continuous": bool(continuous), "block_solve": bool(block_solve)}

    def add(self, body_defs, fixture_body, shape_defs, fixture_defs):
        bd = _as(body_defs, BODY_DEF)
        sd = _as(shape_defs, SHAPE_DEF)
        fd = _as(fixture_defs, FIXTURE_DEF)
        fb = np.ascontiguousarray(fixture_body, np.int32)
        proxies = np.zeros(len(fb), np.int32)
        _check(_lib.b2g_world_add(self._h, len(bd), _ptr(bd), len(fb), _ptr(fb), _ptr(sd), _ptr(fd), _ptr(proxies)))
        self._body_defs.append(bd.copy())
        self._fixture_body.append(fb + self._n_bodies)
        self._fixture_defs.append(fd.copy())
        self._n_bodies += len(bd)
        return proxies

    def add_joints(self, joint_defs):
        """Joints of any kind (JOINT_DEF; bodies and a gear's joints are creation indices)."""
        jd = _as(joint_defs, JOINT_DEF)
        _check(_lib.b2g_world_add_joints(self._h, len(jd), _ptr(jd)))
        self._n_joints = getattr(self, "_n_joints", 0) + len(jd)
        self._joint_defs = getattr(self, "_joint_defs", []) + [jd.copy()]

    def joint_definitions(self):
        """The joint definitions handed to add_joints, in creation order."""
        parts = getattr(self, "_joint_defs", [])
        return np.concatenate(parts) if parts else np.zeros(0, JOINT_DEF)

    def set_joint_impulses(self, impulses):
        imp = np.ascontiguousarray(impulses, np.float32).reshape(-1, 5)
        _check(_lib.b2g_world_set_joint_impulses(self._h, _ptr(imp), len(imp)))

    @property
    def joint_count(self):
        return getattr(self, "_n_joints", 0)

    def set_joint_target(self, joint, target):
        """B2mouseJoint::set_target: a new target wakes body B."""
        _check(_lib.b2g_world_set_joint_target(self._h, int(joint), float(target[0]), float(target[1])))

    def joint_impulses(self):
        """Warm-start impulses per joint: (n, 5) (distance: impulse, lower, upper; revolute and
        prismatic: impulse.x, impulse.y, motor, lower, upper; weld: impulse x, y, z; wheel: impulse,
        spring, motor, lower, upper; friction and motor: linear x, y, angular; mouse: x, y; pulley and
        gear: impulse)."""
        out = np.zeros((max(self.joint_count, 1), 5), np.float32)
        _check(_lib.b2g_world_get_joint_impulses(self._h, _ptr(out), self.joint_count))
        return out[:self.joint_count]

    def definitions(self):
        """(body_defs, fixture_body, fixture_defs) handed to add(), concatenated (fixture_body
        holds world-wide body indices)."""
        cat = lambda parts, dt: np.concatenate(parts) if parts else np.zeros(0, dt)
        return (cat(self._body_defs, BODY_DEF), cat(self._fixture_body, np.int32), cat(self._fixture_defs, FIXTURE_DEF))

    def add_batched(self, body_defs, body_world, fixture_body, shape_defs, fixture_defs):
        """Bodies of many independent worlds (world id per body) stepped together."""
        bd = _as(body_defs, BODY_DEF)
        sd = _as(shape_defs, SHAPE_DEF)
        fd = _as(fixture_defs, FIXTURE_DEF)
        bw = np.ascontiguousarray(body_world, np.int32)
        fb = np.ascontiguousarray(fixture_body, np.int32)
        proxies = np.zeros(len(fb), np.int32)
        _check(_lib.b2g_world_add_batched(self._h, len(bd), _ptr(bd), _ptr(bw), len(fb), _ptr(fb), _ptr(sd), _ptr(fd),
                                          _ptr(proxies)))
        self._batched = True
        return proxies

    @property
    def world_count(self):
        return self._count(_lib.b2g_world_world_count)

    def step(self, dt=1.0 / 60.0, velocity_iterations=8, position_iterations=3):
        _check(_lib.b2g_world_step(self._h, float(dt), int(velocity_iterations), int(position_iterations)))

    def step_phase(self, phase, dt=1.0 / 60.0, velocity_iterations=8, position_iterations=3):
        _check(_lib.b2g_world_step_phase(self._h, int(phase), float(dt), int(velocity_iterations),
                                         int(position_iterations)))

    def _count(self, fn):
        out = ctypes.c_int32()
        _check(fn(self._h, ctypes.byref(out)))
        return out.value

    @property
    def body_count(self):
        return self._count(_lib.b2g_world_body_count)

    @property
    def fixture_count(self):
        return self._count(_lib.b2g_world_fixture_count)

    @property
    def contact_count(self):
        return self._count(_lib.b2g_world_contact_count)

    def body_states(self):
        out = np.zeros(self.body_count, BODY_STATE)
        _check(_lib.b2g_world_get_body_states(self._h, _ptr(out), len(out)))
        return out

    def set_body_states(self, states):
        st = _as(states, BODY_STATE)
        _check(_lib.b2g_world_set_body_states(self._h, _ptr(st), len(st)))

    def mass(self):
        out = np.zeros((self.body_count, 6), np.float32)
        _check(_lib.b2g_world_get_mass(self._h, _ptr(out), len(out)))
        return out

    def fat_aabbs(self):
        out = np.zeros((self.fixture_count, 4), np.float32)
        _check(_lib.b2g_world_get_fat_aabbs(self._h, _ptr(out), len(out)))
        return out

    def set_fat_aabbs(self, fat):
        fat = np.ascontiguousarray(fat, np.float32).reshape(-1, 4)
        _check(_lib.b2g_world_set_fat_aabbs(self._h, _ptr(fat), len(fat)))

    def contacts(self):
        n = self.contact_count
        fab = np.zeros((max(n, 1), 2), np.int32)
        flags = np.zeros(max(n, 1), np.int32)
        man = np.zeros(max(n, 1), MANIFOLD)
        cnt = ctypes.c_int32()
        _check(_lib.b2g_world_get_contacts(self._h, n, _ptr(fab), _ptr(flags), _ptr(man), ctypes.byref(cnt)))
        k = min(cnt.value, n)
        return fab[:k], flags[:k], man[:k]

    def shapes(self):
        """Every fixture's shape as the device holds it (SHAPE_DEF; polygons as SHAPE_POLYGON_EXACT)."""
        out = np.zeros(self.fixture_count, SHAPE_DEF)
        _check(_lib.b2g_world_get_shapes(self._h, _ptr(out), len(out)))
        return out

    def contact_colours(self):
        """Per contact of contacts(): the solve colour its predecessor had (-1: none)."""
        n = self.contact_count
        out = np.full(max(n, 1), -1, np.int32)
        cnt = ctypes.c_int32()
        _check(_lib.b2g_world_get_contact_colours(self._h, n, _ptr(out), ctypes.byref(cnt)))
        return out[:min(cnt.value, n)]

    def destroyed_contacts(self):
        """Touching contacts of the last solve that left the contact set (the next collide
        wakes their bodies): (fixtures_ab, manifolds)."""
        cnt = ctypes.c_int32()
        _check(_lib.b2g_world_get_destroyed_contacts(self._h, 0, None, None, ctypes.byref(cnt)))
        n = cnt.value
        fab = np.zeros((max(n, 1), 2), np.int32)
        man = np.zeros(max(n, 1), MANIFOLD)
        _check(_lib.b2g_world_get_destroyed_contacts(self._h, n, _ptr(fab), _ptr(man), ctypes.byref(cnt)))
        return fab[:n], man[:n]

    def set_contacts(self, fixtures_ab, manifolds, colours=None):
        """Rebuild the contact set from the current fat AABBs and give its contacts the given
        manifolds (with impulses) and colours, matched by fixtures (A, B)."""
        fab = np.ascontiguousarray(fixtures_ab, np.int32).reshape(-1, 2)
        man = np.ascontiguousarray(manifolds, MANIFOLD)
        col = None if colours is None else np.ascontiguousarray(colours, np.int32)
        if len(man) != len(fab) or (col is not None and len(col) != len(fab)):
            raise ValueError("set_contacts: fixtures, manifolds and colours differ in length")
        _check(_lib.b2g_world_set_contacts(self._h, len(fab), _ptr(fab), _ptr(man), _ptr(col)))

    def solver_state(self):
        st = _SolverState()
        _check(_lib.b2g_world_get_solver_state(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in _SolverState._fields_}

    def set_solver_state(self, state):
        st = _SolverState(**{name: state[name] for name, _ in _SolverState._fields_})
        _check(_lib.b2g_world_set_solver_state(self._h, ctypes.byref(st)))

    def solver_order(self):
        n = ctypes.c_int32()
        nc = ctypes.c_int32()
        _check(_lib.b2g_world_get_solver_order(self._h, 0, None, None, 0, ctypes.byref(n), ctypes.byref(nc)))
        order = np.zeros(max(n.value, 1), np.int32)
        offs = np.zeros(nc.value + 1, np.int32)
        _check(_lib.b2g_world_get_solver_order(self._h, n.value, _ptr(order), _ptr(offs), len(offs), ctypes.byref(n),
                                               ctypes.byref(nc)))
        return order[:n.value], offs

    def profile(self):
        p = _Profile()
        _check(_lib.b2g_world_get_profile(self._h, ctypes.byref(p)))
        return {name: getattr(p, name) for name, _ in _Profile._fields_ if name != "pad"}

    def tune(self, param, value):
        """Set a tuning knob (TUNE_*) of this world."""
        _check(_lib.b2g_world_tune(self._h, int(param), int(value)))

    def time_kernel(self, name):
        """Record HIP events around every launch of one kernel (a KERNELS name; "none" stops)."""
        _check(_lib.b2g_world_time_kernel(self._h, KERNELS[name]))

    def kernel_time(self):
        """(device ms summed over launches, launches, items, aux) of the timed kernel since the
        last call; resets the totals. aux is the traversal's pair count."""
        ms, launches, items, aux = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(_lib.b2g_world_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(launches), ctypes.byref(items),
                                          ctypes.byref(aux)))
        return ms.value, launches.value, items.value, aux.value

    # ---- adjoint -------------------------------------------------------------------------
    def tape_begin(self, checkpoint_every=16):
        """Record the following steps for tape_backward (body-state checkpoints every
        checkpoint_every steps)."""
        _check(_lib.b2g_tape_begin(self._h, int(checkpoint_every)))

    def tape_end(self):
        _check(_lib.b2g_tape_end(self._h))

    def tape_free(self):
        _check(_lib.b2g_tape_free(self._h))

    def tape_info(self):
        steps, slots = ctypes.c_int32(), ctypes.c_int64()
        _check(_lib.b2g_tape_info(self._h, ctypes.byref(steps), ctypes.byref(slots)))
        return steps.value, slots.value

    def tape_replay(self):
        """Final body state (n, 6: c.x, c.y, a, v.x, v.y, w) of the frozen replay of the tape."""
        out = np.zeros((self.body_count, 6), np.float32)
        _check(_lib.b2g_tape_replay(self._h, _ptr(out)))
        return out

    def tape_backward(self, cotangent):
        """Gradients of a loss with respect to the body state at the start of the recorded run
        and to the fixture frictions, given dL/d(final state) as an (n, 6) array."""
        cot = np.ascontiguousarray(cotangent, np.float32).reshape(self.body_count, 6)
        g = np.zeros_like(cot)
        gf = np.zeros(max(self.fixture_count, 1), np.float32)
        _check(_lib.b2g_tape_backward(self._h, _ptr(cot), _ptr(g), _ptr(gf)))
        return g, gf[:self.fixture_count]

    def tape_export(self, step):
        """One recorded step: slots (n, 10) int32, impulses (n, 8) float32, colour offsets and
        step parameters (dt, dt_ratio, velocity/position iterations, warm, block)."""
        n, nc = ctypes.c_int32(), ctypes.c_int32()
        params = np.zeros(6, np.float32)
        _check(_lib.b2g_tape_export(self._h, int(step), 0, None, None, None, 0, ctypes.byref(n), ctypes.byref(nc),
                                    _ptr(params)))
        slots = np.zeros((max(n.value, 1), 10), np.int32)
        imp = np.zeros((max(n.value, 1), 8), np.float32)
        offs = np.zeros(nc.value + 1, np.int32)
        _check(_lib.b2g_tape_export(self._h, int(step), n.value, _ptr(slots), _ptr(imp), _ptr(offs), len(offs),
                                    ctypes.byref(n), ctypes.byref(nc), _ptr(params)))
        return slots[:n.value], imp[:n.value], offs, params

    def set_fixture_frictions(self, friction):
        """B2fixture::set_friction for every fixture (array of the fixture count)."""
        fr = np.ascontiguousarray(friction, np.float32)
        _check(_lib.b2g_world_set_fixture_frictions(self._h, _ptr(fr), len(fr)))

    def set_bullet(self, body, flag=True):
        """B2body::set_bullet for one body (continuous collision also against dynamic bodies)."""
        _check(_lib.b2g_world_set_bullet(self._h, int(body), int(bool(flag))))

    def set_sub_stepping(self, flag=True):
        """B2world::set_sub_stepping: a TOI pass ends after its first solved event and the next
        steps resume it without solving until a pass finds no event."""
        _check(_lib.b2g_world_set_sub_stepping(self._h, int(bool(flag))))

    def sub_stepping(self):
        """(sub-stepping on, the last TOI pass finished)"""
        f, c = ctypes.c_int32(), ctypes.c_int32()
        _check(_lib.b2g_world_get_sub_stepping(self._h, ctypes.byref(f), ctypes.byref(c)))
        return bool(f.value), bool(c.value)

    def collide_pairs(self, pairs, xf_a, xf_b):
        pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        n = len(pairs)
        xa = np.ascontiguousarray(xf_a, np.float32).reshape(n, 4)
        xb = np.ascontiguousarray(xf_b, np.float32).reshape(n, 4)
        out = np.zeros(max(n, 1), MANIFOLD)
        _check(_lib.b2g_collide_pairs(self._h, n, _ptr(pairs), _ptr(xa), _ptr(xb), _ptr(out)))
        return out[:n]

    def collide_pairs_noalloc(self, pairs, xf_a, xf_b):
        """b2g_collide_pairs_noalloc: the same as collide_pairs with the device staging in the
        library's fixed scratch area (no device allocation); at most NOALLOC_MAX_PAIRS pairs."""
        pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        n = len(pairs)
        xa = np.ascontiguousarray(xf_a, np.float32).reshape(n, 4)
        xb = np.ascontiguousarray(xf_b, np.float32).reshape(n, 4)
        out = np.zeros(max(n, 1), MANIFOLD)
        _check(_lib.b2g_collide_pairs_noalloc(self._h, n, _ptr(pairs), _ptr(xa), _ptr(xb), _ptr(out)))
        return out[:n]

    def broadphase_pairs(self, fat):
        fat = np.ascontiguousarray(fat, np.float32).reshape(-1, 4)
        cnt = ctypes.c_int64()
        _check(_lib.b2g_broadphase_pairs(self._h, _ptr(fat), 0, None, ctypes.byref(cnt)))
        out = np.zeros((max(cnt.value, 1), 2), np.int32)
        _check(_lib.b2g_broadphase_pairs(self._h, _ptr(fat), cnt.value, _ptr(out), ctypes.byref(cnt)))
        return out[:cnt.value]


# ---- reference-style API ---------------------------------------------------------------------

class B2vec2(tuple):
    def __new__(cls, x=0.0, y=0.0):
        return tuple.__new__(cls, (float(x), float(y)))

    @property
    def x(self):
        return self[0]

    @property
    def y(self):
        return self[1]


class B2circleShape:
    """src/shapes/b2_circle_shape.rs: m_p (local centre) and m_radius."""

    def __init__(self, radius=0.0, p=(0.0, 0.0)):
        self.m_radius = float(radius)
        self.m_p = B2vec2(*p)

    def _def(self):
        d = np.zeros((), SHAPE_DEF)
        d["type"] = SHAPE_CIRCLE
        d["radius"] = self.m_radius
        d["params"][:2] = self.m_p
        return d


class B2polygonShape:
    """src/shapes/b2_polygon_shape.rs: set_as_box, set_as_box_angle, set (convex hull)."""

    def __init__(self):
        self._d = None

    def set_as_box(self, hx, hy):
        d = np.zeros((), SHAPE_DEF)
        d["type"] = SHAPE_BOX
        d["params"][:2] = (hx, hy)
        self._d = d

    def set_as_box_angle(self, hx, hy, center, angle):
        d = np.zeros((), SHAPE_DEF)
        d["type"] = SHAPE_BOX
        d["flags"] = 1
        d["params"][:5] = (hx, hy, center[0], center[1], angle)
        self._d = d

    def set(self, vertices):
        v = np.asarray(vertices, np.float32).reshape(-1, 2)
        if not 3 <= len(v) <= 8:
            raise ValueError("a polygon needs 3..8 vertices")
        d = np.zeros((), SHAPE_DEF)
        d["type"] = SHAPE_POLYGON
        d["count"] = len(v)
        d["params"][:2 * len(v)] = v.ravel()
        self._d = d

    def _def(self):
        if self._d is None:
            raise ValueError("polygon shape not set")
        return self._d


class B2edgeShape:
    """src/shapes/b2_edge_shape.rs: set_two_sided, set_one_sided."""

    def __init__(self):
        self._d = None

    def set_two_sided(self, v1, v2):
        d = np.zeros((), SHAPE_DEF)
        d["type"] = SHAPE_EDGE
        d["params"][:4] = (v1[0], v1[1], v2[0], v2[1])
        self._d = d

    def set_one_sided(self, v0, v1, v2, v3):
        d = np.zeros((), SHAPE_DEF)
        d["type"] = SHAPE_EDGE
        d["flags"] = 1
        d["params"][:8] = (v1[0], v1[1], v2[0], v2[1], v0[0], v0[1], v3[0], v3[1])
        self._d = d

    def _def(self):
        if self._d is None:
            raise ValueError("edge shape not set")
        return self._d


class B2filter:
    def __init__(self, category_bits=0x0001, mask_bits=0xFFFF, group_index=0):
        self.category_bits = category_bits
        self.mask_bits = mask_bits
        self.group_index = group_index


class B2fixtureDef:
    """src/b2_fixture.rs:44-58 defaults: friction 0.2, restitution 0, threshold 1, density 0."""

    def __init__(self, shape=None, friction=0.2, restitution=0.0, restitution_threshold=1.0, density=0.0,
                 is_sensor=False, filter=None):
        self.shape = shape
        self.friction = friction
        self.restitution = restitution
        self.restitution_threshold = restitution_threshold
        self.density = density
        self.is_sensor = is_sensor
        self.filter = filter or B2filter()

    def _def(self):
        d = np.zeros((), FIXTURE_DEF)
        d["friction"] = self.friction
        d["restitution"] = self.restitution
        d["restitution_threshold"] = self.restitution_threshold
        d["density"] = self.density
        d["is_sensor"] = int(self.is_sensor)
        d["category"] = self.filter.category_bits
        d["mask"] = self.filter.mask_bits
        d["group"] = self.filter.group_index
        return d


class B2bodyDef:
    """src/b2_body.rs:39-58 defaults."""

    def __init__(self, body_type=B2_STATIC_BODY, position=(0.0, 0.0), angle=0.0, linear_velocity=(0.0, 0.0),
                 angular_velocity=0.0, linear_damping=0.0, angular_damping=0.0, allow_sleep=True, awake=True,
                 fixed_rotation=False, bullet=False, gravity_scale=1.0):
        self.body_type = body_type
        self.position = position
        self.angle = angle
        self.linear_velocity = linear_velocity
        self.angular_velocity = angular_velocity
        self.linear_damping = linear_damping
        self.angular_damping = angular_damping
        self.allow_sleep = allow_sleep
        self.awake = awake
        self.fixed_rotation = fixed_rotation
        self.bullet = bullet
        self.gravity_scale = gravity_scale

    def _def(self):
        d = np.zeros((), BODY_DEF)
        d["type"] = self.body_type
        d["allow_sleep"] = int(self.allow_sleep)
        d["awake"] = int(self.awake)
        d["fixed_rotation"] = int(self.fixed_rotation)
        d["position"] = self.position
        d["angle"] = self.angle
        d["linear_velocity"] = self.linear_velocity
        d["angular_velocity"] = self.angular_velocity
        d["linear_damping"] = self.linear_damping
        d["angular_damping"] = self.angular_damping
        d["gravity_scale"] = self.gravity_scale
        d["bullet"] = int(self.bullet)
        return d


class B2body:
    """Handle on one body of a :class:`B2world` (src/b2_body.rs getters)."""

    def __init__(self, world, index):
        self._world = world
        self.index = index

    def create_fixture(self, fixture_def):
        return self._world._queue_fixture(self.index, fixture_def)

    def create_fixture_by_shape(self, shape, density):
        return self.create_fixture(B2fixtureDef(shape=shape, density=density))

    def _state(self):
        return self._world._states()[self.index]

    def get_position(self):
        return B2vec2(*self._state()["p"])

    def get_angle(self):
        return float(self._state()["a"])

    def get_world_center(self):
        return B2vec2(*self._state()["c"])

    def get_linear_velocity(self):
        return B2vec2(*self._state()["v"])

    def get_angular_velocity(self):
        return float(self._state()["w"])

    def get_mass(self):
        return float(self._world._world.mass()[self.index, 0])

    def get_inertia(self):
        m = self._world._world.mass()[self.index]
        return float(m[2] + m[0] * (m[4] * m[4] + m[5] * m[5]))

    def is_awake(self):
        return bool(self._state()["awake"] != 0)

    def set_bullet(self, flag):
        """B2body::set_bullet (src/b2_body.rs:504-506)."""
        self._world._set_bullet(self.index, bool(flag))

    def is_bullet(self):
        return self._world._bullets.get(self.index, False)

    def get_sleep_time(self):
        return float(self._state()["sleep_time"])


def _local_point(body, p):
    """B2body::get_local_point: mul_t(xf, p) in f32 (src/b2_body.rs)."""
    st = body._state()
    f = np.float32
    dx, dy = f(p[0]) - f(st["p"][0]), f(p[1]) - f(st["p"][1])
    s, c = f(st["q"][0]), f(st["q"][1])
    return (float(c * dx + s * dy), float(-s * dx + c * dy))


class B2distanceJointDef:
    """``B2distanceJointDef`` (src/joints/b2_distance_joint.rs:11-60)."""

    def __init__(self):
        self.body_a = self.body_b = None
        self.collide_connected = False
        self.local_anchor_a = (0.0, 0.0)
        self.local_anchor_b = (0.0, 0.0)
        self.length, self.min_length, self.max_length = 1.0, 0.0, float(np.finfo(np.float32).max)
        self.stiffness = self.damping = 0.0

    def initialize(self, body_a, body_b, anchor_a, anchor_b):
        """b2_distance_joint_def_initialize (src/private/dynamics/joints/b2_distance_joint.rs:23-40)."""
        self.body_a, self.body_b = body_a, body_b
        self.local_anchor_a = _local_point(body_a, anchor_a)
        self.local_anchor_b = _local_point(body_b, anchor_b)
        d = np.float32(anchor_b[0]) - np.float32(anchor_a[0]), np.float32(anchor_b[1]) - np.float32(anchor_a[1])
        self.length = float(max(np.sqrt(d[0] * d[0] + d[1] * d[1]), np.float32(0.005)))
        self.min_length = self.max_length = self.length

    def _def(self):
        d = np.zeros((), JOINT_DEF)
        d["type"] = JOINT_DISTANCE
        d["body_a"], d["body_b"] = self.body_a.index, self.body_b.index
        d["collide_connected"] = int(self.collide_connected)
        d["local_anchor_a"], d["local_anchor_b"] = self.local_anchor_a, self.local_anchor_b
        d["length"], d["min_length"], d["max_length"] = self.length, self.min_length, self.max_length
        d["stiffness"], d["damping"] = self.stiffness, self.damping
        return d


class B2revoluteJointDef:
    """``B2revoluteJointDef`` (src/joints/b2_revolute_joint.rs:10-95)."""

    def __init__(self):
        self.body_a = self.body_b = None
        self.collide_connected = False
        self.local_anchor_a = (0.0, 0.0)
        self.local_anchor_b = (0.0, 0.0)
        self.reference_angle = 0.0
        self.enable_limit, self.lower_angle, self.upper_angle = False, 0.0, 0.0
        self.enable_motor, self.motor_speed, self.max_motor_torque = False, 0.0, 0.0

    def initialize(self, body_a, body_b, anchor):
        """B2revoluteJointDef::initialize (src/joints/b2_revolute_joint.rs:76-82)."""
        self.body_a, self.body_b = body_a, body_b
        self.local_anchor_a = _local_point(body_a, anchor)
        self.local_anchor_b = _local_point(body_b, anchor)
        self.reference_angle = float(np.float32(body_b.get_angle()) - np.float32(body_a.get_angle()))

    def _def(self):
        d = np.zeros((), JOINT_DEF)
        d["type"] = JOINT_REVOLUTE
        d["body_a"], d["body_b"] = self.body_a.index, self.body_b.index
        d["collide_connected"] = int(self.collide_connected)
        d["local_anchor_a"], d["local_anchor_b"] = self.local_anchor_a, self.local_anchor_b
        d["reference_angle"] = self.reference_angle
        d["enable_limit"], d["lower_angle"], d["upper_angle"] = int(self.enable_limit), self.lower_angle, self.upper_angle
        d["enable_motor"], d["motor_speed"] = int(self.enable_motor), self.motor_speed
        d["max_motor_torque"] = self.max_motor_torque
        return d


class B2weldJointDef:
    """``B2weldJointDef`` (src/joints/b2_weld_joint.rs:10-60)."""

    def __init__(self):
        self.body_a = self.body_b = None
        self.collide_connected = False
        self.local_anchor_a = (0.0, 0.0)
        self.local_anchor_b = (0.0, 0.0)
        self.reference_angle = 0.0
        self.stiffness = self.damping = 0.0

    def initialize(self, body_a, body_b, anchor):
        """B2weldJointDef::initialize (src/joints/b2_weld_joint.rs:54-60)."""
        self.body_a, self.body_b = body_a, body_b
        self.local_anchor_a = _local_point(body_a, anchor)
        self.local_anchor_b = _local_point(body_b, anchor)
        self.reference_angle = float(np.float32(body_b.get_angle()) - np.float32(body_a.get_angle()))

    def _def(self):
        d = np.zeros((), JOINT_DEF)
        d["type"] = JOINT_WELD
        d["body_a"], d["body_b"] = self.body_a.index, self.body_b.index
        d["collide_connected"] = int(self.collide_connected)
        d["local_anchor_a"], d["local_anchor_b"] = self.local_anchor_a, self.local_anchor_b
        d["reference_angle"], d["stiffness"], d["damping"] = self.reference_angle, self.stiffness, self.damping
        return d


class B2prismaticJointDef:
    """``B2prismaticJointDef`` (src/joints/b2_prismatic_joint.rs): anchors, the axis in body A's
    frame, reference angle, translation limits, motor (speed, max force)."""

    def __init__(self):
        self.body_a = self.body_b = None
        self.collide_connected = False
        self.local_anchor_a = (0.0, 0.0)
        self.local_anchor_b = (0.0, 0.0)
        self.local_axis_a = (1.0, 0.0)
        self.reference_angle = 0.0
        self.enable_limit, self.lower_translation, self.upper_translation = False, 0.0, 0.0
        self.enable_motor, self.max_motor_force, self.motor_speed = False, 0.0, 0.0

    def initialize(self, body_a, body_b, anchor, axis):
        """B2prismaticJointDef::initialize (src/joints/b2_prismatic_joint.rs:75-90): anchors from the
        bodies' poses, the axis as a local vector of body A, the reference angle from their angles."""
        self.body_a, self.body_b = body_a, body_b
        self.local_anchor_a = _local_point(body_a, anchor)
        self.local_anchor_b = _local_point(body_b, anchor)
        st = body_a._state()
        f = np.float32
        s, c = f(st["q"][0]), f(st["q"][1])
        ax, ay = f(axis[0]), f(axis[1])
        self.local_axis_a = (float(c * ax + s * ay), float(-s * ax + c * ay))  # get_local_vector
        self.reference_angle = float(np.float32(body_b.get_angle()) - np.float32(body_a.get_angle()))

    def _def(self):
        d = np.zeros((), JOINT_DEF)
        d["type"] = JOINT_PRISMATIC
        d["body_a"], d["body_b"] = self.body_a.index, self.body_b.index
        d["collide_connected"] = int(self.collide_connected)
        d["local_anchor_a"], d["local_anchor_b"] = self.local_anchor_a, self.local_anchor_b
        d["local_axis_a"] = self.local_axis_a
        d["reference_angle"] = self.reference_angle
        d["enable_limit"], d["lower_angle"], d["upper_angle"] = (int(self.enable_limit), self.lower_translation,
                                                                 self.upper_translation)
        d["enable_motor"], d["motor_speed"], d["max_motor_torque"] = (int(self.enable_motor), self.motor_speed,
                                                                       self.max_motor_force)
        return d


class B2fixture:
    """Handle on one fixture of a :class:`B2world` (fixtures are numbered in creation order)."""

    def __init__(self, world, index):
        self._world = world
        self.index = index

    def get_body(self):
        return B2body(self._world, self._world._fixture_info[self.index][0])

    def is_sensor(self):
        return self._world._fixture_info[self.index][1]

    def __eq__(self, other):
        return isinstance(other, B2fixture) and other._world is self._world and other.index == self.index

    def __hash__(self):
        return hash((id(self._world), self.index))


class B2contact:
    """One contact of :meth:`B2world.get_contact_list` (src/b2_contact.rs getters)."""

    def __init__(self, world, fixture_a, fixture_b, touching, manifold):
        self._world = world
        self._fa, self._fb = fixture_a, fixture_b
        self._touching = touching
        self._manifold = manifold

    def get_fixture_a(self):
        return B2fixture(self._world, self._fa)

    def get_fixture_b(self):
        return B2fixture(self._world, self._fb)

    def is_touching(self):
        return self._touching

    def get_manifold(self):
        return self._manifold


def _local_vector(body, v):
    """B2body::get_local_vector: mul_t(q, v) in f32."""
    st = body._state()
    f = np.float32
    s, c = f(st["q"][0]), f(st["q"][1])
    return (float(c * f(v[0]) + s * f(v[1])), float(-s * f(v[0]) + c * f(v[1])))


class _JointDefBase:
    kind = 0

    def _base(self):
        d = np.zeros((), JOINT_DEF)
        d["type"] = self.kind
        d["body_a"], d["body_b"] = self.body_a.index, self.body_b.index
        d["collide_connected"] = int(self.collide_connected)
        return d


class B2wheelJointDef(_JointDefBase):
    """``B2wheelJointDef`` (src/joints/b2_wheel_joint.rs:39-92): anchors, the axis in body A's
    frame, translation limits, motor torque and speed, a spring (stiffness, damping)."""
    kind = JOINT_WHEEL

    def __init__(self):
        self.body_a = self.body_b = None
        self.collide_connected = False
        self.local_anchor_a = (0.0, 0.0)
        self.local_anchor_b = (0.0, 0.0)
        self.local_axis_a = (1.0, 0.0)
        self.enable_limit, self.lower_translation, self.upper_translation = False, 0.0, 0.0
        self.enable_motor, self.max_motor_torque, self.motor_speed = False, 0.0, 0.0
        self.stiffness = self.damping = 0.0

    def initialize(self, body_a, body_b, anchor, axis):
        self.body_a, self.body_b = body_a, body_b
        self.local_anchor_a = _local_point(body_a, anchor)
        self.local_anchor_b = _local_point(body_b, anchor)
        self.local_axis_a = _local_vector(body_a, axis)

    def _def(self):
        d = self._base()
        d["local_anchor_a"], d["local_anchor_b"] = self.local_anchor_a, self.local_anchor_b
        d["local_axis_a"] = self.local_axis_a
        d["enable_limit"], d["lower_angle"], d["upper_angle"] = (int(self.enable_limit), self.lower_translation,
                                                                 self.upper_translation)
        d["enable_motor"], d["max_motor_torque"], d["motor_speed"] = (int(self.enable_motor), self.max_motor_torque,
                                                                       self.motor_speed)
        d["stiffness"], d["damping"] = self.stiffness, self.damping
        return d


class B2frictionJointDef(_JointDefBase):
    """``B2frictionJointDef`` (src/joints/b2_friction_joint.rs): anchors, max force and torque."""
    kind = JOINT_FRICTION

    def __init__(self):
        self.body_a = self.body_b = None
        self.collide_connected = False
        self.local_anchor_a = (0.0, 0.0)
        self.local_anchor_b = (0.0, 0.0)
        self.max_force = self.max_torque = 0.0

    def initialize(self, body_a, body_b, anchor):
        self.body_a, self.body_b = body_a, body_b
        self.local_anchor_a = _local_point(body_a, anchor)
        self.local_anchor_b = _local_point(body_b, anchor)

    def _def(self):
        d = self._base()
        d["local_anchor_a"], d["local_anchor_b"] = self.local_anchor_a, self.local_anchor_b
        d["max_force"], d["max_torque"] = self.max_force, self.max_torque
        return d


class B2motorJointDef(_JointDefBase):
    """``B2motorJointDef`` (src/joints/b2_motor_joint.rs): linear and angular offset of body B from
    body A, max force and torque, correction factor."""
    kind = JOINT_MOTOR

    def __init__(self):
        self.body_a = self.body_b = None
        self.collide_connected = False
        self.linear_offset = (0.0, 0.0)
        self.angular_offset = 0.0
        self.max_force, self.max_torque, self.correction_factor = 1.0, 1.0, 0.3

    def initialize(self, body_a, body_b):
        """B2motorJointDef::initialize: the current offset of body B from body A."""
        self.body_a, self.body_b = body_a, body_b
        self.linear_offset = _local_point(body_a, tuple(body_b._state()["p"]))
        self.angular_offset = float(np.float32(body_b.get_angle()) - np.float32(body_a.get_angle()))

    def _def(self):
        d = self._base()
        d["linear_offset"], d["angular_offset"] = self.linear_offset, self.angular_offset
        d["max_force"], d["max_torque"], d["correction_factor"] = self.max_force, self.max_torque, self.correction_factor
        return d


class B2mouseJointDef(_JointDefBase):
    """``B2mouseJointDef`` (src/joints/b2_mouse_joint.rs:35-52): body B's point pulled to the target
    (body A is a reference body, usually the ground)."""
    kind = JOINT_MOUSE

    def __init__(self):
        self.body_a = self.body_b = None
        self.collide_connected = False
        self.target = (0.0, 0.0)
        self.max_force = self.stiffness = self.damping = 0.0

    def _def(self):
        d = self._base()
        d["target"] = self.target
        d["max_force"], d["stiffness"], d["damping"] = self.max_force, self.stiffness, self.damping
        return d


class B2pulleyJointDef(_JointDefBase):
    """``B2pulleyJointDef`` (src/joints/b2_pulley_joint.rs): ground anchors, anchors, lengths and
    ratio; collide_connected defaults to true."""
    kind = JOINT_PULLEY

    def __init__(self):
        self.body_a = self.body_b = None
        self.collide_connected = True
        self.ground_anchor_a, self.ground_anchor_b = (-1.0, 1.0), (1.0, 1.0)
        self.local_anchor_a, self.local_anchor_b = (-1.0, 0.0), (1.0, 0.0)
        self.length_a = self.length_b = 0.0
        self.ratio = 1.0

    def initialize(self, body_a, body_b, ground_anchor_a, ground_anchor_b, anchor_a, anchor_b, ratio):
        """B2pulleyJointDef::initialize (src/joints/b2_pulley_joint.rs:59-82)."""
        self.body_a, self.body_b = body_a, body_b
        self.ground_anchor_a, self.ground_anchor_b = tuple(ground_anchor_a), tuple(ground_anchor_b)
        self.local_anchor_a = _local_point(body_a, anchor_a)
        self.local_anchor_b = _local_point(body_b, anchor_b)
        f = np.float32
        da = (f(anchor_a[0]) - f(ground_anchor_a[0]), f(anchor_a[1]) - f(ground_anchor_a[1]))
        db = (f(anchor_b[0]) - f(ground_anchor_b[0]), f(anchor_b[1]) - f(ground_anchor_b[1]))
        self.length_a = float(np.sqrt(da[0] * da[0] + da[1] * da[1]))
        self.length_b = float(np.sqrt(db[0] * db[0] + db[1] * db[1]))
        self.ratio = ratio

    def _def(self):
        d = self._base()
        d["ground_anchor_a"], d["ground_anchor_b"] = self.ground_anchor_a, self.ground_anchor_b
        d["local_anchor_a"], d["local_anchor_b"] = self.local_anchor_a, self.local_anchor_b
        d["length_a"], d["length_b"], d["ratio"] = self.length_a, self.length_b, self.ratio
        return d


class B2gearJointDef(_JointDefBase):
    """``B2gearJointDef`` (src/joints/b2_gear_joint.rs): two revolute or prismatic joints and the
    ratio; body_a and body_b are the joints' bodies B."""
    kind = JOINT_GEAR

    def __init__(self):
        self.body_a = self.body_b = None
        self.collide_connected = False
        self.joint1 = self.joint2 = None
        self.ratio = 1.0

    def _def(self):
        d = self._base()
        d["joint1"], d["joint2"], d["ratio"] = self.joint1.index, self.joint2.index, self.ratio
        return d


class B2joint:
    """Handle on one joint of a :class:`B2world`."""

    def __init__(self, world, index):
        self._world = world
        self.index = index

    def get_impulses(self):
        return self._world._world.joint_impulses()[self.index]

    def set_target(self, target):
        """B2mouseJoint::set_target."""
        self._world._world.set_joint_target(self.index, target)
        self._world._cache = None


class B2world:
    """``B2world`` (src/b2_world.rs) stepping on the GPU.

    Bodies and fixtures are queued on the host and handed to the device in one bulk call on
    the next step or query, in creation order, so proxy ids and mass data match the reference.
    """

    def __init__(self, gravity=(0.0, -10.0)):
        self._world = World(gravity)
        self._bodies = []
        self._fixtures = []
        self._n_flushed_bodies = 0
        self._n_flushed_fixtures = 0
        self._fixture_info = []  # per fixture: (body index, is_sensor)
        self._bullets = {}
        self._cache = None
        # B2world::new defaults (src/private/dynamics/b2_world.rs:37-43): warm starting, sleeping
        # and continuous physics on
        self._warm_starting = True
        self._allow_sleep = True
        self._continuous = True
        self._apply_flags()

    @classmethod
    def new(cls, gravity):
        return cls(gravity)

    def set_warm_starting(self, flag):
        self._warm_starting = bool(flag)
        self._apply_flags()

    def set_continuous_physics(self, flag):
        self._continuous = bool(flag)
        self._apply_flags()

    def get_continuous_physics(self):
        return self._continuous

    def set_sub_stepping(self, flag):
        self._world.set_sub_stepping(flag)

    def get_sub_stepping(self):
        return self._world.sub_stepping()[0]

    def set_allow_sleeping(self, flag):
        # B2world::set_allow_sleeping (src/private/dynamics/b2_world.rs:342-353): switching it off
        # wakes every body
        self._flush()
        self._allow_sleep = bool(flag)
        self._apply_flags()

    def get_allow_sleeping(self):
        return self._allow_sleep

    def _apply_flags(self):
        self._world.set_flags(self._warm_starting, self._allow_sleep, self._continuous, True)

    def create_body(self, body_def):
        self._bodies.append(body_def._def())
        self._cache = None
        index = self._n_flushed_bodies + len(self._bodies) - 1
        if body_def.bullet:
            self._bullets[index] = True
        return B2body(self, index)

    def _set_bullet(self, index, flag):
        k = index - self._n_flushed_bodies
        if k >= 0:
            self._bodies[k]["bullet"] = int(flag)
        else:
            self._world.set_bullet(index, flag)
        self._bullets[index] = flag

    def _queue_fixture(self, body_index, fixture_def):
        if fixture_def.shape is None:
            raise ValueError("fixture definition needs a shape")
        self._fixtures.append((body_index, fixture_def.shape._def(), fixture_def._def()))
        self._fixture_info.append((body_index, bool(fixture_def.is_sensor)))
        self._cache = None
        return B2fixture(self, len(self._fixture_info) - 1)

    def _flush(self):
        if not self._bodies and not self._fixtures:
            return
        base = self._n_flushed_bodies
        bd = np.array(self._bodies, BODY_DEF) if self._bodies else np.zeros(0, BODY_DEF)
        fb = np.array([b - base for b, _, _ in self._fixtures], np.int32)
        sd = np.array([s for _, s, _ in self._fixtures], SHAPE_DEF) if self._fixtures else np.zeros(0, SHAPE_DEF)
        fd = np.array([f for _, _, f in self._fixtures], FIXTURE_DEF) if self._fixtures else np.zeros(0, FIXTURE_DEF)
        self._world.add(bd, fb, sd, fd)
        self._n_flushed_bodies += len(self._bodies)
        self._n_flushed_fixtures += len(self._fixtures)
        self._bodies = []
        self._fixtures = []

    def step(self, time_step, velocity_iterations, position_iterations):
        self._flush()
        self._world.step(time_step, velocity_iterations, position_iterations)
        self._cache = None

    def _states(self):
        self._flush()
        if self._cache is None:
            self._cache = self._world.body_states()
        return self._cache

    def create_joint(self, joint_def):
        """B2world::create_joint (src/private/dynamics/b2_world.rs:156-275): distance and revolute."""
        self._flush()
        self._world.add_joints(np.array([joint_def._def()], JOINT_DEF))
        self._cache = None
        return B2joint(self, self._world.joint_count - 1)

    def get_body_count(self):
        self._flush()
        return self._world.body_count

    def get_contact_count(self):
        self._flush()
        return self._world.contact_count

    def get_contact_list(self):
        """B2world::get_contact_list: the contacts of the current set, each with its fixtures,
        touching flag (for a sensor: the shapes overlap) and manifold."""
        self._flush()
        fab, flags, man = self._world.contacts()
        return [B2contact(self, int(a), int(b), bool(t), m) for (a, b), t, m in zip(fab, flags, man)]

    def get_profile(self):
        return self._world.profile()
