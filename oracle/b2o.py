"""ctypes wrapper of the CPU oracle (test infrastructure only).

The oracle is a sequential C++ restatement of the box2d-rs ``B2world::step`` path (see the
C++ sources next to this file). Only the test-suite, the smoke check and the CPU-baseline
leg of the benchmark import this module, and only as the checker: the product library never
loads the oracle.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

CIRCLE, EDGE, POLYGON, BOX = 0, 1, 2, 4
STATIC, KINEMATIC, DYNAMIC = 0, 1, 2


def _dtypes(real):
    body_def = np.dtype([("type", "<i4"), ("allow_sleep", "<i4"), ("awake", "<i4"), ("fixed_rotation", "<i4"),
                         ("position", real, 2), ("angle", real), ("linear_velocity", real, 2),
                         ("angular_velocity", real), ("linear_damping", real), ("angular_damping", real),
                         ("gravity_scale", real), ("bullet", "<i4")], align=True)
    shape_def = np.dtype([("type", "<i4"), ("count", "<i4"), ("flags", "<i4"), ("radius", real),
                          ("params", real, 36)], align=True)
    fixture_def = np.dtype([("friction", real), ("restitution", real), ("restitution_threshold", real),
                            ("density", real), ("is_sensor", "<i4"), ("category", "<u2"), ("mask", "<u2"),
                            ("group", "<i2"), ("pad0", "<i2"), ("pad1", "<i4")], align=True)
    body_state = np.dtype([("p", real, 2), ("q", real, 2), ("c", real, 2), ("a", real), ("v", real, 2),
                           ("w", real), ("c0", real, 2), ("a0", real), ("sleep_time", real), ("awake", real),
                           ("pad", real)], align=True)
    point = np.dtype([("local_point", real, 2), ("normal_impulse", real), ("tangent_impulse", real),
                      ("id", "<u4"), ("pad", "<i4")], align=True)
    manifold = np.dtype([("type", "<i4"), ("point_count", "<i4"), ("local_normal", real, 2),
                         ("local_point", real, 2), ("points", point, 2)], align=True)
    shape = np.dtype([("type", "<i4"), ("count", "<i4"), ("one_sided", "<i4"), ("pad", "<i4"), ("radius", real),
                      ("p", real, 2), ("centroid", real, 2), ("v0", real, 2), ("v1", real, 2), ("v2", real, 2),
                      ("v3", real, 2), ("vertices", real, 16), ("normals", real, 16)], align=True)
    joint_def = np.dtype([("type", "<i4"), ("body_a", "<i4"), ("body_b", "<i4"), ("collide_connected", "<i4"),
                          ("local_anchor_a", real, 2), ("local_anchor_b", real, 2), ("length", real),
                          ("min_length", real), ("max_length", real), ("stiffness", real), ("damping", real),
                          ("reference_angle", real), ("lower_angle", real), ("upper_angle", real),
                          ("max_motor_torque", real), ("motor_speed", real), ("enable_limit", "<i4"),
                          ("enable_motor", "<i4"), ("local_axis_a", real, 2), ("max_force", real),
                          ("max_torque", real), ("linear_offset", real, 2), ("angular_offset", real),
                          ("correction_factor", real), ("target", real, 2), ("ground_anchor_a", real, 2),
                          ("ground_anchor_b", real, 2), ("length_a", real), ("length_b", real), ("ratio", real),
                          ("joint1", "<i4"), ("joint2", "<i4")], align=True)
    joint_state = np.dtype([("impulse", real, 5), ("pad", real, 3)], align=True)
    return dict(body_def=body_def, shape_def=shape_def, fixture_def=fixture_def, body_state=body_state,
                manifold=manifold, shape=shape, joint_def=joint_def, joint_state=joint_state)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Oracle:
    """One instantiation of the oracle library (f32 or f64)."""

    def __init__(self, precision="f32"):
        path = os.path.join(_BUILD, "liboracle_%s.so" % precision)
        if not os.path.exists(path):
            raise FileNotFoundError("oracle library not built: %s (run make -C oracle)" % path)
        self.lib = ctypes.CDLL(path)
        self.real = np.float32 if precision == "f32" else np.float64
        self.creal = ctypes.c_float if precision == "f32" else ctypes.c_double
        assert self.lib.b2o_real_size() == np.dtype(self.real).itemsize
        self.dt = _dtypes(self.real)
        L, R, P = self.lib, self.creal, ctypes.c_void_p
        I = ctypes.c_int
        L.b2o_world_new.restype = P
        L.b2o_world_new.argtypes = [R, R]
        L.b2o_world_free.argtypes = [P]
        L.b2o_world_set_flags.argtypes = [P, I, I, I]
        L.b2o_world_add.argtypes = [P, I, P, I, P, P, P, P]
        L.b2o_world_step.argtypes = [P, R, I, I]
        L.b2o_world_add_joints.argtypes = [P, I, P]
        L.b2o_world_get_joints.argtypes = [P, P]
        L.b2o_world_collide.argtypes = [P]
        L.b2o_world_run.argtypes = [P, R, I, I, I]
        L.b2o_world_run.restype = ctypes.c_double
        for f in ("b2o_world_body_count", "b2o_world_fixture_count", "b2o_world_contact_count"):
            getattr(L, f).argtypes = [P]
        L.b2o_world_get_bodies.argtypes = [P, P]
        L.b2o_world_set_bodies.argtypes = [P, P]
        L.b2o_world_set_awake.argtypes = [P, I, I]
        L.b2o_world_set_bullet.argtypes = [P, I, I]
        L.b2o_world_get_mass.argtypes = [P, P]
        L.b2o_world_get_proxies.argtypes = [P, P, P]
        L.b2o_world_get_contacts.argtypes = [P, I, P, P, P, P]
        L.b2o_world_get_profile.argtypes = [P, P]
        L.b2o_make_shape.argtypes = [P, P]
        L.b2o_shape_mass.argtypes = [P, R, P]
        L.b2o_shape_aabb.argtypes = [P, P, P]
        L.b2o_collide.argtypes = [I, P, P, P, P, P]
        L.b2o_distance.argtypes = [I, P, P, P, P, I, P]
        L.b2o_time_of_impact.argtypes = [I, P, P, P, P, R, P]
        L.b2o_world_set_continuous.argtypes = [P, I]
        L.b2o_world_toi_events.argtypes = [P]
        L.b2o_world_toi_events.restype = I
        L.b2o_world_set_sub_stepping.argtypes = [P, I]
        L.b2o_world_set_joint_target.argtypes = [P, I, R, R]
        L.b2o_world_step_complete.argtypes = [P]
        L.b2o_world_step_complete.restype = I
        L.b2o_world_manifold.argtypes = [I, P, P, P, P, P, P]
        L.b2o_rot.argtypes = [I, P, P]
        L.b2o_brute_pairs.argtypes = [I, P, ctypes.c_longlong, P]
        L.b2o_brute_pairs.restype = ctypes.c_longlong
        L.b2o_solver_replay.argtypes = [I, P, P, P, I, P, P, P, P, P, P, P, P, I]
        L.b2o_solver_replay.restype = I

    # -- conversion helpers ---------------------------------------------------------------
    def cast(self, arr, kind):
        """Convert a structured array (any real precision) to this oracle's layout."""
        dt = self.dt[kind]
        if arr.dtype == dt:
            return np.ascontiguousarray(arr)
        out = np.zeros(arr.shape, dt)
        for name in dt.names:
            if name in arr.dtype.names:
                out[name] = arr[name]
        return out

    # -- shapes and stateless collision -----------------------------------------------------
    def make_shapes(self, shape_defs):
        sd = self.cast(shape_defs, "shape_def")
        out = np.zeros(len(sd), self.dt["shape"])
        status = np.zeros(len(sd), np.int32)
        for i in range(len(sd)):
            status[i] = self.lib.b2o_make_shape(_ptr(sd[i:i + 1]), _ptr(out[i:i + 1]))
        return out, status

    def shape_mass(self, shapes, density):
        out = np.zeros((len(shapes), 4), self.real)
        for i in range(len(shapes)):
            self.lib.b2o_shape_mass(_ptr(shapes[i:i + 1]), self.creal(density), _ptr(out[i]))
        return out

    def shape_aabb(self, shapes, xf):
        xf = np.ascontiguousarray(xf, self.real).reshape(-1, 4)
        out = np.zeros((len(shapes), 4), self.real)
        for i in range(len(shapes)):
            self.lib.b2o_shape_aabb(_ptr(shapes[i:i + 1]), _ptr(xf[i]), _ptr(out[i]))
        return out

    def collide(self, shapes_a, xf_a, shapes_b, xf_b):
        n = len(shapes_a)
        xa = np.ascontiguousarray(xf_a, self.real).reshape(n, 4)
        xb = np.ascontiguousarray(xf_b, self.real).reshape(n, 4)
        sa = np.ascontiguousarray(shapes_a)
        sb = np.ascontiguousarray(shapes_b)
        out = np.zeros(n, self.dt["manifold"])
        self.lib.b2o_collide(n, _ptr(sa), _ptr(xa), _ptr(sb), _ptr(xb), _ptr(out))
        return out

    def distance(self, shapes_a, xf_a, shapes_b, xf_b, use_radii=False):
        """GJK distance (src/private/collision/b2_distance.rs:404): rows of (distance, point_a,
        point_b, iterations)."""
        n = len(shapes_a)
        xa = np.ascontiguousarray(xf_a, self.real).reshape(n, 4)
        xb = np.ascontiguousarray(xf_b, self.real).reshape(n, 4)
        sa = np.ascontiguousarray(shapes_a)
        sb = np.ascontiguousarray(shapes_b)
        out = np.zeros((n, 6), self.real)
        self.lib.b2o_distance(n, _ptr(sa), _ptr(xa), _ptr(sb), _ptr(xb), int(use_radii), _ptr(out))
        return out

    def time_of_impact(self, shapes_a, sweep_a, shapes_b, sweep_b, t_max=1.0):
        """Conservative advancement (src/private/collision/b2_time_of_impact.rs:232). A sweep row is
        (local_center(2), c0(2), c(2), a0, a, alpha0). Returns (state, t) arrays; states as
        B2toioutputState: 1 failed, 2 overlapped, 3 touching, 4 separated."""
        n = len(shapes_a)
        wa = np.ascontiguousarray(sweep_a, self.real).reshape(n, 9)
        wb = np.ascontiguousarray(sweep_b, self.real).reshape(n, 9)
        sa = np.ascontiguousarray(shapes_a)
        sb = np.ascontiguousarray(shapes_b)
        out = np.zeros((n, 2), self.real)
        self.lib.b2o_time_of_impact(n, _ptr(sa), _ptr(wa), _ptr(sb), _ptr(wb), t_max, _ptr(out))
        return out[:, 0].astype(np.int32), out[:, 1].copy()

    def world_manifold(self, manifolds, xf_a, ra, xf_b, rb):
        n = len(manifolds)
        m = self.cast(manifolds, "manifold")
        xa = np.ascontiguousarray(xf_a, self.real).reshape(n, 4)
        xb = np.ascontiguousarray(xf_b, self.real).reshape(n, 4)
        ra = np.ascontiguousarray(ra, self.real)
        rb = np.ascontiguousarray(rb, self.real)
        out = np.zeros((n, 8), self.real)
        self.lib.b2o_world_manifold(n, _ptr(m), _ptr(xa), _ptr(ra), _ptr(xb), _ptr(rb), _ptr(out))
        return out

    def rot(self, angles):
        a = np.ascontiguousarray(angles, self.real)
        sc = np.zeros((len(a), 2), self.real)
        self.lib.b2o_rot(len(a), _ptr(a), _ptr(sc))
        return sc

    def brute_pairs(self, fat):
        fat = np.ascontiguousarray(fat, self.real).reshape(-1, 4)
        n = len(fat)
        k = self.lib.b2o_brute_pairs(n, _ptr(fat), 0, None)
        out = np.zeros((max(k, 1), 2), np.int32)
        self.lib.b2o_brute_pairs(n, _ptr(fat), k, _ptr(out))
        return out[:k]

    def solver_replay(self, body_type, body_props, state, contact_bodies, contact_params, manifolds, order,
                      dt, dt_ratio, gravity, vel_iters=8, pos_iters=3, warm_starting=True, block_solve=True,
                      body_flags=None, allow_sleep=False):
        nb = len(body_type)
        nc = len(contact_bodies)
        bt = np.ascontiguousarray(body_type, np.int32)
        bp = np.ascontiguousarray(body_props, self.real).reshape(nb, 10)
        st = self.cast(state, "body_state").copy()
        cb = np.ascontiguousarray(contact_bodies, np.int32).reshape(nc, 2)
        cp = np.ascontiguousarray(contact_params, self.real).reshape(nc, 6)
        mf = self.cast(manifolds, "manifold").copy()
        od = np.ascontiguousarray(order, np.int32)
        sp = np.array([dt, dt_ratio, gravity[0], gravity[1]], self.real)
        it = np.array([vel_iters, pos_iters, int(warm_starting), int(block_solve)], np.int32)
        isl = np.zeros(nb, np.int32)
        fl = None if body_flags is None else np.ascontiguousarray(body_flags, np.int32)
        n_islands = self.lib.b2o_solver_replay(nb, _ptr(bt), _ptr(bp), _ptr(st), nc, _ptr(cb), _ptr(cp), _ptr(mf),
                                               _ptr(od), _ptr(sp), _ptr(it), _ptr(isl),
                                               None if fl is None else _ptr(fl), int(allow_sleep))
        return st, mf, isl, n_islands

    def world(self, gravity=(0.0, -10.0)):
        return OracleWorld(self, gravity)


class OracleWorld:
    """Sequential reference-order world (B2world::new / create_body / create_fixture / step)."""

    def __init__(self, oracle, gravity):
        self.o = oracle
        self.h = ctypes.c_void_p(oracle.lib.b2o_world_new(oracle.creal(gravity[0]), oracle.creal(gravity[1])))
        self.set_flags(warm_starting=True, allow_sleep=False, block_solve=True)

    def __del__(self):
        if getattr(self, "h", None):
            self.o.lib.b2o_world_free(self.h)
            self.h = None

    def set_flags(self, warm_starting=True, allow_sleep=False, block_solve=True, continuous=False):
        self.o.lib.b2o_world_set_flags(self.h, int(warm_starting), int(allow_sleep), int(block_solve))
        self.o.lib.b2o_world_set_continuous(self.h, int(continuous))

    @property
    def toi_events(self):
        """TOI events solved in the last step."""
        return self.o.lib.b2o_world_toi_events(self.h)

    def set_sub_stepping(self, flag=True):
        """One solved TOI event per step; the next step resumes the pass without solving."""
        self.o.lib.b2o_world_set_sub_stepping(self.h, int(flag))

    @property
    def step_complete(self):
        """False while a sub-stepped TOI pass is unfinished."""
        return bool(self.o.lib.b2o_world_step_complete(self.h))

    def add(self, body_defs, fixture_body, shape_defs, fixture_defs):
        bd = self.o.cast(body_defs, "body_def")
        sd = self.o.cast(shape_defs, "shape_def")
        fd = self.o.cast(fixture_defs, "fixture_def")
        fb = np.ascontiguousarray(fixture_body, np.int32)
        proxies = np.zeros(len(fb), np.int32)
        st = self.o.lib.b2o_world_add(self.h, len(bd), _ptr(bd), len(fb), _ptr(fb), _ptr(sd), _ptr(fd), _ptr(proxies))
        if st != 0:
            raise ValueError("degenerate polygon in shape definitions")
        return proxies

    def add_joints(self, joint_defs):
        jd = self.o.cast(joint_defs, "joint_def")
        made = self.o.lib.b2o_world_add_joints(self.h, len(jd), _ptr(jd))
        if made != len(jd):
            raise ValueError("joint %d: bad bodies or type" % made)
        self.n_joints = getattr(self, "n_joints", 0) + len(jd)

    def set_joint_target(self, joint, target):
        """B2mouseJoint::set_target."""
        self.o.lib.b2o_world_set_joint_target(self.h, int(joint), float(target[0]), float(target[1]))

    def joints(self):
        out = np.zeros(getattr(self, "n_joints", 0), self.o.dt["joint_state"])
        if len(out):
            self.o.lib.b2o_world_get_joints(self.h, _ptr(out))
        return out

    def step(self, dt=1.0 / 60.0, vel_iters=8, pos_iters=3, steps=1):
        if steps == 1:
            self.o.lib.b2o_world_step(self.h, self.o.creal(dt), vel_iters, pos_iters)
            return None
        return self.o.lib.b2o_world_run(self.h, self.o.creal(dt), vel_iters, pos_iters, steps)

    def collide(self):
        """Collide phase only: afterwards the contact list equals the fat-AABB overlap set."""
        self.o.lib.b2o_world_collide(self.h)

    @property
    def body_count(self):
        return self.o.lib.b2o_world_body_count(self.h)

    @property
    def fixture_count(self):
        return self.o.lib.b2o_world_fixture_count(self.h)

    @property
    def contact_count(self):
        return self.o.lib.b2o_world_contact_count(self.h)

    def bodies(self):
        out = np.zeros(self.body_count, self.o.dt["body_state"])
        self.o.lib.b2o_world_get_bodies(self.h, _ptr(out))
        return out

    def set_bodies(self, state):
        st = self.o.cast(state, "body_state")
        self.o.lib.b2o_world_set_bodies(self.h, _ptr(st))

    def set_awake(self, body, flag=True):
        """B2body::set_awake on one body (src/b2_body.rs:783-801)."""
        self.o.lib.b2o_world_set_awake(self.h, int(body), int(flag))

    def set_bullet(self, body, flag=True):
        """B2body::set_bullet."""
        self.o.lib.b2o_world_set_bullet(self.h, int(body), int(flag))

    def mass(self):
        out = np.zeros((self.body_count, 6), self.o.real)
        self.o.lib.b2o_world_get_mass(self.h, _ptr(out))
        return out

    def proxies(self):
        n = self.fixture_count
        fat = np.zeros((n, 4), self.o.real)
        pid = np.zeros(n, np.int32)
        self.o.lib.b2o_world_get_proxies(self.h, _ptr(fat), _ptr(pid))
        return fat, pid

    def contacts(self):
        n = self.contact_count
        fab = np.zeros((n, 2), np.int32)
        flags = np.zeros(n, np.int32)
        man = np.zeros(n, self.o.dt["manifold"])
        mat = np.zeros((n, 3), self.o.real)
        k = self.o.lib.b2o_world_get_contacts(self.h, n, _ptr(fab), _ptr(flags), _ptr(man), _ptr(mat))
        return fab[:k], flags[:k], man[:k], mat[:k]

    def profile(self):
        out = np.zeros(4, np.float64)
        self.o.lib.b2o_world_get_profile(self.h, _ptr(out))
        return out
