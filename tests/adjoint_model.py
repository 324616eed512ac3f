"""Host f64 model of a recorded run of frozen steps and its reverse pass (tests/host/adjoint_check.cpp),
plus a recorder that takes the frozen structure from the oracle's own run (CPU only).

The structure of a step is what b2g_tape_export returns for the GPU: per touching contact
(slot) its bodies, fixtures, manifold type | point count << 8, the slots its two points carry
their warm-start impulses from (slot * 2 + point over all steps, or -1), the number of position
iterations and the two feature ids; plus the carried impulses and the step parameters.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "host", "_build", "libadjoint_check.so")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def state6(states):
    """(n, 6) float64: c.x, c.y, a, v.x, v.y, w from body-state records."""
    s = np.zeros((len(states), 6))
    s[:, 0:2], s[:, 2], s[:, 3:5], s[:, 5] = states["c"], states["a"], states["v"], states["w"]
    return s


class HostAdjoint:
    def __init__(self, oracle32, scene, mass, gravity=(0.0, -10.0)):
        self.lib = ctypes.CDLL(LIB)
        self.lib.ac_new.restype = ctypes.c_void_p
        self.lib.ac_new.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        self.lib.ac_free.argtypes = [ctypes.c_void_p]
        self.lib.ac_add_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p]
        self.lib.ac_forward.argtypes = [ctypes.c_void_p] * 4
        self.lib.ac_backward.argtypes = [ctypes.c_void_p] * 6
        self.lib.ac_mismatch_bodies.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        bd = scene.body_defs
        self.nb, self.nf = len(bd), len(scene.fixture_defs)
        m = np.zeros((self.nb, 8))
        m[:, 0], m[:, 1], m[:, 2:4], m[:, 4] = mass[:, 1], mass[:, 3], mass[:, 4:6], mass[:, 0]
        m[:, 5], m[:, 6], m[:, 7] = bd["gravity_scale"], bd["linear_damping"], bd["angular_damping"]
        self._shapes, status = oracle32.make_shapes(scene.shape_defs)
        assert np.all(status == 0)
        fd = scene.fixture_defs
        mat = np.stack([fd["friction"], fd["restitution"], fd["restitution_threshold"]], 1).astype(np.float64)
        self.friction = mat[:, 0].copy()
        bt = np.ascontiguousarray(bd["type"], np.int32)
        self.btype = bt
        force = np.zeros((self.nb, 3))
        self.h = self.lib.ac_new(self.nb, _p(bt), _p(m), _p(force), self.nf, _p(self._shapes), _p(mat),
                                 float(gravity[0]), float(gravity[1]))
        self.steps = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ac_free(self.h)
            self.h = None

    def add_step(self, slots, lam0, params):
        sl = np.ascontiguousarray(slots, np.int32).reshape(-1, 10)
        l0 = np.ascontiguousarray(lam0, np.float64).reshape(-1, 4)
        pr = np.ascontiguousarray(params, np.float64)
        self.lib.ac_add_step(self.h, len(sl), _p(sl), _p(l0), _p(pr))
        self.steps += 1

    def forward(self, state0, friction=None):
        s0 = np.ascontiguousarray(state0, np.float64)
        fr = np.ascontiguousarray(self.friction if friction is None else friction, np.float64)
        out = np.zeros((self.nb, 6))
        mism = self.lib.ac_forward(self.h, _p(s0), _p(fr), _p(out))
        return out, mism

    def mismatch_bodies(self):
        """Bool per body: in a slot whose recomputed manifold (type, count or feature ids)
        differed from the recording in the last forward or backward pass."""
        flags = np.zeros(self.nb, np.int32)
        self.lib.ac_mismatch_bodies(self.h, _p(flags))
        return flags != 0

    def backward(self, state0, cot, friction=None):
        s0 = np.ascontiguousarray(state0, np.float64)
        fr = np.ascontiguousarray(self.friction if friction is None else friction, np.float64)
        c = np.ascontiguousarray(cot, np.float64)
        g = np.zeros((self.nb, 6))
        gf = np.zeros(self.nf)
        mism = self.lib.ac_backward(self.h, _p(s0), _p(fr), _p(c), _p(g), _p(gf))
        return g, gf, mism


def oracle_structure(oracle, scene, steps, dt=1.0 / 60.0, pos_iters=3):
    """Frozen structure of `steps` steps of the oracle's own run: touching contacts in contact-list
    order, warm-start carries matched by fixture pair and feature id, every constraint taking all
    position iterations."""
    ow = oracle.world()
    ow.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
    body = np.asarray(scene.fixture_body)
    prev, total, out = {}, 0, []
    for n in range(steps):
        ow.collide()
        fab, flags, man, _ = ow.contacts()
        touch = np.nonzero(flags & 1)[0]
        slots = np.zeros((len(touch), 10), np.int32)
        lam0 = np.zeros((len(touch), 4))
        cur = {}
        for i, k in enumerate(touch):
            fa, fb = int(fab[k, 0]), int(fab[k, 1])
            m = man[k]
            cnt = int(m["point_count"])
            ids = [int(m["points"]["id"][p]) for p in range(2)]
            src = [-1, -1]
            old = prev.get((fa, fb))
            if old is not None:
                for p in range(cnt):
                    for q in range(old[2]):
                        if ids[p] == old[1][q]:
                            src[p] = 2 * old[0] + q
                            break
            slots[i] = [body[fa], body[fb], fa, fb, int(m["type"]) | cnt << 8, src[0], src[1], pos_iters,
                        ids[0], ids[1]]
            lam0[i] = [m["points"]["normal_impulse"][0], m["points"]["tangent_impulse"][0],
                       m["points"]["normal_impulse"][1], m["points"]["tangent_impulse"][1]]
            cur[(fa, fb)] = (total + i, ids, cnt)
        params = np.array([dt, 0.0 if n == 0 else 1.0, 8, pos_iters, 1, 1])
        out.append((slots, lam0, params))
        prev = cur
        total += len(touch)
        ow.step(dt, 8, pos_iters)
    return out


def tape_structure(w):
    """The recorded structure of every step of a GPU tape: [(slots, lam0, params)] with the
    b2g_tape_export layouts (carry sources numbered over the whole batch)."""
    out = []
    for k in range(w.tape_info()[0]):
        slots, imp, offs, params = w.tape_export(k)
        out.append((np.asarray(slots).reshape(-1, 10), np.asarray(imp)[:, :4], np.asarray(params)))
    return out


def world_slice(batch, structure, wi):
    """World wi's part of a batch's recorded structure, with bodies, fixtures and carry sources
    renumbered to the world's own (the order of scenes.world_of): [(slots, lam0, params)]."""
    bsel = np.nonzero(batch.body_world == wi)[0]
    bmap = np.full(len(batch.body_world), -1, np.int64)
    bmap[bsel] = np.arange(len(bsel))
    fsel = np.nonzero(np.isin(batch.fixture_body, bsel))[0]
    fmap = np.full(len(batch.fixture_body), -1, np.int64)
    fmap[fsel] = np.arange(len(fsel))
    g2l = []  # per recorded slot of the batch: its number in the world's structure, or -1
    out, nloc = [], 0
    for slots, lam0, params in structure:
        keep = bmap[slots[:, 0]] >= 0
        loc = np.full(len(slots), -1, np.int64)
        loc[keep] = nloc + np.arange(int(keep.sum()))
        s = slots[keep].astype(np.int64).copy()
        s[:, 0], s[:, 1] = bmap[s[:, 0]], bmap[s[:, 1]]
        s[:, 2], s[:, 3] = fmap[s[:, 2]], fmap[s[:, 3]]
        assert np.all(s[:, :4] >= 0), "a slot joins two worlds"
        prev = np.concatenate(g2l) if g2l else np.zeros(0, np.int64)
        for col in (5, 6):
            src = s[:, col]
            has = src >= 0
            to = prev[src[has] // 2]
            assert np.all(to >= 0), "a carry source lies in another world"
            src[has] = 2 * to + src[has] % 2
        out.append((s.astype(np.int32), lam0[keep], params))
        g2l.append(loc)
        nloc += int(keep.sum())
    return out


def fd_check_world(oracle32, batch, wi, mass, state0, structure, g, gf, rng, h=1e-3, n_other=5, n_friction=4):
    """GPU gradients of world wi against central differences of the host f64 frozen model on the
    GPU's recorded structure, for the loss |c_body0 - goal|^2: body 0's initial velocity (3
    components), n_other random initial-velocity components of other bodies and n_friction
    random fixture frictions. Returns a dict: 'branch' False when the f64 model leaves the
    recorded branches (its gradient is then of another function); the relative error
    max |g - fd| / max |fd| over the checked parameters ('rel_err') and over the smooth ones
    ('rel_err_smooth': those whose central differences at h and h / 10 agree within 1e-4 of the
    world's scale; a clamp switching inside the step shows as a disagreement there)."""
    from b2g import scenes
    sc = scenes.world_of(batch, wi)
    bsel = np.nonzero(batch.body_world == wi)[0]
    fsel = np.nonzero(np.isin(batch.fixture_body, bsel))[0]
    host = HostAdjoint(oracle32, sc, mass[bsel])
    for sl, l0, pr in world_slice(batch, structure, wi):
        host.add_step(sl, l0, pr)
    s0 = np.asarray(state0, np.float64)[bsel]
    out, mism = host.forward(s0)
    body0 = int(batch.meta["body0"][wi] - bsel[0])
    goal = batch.meta["goal"][wi]
    if mism:
        return {"world": int(wi), "branch": False, "mismatches": int(mism)}

    def loss(s, fr=None):
        o, m = host.forward(s, fr)
        d = o[body0, 0:2] - goal
        return float(d @ d), m

    dyn = np.nonzero(sc.body_defs["type"] == 2)[0]
    params = [("s", body0, j) for j in (3, 4, 5)]
    others = [("s", int(b), j) for b in dyn if b != body0 for j in (3, 4, 5)]
    params += [others[i] for i in rng.choice(len(others), n_other, replace=False)]
    params += [("f", int(f), 0) for f in rng.choice(len(fsel), n_friction, replace=False)]
    rows = []
    branch = True
    for kind, i, j in params:
        fd = []
        for step in (h, h / 10):
            if kind == "s":
                sp, sm = s0.copy(), s0.copy()
                sp[i, j] += step
                sm[i, j] -= step
                (lp, mp), (lm, mm) = loss(sp), loss(sm)
            else:
                fp, fm = host.friction.copy(), host.friction.copy()
                fp[i] += step
                fm[i] -= step
                (lp, mp), (lm, mm) = loss(s0, fp), loss(s0, fm)
            branch &= mp == 0 and mm == 0
            fd.append((lp - lm) / (2 * step))
        gg = float(g[bsel][i, j]) if kind == "s" else float(gf[fsel][i])
        rows.append((gg, fd[0], fd[1]))
    rows = np.array(rows)
    scale = max(float(np.abs(rows[:, 1]).max()), 1e-12)
    # a clamp (friction cone, non-negative normal impulse, block-solver case) switching inside
    # [-h, h] makes the central difference an average over a kink; the two step sizes then disagree
    smooth = np.abs(rows[:, 1] - rows[:, 2]) <= 1e-4 * scale
    err = np.abs(rows[:, 0] - rows[:, 1]) / scale
    return {"world": int(wi), "branch": bool(branch), "scale": scale, "params": len(rows),
            "smooth": int(smooth.sum()), "rel_err": float(err.max()),
            "rel_err_smooth": float(err[smooth].max()) if smooth.any() else None,
            "fd_steps_disagree": float(np.abs(rows[:, 1] - rows[:, 2]).max() / scale)}
