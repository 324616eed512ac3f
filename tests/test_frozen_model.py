"""The library's templated step functions on the host, pinned against the oracle (CPU only).

box2d-rs_amd/csrc/b2g_step_t.h holds the island-solve functions (integrate, initialize, warm
start, velocity and position solves: src/private/dynamics/b2_island_private.rs:135-252,
b2_contact_solver_private.rs:113-730) as templates over the scalar; the adjoint evaluates them
on dual numbers. Instantiated in f64 and run as one solve in a given constraint order
(tests/host/frozen_model.h), they must reproduce the oracle's f64 island solver replayed in the
same order.
"""
import ctypes
import os

import numpy as np
import pytest

from helpers import ROOT

HOST_DIR = os.path.join(ROOT, "tests", "host")

DT = 1.0 / 60.0


def _lib():
    path = os.path.join(HOST_DIR, "_build", "libstep_check.so")
    if not os.path.exists(path):
        pytest.skip("tests/host not built")
    return ctypes.CDLL(path)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def solve_inputs(oracle64, scene, steps):
    """State and touching contacts of an f64 oracle world after `steps` steps and one collide."""
    ow = oracle64.world()
    ow.add(scene.body_defs, scene.fixture_body, scene.shape_defs, scene.fixture_defs)
    for _ in range(steps):
        ow.step(DT, 8, 3)
    ow.collide()
    st = ow.bodies()
    mass = ow.mass()
    fab, flags, man, mat = ow.contacts()
    touch = np.nonzero(flags & 1)[0]
    bd = scene.body_defs
    nb = len(st)
    props = np.zeros((nb, 10))
    props[:, 0], props[:, 1], props[:, 2] = mass[:, 1], mass[:, 3], mass[:, 0]
    props[:, 3:5] = mass[:, 4:6]
    props[:, 5] = bd["gravity_scale"]
    props[:, 6] = bd["linear_damping"]
    props[:, 7] = bd["angular_damping"]
    body = np.asarray(scene.fixture_body)
    sd = scene.shape_defs
    fa, fb = fab[touch, 0], fab[touch, 1]
    rad = lambda f: np.where(sd["type"][f] == 0, sd["radius"][f], 2.0 * 0.005)
    params = np.zeros((len(touch), 6))
    params[:, 0:3] = mat[touch]
    params[:, 4], params[:, 5] = rad(fa), rad(fb)
    cb = np.stack([body[fa], body[fb]], 1).astype(np.int32)
    return st, props, cb, params, man[touch]


def run_model(btype, props, st, cb, params, man, order, dt_ratio, iters):
    lib = _lib()
    nb, nc = len(st), len(cb)
    state = np.zeros((nb, 6))
    state[:, 0:2], state[:, 2], state[:, 3:5], state[:, 5] = st["c"], st["a"], st["v"], st["w"]
    hdr = np.stack([man["type"], man["point_count"]], 1).astype(np.int32)
    geom = np.zeros((nc, 8))
    geom[:, 0:2], geom[:, 2:4] = man["local_normal"], man["local_point"]
    geom[:, 4:6], geom[:, 6:8] = man["points"]["local_point"][:, 0], man["points"]["local_point"][:, 1]
    lam = np.zeros((nc, 4))
    lam[:, 0], lam[:, 1] = man["points"]["normal_impulse"][:, 0], man["points"]["tangent_impulse"][:, 0]
    lam[:, 2], lam[:, 3] = man["points"]["normal_impulse"][:, 1], man["points"]["tangent_impulse"][:, 1]
    step = np.array([DT, dt_ratio, 0.0, -10.0])
    it = np.array(iters, np.int32)
    runs = np.zeros(nc, np.int32)
    bt = np.ascontiguousarray(btype, np.int32)
    cbc = np.ascontiguousarray(cb, np.int32)
    lib.sc_solve_f64(nb, _p(bt), _p(np.ascontiguousarray(props)), _p(state), nc, _p(cbc),
                     _p(np.ascontiguousarray(params)), _p(hdr), _p(geom), _p(lam), _p(np.ascontiguousarray(order)),
                     _p(step), _p(it), _p(runs))
    return state, lam, runs


@pytest.mark.parametrize("name,steps,block", [("heap", 30, True), ("pyramid", 120, True), ("heap", 90, False)])
def test_templated_solve_matches_oracle_f64(oracle64, name, steps, block):
    from b2g import scenes
    sc = scenes.heap_world(300, seed=5, half_extent=10.0) if name == "heap" else scenes.pyramid()
    st, props, cb, params, man = solve_inputs(oracle64, sc, steps)
    nc = len(cb)
    assert nc > 20
    order = np.random.default_rng(steps).permutation(nc).astype(np.int32)
    dt_ratio = 1.0 if steps > 0 else 0.0
    btype = sc.body_defs["type"]
    state, lam, runs = run_model(btype, props, st, cb, params, man, order, dt_ratio, [8, 3, 1, int(block)])
    ref_st, ref_man, _, _ = oracle64.solver_replay(btype, props, st, cb, params, man, order, DT, dt_ratio,
                                                   (0.0, -10.0), 8, 3, True, block)
    dyn = btype != 0
    got = {"c": state[:, 0:2], "a": state[:, 2], "v": state[:, 3:5], "w": state[:, 5]}
    for f in ("c", "a", "v", "w"):
        err = np.abs(got[f][dyn] - ref_st[f][dyn]).max()
        assert err <= 1e-9 * max(1.0, np.abs(ref_st[f][dyn]).max()), (f, err)
    ref_lam = np.stack([ref_man["points"]["normal_impulse"][:, 0], ref_man["points"]["tangent_impulse"][:, 0],
                        ref_man["points"]["normal_impulse"][:, 1], ref_man["points"]["tangent_impulse"][:, 1]], 1)
    assert np.abs(lam - ref_lam).max() <= 1e-9 * max(1.0, np.abs(ref_lam).max())
    assert runs.max() >= 1
