"""The GPU narrowphase code (b2g_collide.h), built for the CPU, agrees bit-for-bit with the
oracle on random shape pairs near contact: point counts, manifold types, contact-feature ids
and geometry (all five contact classes of the reference registry)."""
import ctypes
import os
import subprocess
import zlib

import numpy as np
import pytest

from helpers import PAIR_KINDS, ROOT, compare_manifolds, extent, near_poses, random_shape_defs

HOST_DIR = os.path.join(ROOT, "tests", "host")


@pytest.fixture(scope="module")
def hostlib():
    lib_path = os.path.join(HOST_DIR, "_build", "libcollide_check.so")
    subprocess.check_call(["make", "-s", "-C", HOST_DIR])
    lib = ctypes.CDLL(lib_path)
    lib.hc_collide.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 5
    lib.hc_aabb.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 3
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("kinds", PAIR_KINDS, ids=["-".join(k) for k in PAIR_KINDS])
def test_host_collide_matches_oracle(oracle, hostlib, kinds):
    rng = np.random.default_rng(zlib.crc32("-".join(kinds).encode()) & 0xffff)
    n = 4000
    sda = random_shape_defs(rng, kinds[0], n, oracle.dt["shape_def"])
    sdb = random_shape_defs(rng, kinds[1], n, oracle.dt["shape_def"])
    sa, st_a = oracle.make_shapes(sda)
    sb, st_b = oracle.make_shapes(sdb)
    keep = (st_a == 0) & (st_b == 0)
    sa, sb, sda, sdb = sa[keep], sb[keep], sda[keep], sdb[keep]
    xa, xb = near_poses(rng, extent(sda), extent(sdb), len(sa), oracle)
    ref = oracle.collide(sa, xa, sb, xb)
    got = np.zeros(len(sa), oracle.dt["manifold"])
    hostlib.hc_collide(len(sa), _p(sa), _p(xa), _p(sb), _p(xb), _p(got))
    touching = int((ref["point_count"] > 0).sum())
    assert touching > len(sa) // 10, "too few touching pairs to be a meaningful check"
    bad = compare_manifolds(got, ref, rtol=0.0, atol=0.0)
    assert not bad, bad[:5]


def test_host_aabb_matches_oracle(oracle, hostlib):
    rng = np.random.default_rng(7)
    for kind in ("circle", "polygon", "edge"):
        sd = random_shape_defs(rng, kind, 500, oracle.dt["shape_def"])
        s, st = oracle.make_shapes(sd)
        s = s[st == 0]
        xf = np.concatenate([rng.uniform(-50, 50, (len(s), 2)).astype(np.float32),
                             oracle.rot(rng.uniform(-7, 7, len(s)).astype(np.float32))], 1)
        ref = oracle.shape_aabb(s, xf)
        got = np.zeros((len(s), 4), np.float32)
        hostlib.hc_aabb(len(s), _p(s), _p(np.ascontiguousarray(xf)), _p(got))
        assert np.array_equal(got, ref)
