"""The C-ABI library loads on the CPU and exports every entry point include/b2g.h declares
(no compute call is made: there is no GPU here)."""
import os
import re
import subprocess

from helpers import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "b2g.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(b2g_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("b2g_world_create", "b2g_world_add", "b2g_world_step", "b2g_world_step_phase",
                     "b2g_collide_pairs", "b2g_broadphase_pairs", "b2g_world_get_contacts"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import b2g
    lib = os.path.join(ROOT, "box2d-rs_amd", "b2g", "libb2g.so")
    out = subprocess.check_output(["nm", "-D", "--defined-only", lib]).decode()
    exported = set(re.findall(r"\sT\s(b2g_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    assert b2g.version().startswith("b2g")


def test_struct_layouts_match_header():
    import b2g
    assert b2g.BODY_DEF.itemsize == 56
    assert b2g.SHAPE_DEF.itemsize == 160
    assert b2g.FIXTURE_DEF.itemsize == 32
    assert b2g.BODY_STATE.itemsize == 64
    assert b2g.MANIFOLD.itemsize == 72
