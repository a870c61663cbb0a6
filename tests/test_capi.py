"""The C ABI library loads and exports every entry point include/natsel_gpu.h declares
(no GPU calls)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "natsel_gpu.h")).read()
    return sorted(set(re.findall(r"\b(nsl_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "nsl_world_step" in syms and "nsl_create_body" in syms
    assert len(syms) >= 15


def test_library_exports_every_declared_symbol():
    from natsel_amd import capi

    path = capi.library_path()
    if not os.path.exists(path):
        pytest.fail("libnatsel_gpu.so not built (run __graft_entry__.build())")
    lib = capi.load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(capi.SIGNATURES) == set(declared_symbols())


def test_default_options_match_reference_defaults():
    from natsel_amd import capi

    lib = capi.load_library()
    o = capi.Options()
    lib.nsl_default_options(o)
    assert tuple(o.gravity) == (0.0, -10.0, 0.0)
    assert (o.contact_hertz, o.contact_damping_ratio, o.contact_speed) == (30.0, 10.0, 3.0)
    assert o.max_linear_speed == 400.0 and o.restitution_iterations == 2


# ctypes structure -> C typedef of include/natsel_gpu.h
C_STRUCTS = {
    "Options": "nsl_options", "BodyDef": "nsl_body_def", "Material": "nsl_material", "Profile": "nsl_profile",
    "KernelStats": "nsl_kernel_stats", "ToiInput": "nsl_toi_input", "ToiOutput": "nsl_toi_output",
    "RevoluteJointDef": "nsl_revolute_joint_def", "BatchStats": "nsl_batch_stats", "ShapeGeometry": "nsl_shape_geometry",
    "CollideCache": "nsl_collide_cache", "CollideQuery": "nsl_collide_query", "LocalManifold": "nsl_local_manifold",
    "Tuning": "nsl_tuning", "BatchTuning": "nsl_batch_tuning", "ColourResidual": "nsl_colour_residual",
}


def test_struct_layouts_match_the_header(tmp_path):
    """sizeof and every field's offsetof, printed by a C program compiled against the header,
    equal the ctypes structures' (a binding that drifts from the header fails here)."""
    import ctypes
    import subprocess

    from natsel_amd import capi

    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "natsel_gpu.h"', "int main(void) {"]
    for py, c in C_STRUCTS.items():
        lines.append(f'  printf("{py} size %zu\\n", sizeof({c}));')
        for name, _ in getattr(capi, py)._fields_:
            lines.append(f'  printf("{py} {name} %zu\\n", offsetof({c}, {name}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    want = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        py, field, value = line.split()
        want[(py, field)] = int(value)
    got = {}
    for py in C_STRUCTS:
        cls = getattr(capi, py)
        got[(py, "size")] = ctypes.sizeof(cls)
        for name, _ in cls._fields_:
            got[(py, name)] = getattr(cls, name).offset
    assert got == want, sorted(k for k in want if got.get(k) != want[k])
    # every typedef'd struct of the header has a binding
    text = open(os.path.join(ROOT, "include", "natsel_gpu.h")).read()
    declared = set(re.findall(r"}\s*(nsl_[a-z_]+)\s*;", text))
    assert declared == set(C_STRUCTS.values()), declared ^ set(C_STRUCTS.values())


def test_collide_rejects_malformed_queries():
    """nsl_collide validates on the host before any device work: a pair outside the register
    table's order and caches whose indices do not fit the shapes are NSL_ERR_INVALID."""
    import ctypes

    from natsel_amd import capi

    lib = capi.load_library()
    q = (capi.CollideQuery * 1)()
    out = (capi.LocalManifold * 1)()

    def status():
        return lib.nsl_collide(0, q, out, 1)

    q[0].a.type, q[0].b.type = capi.SHAPE_SPHERE, capi.SHAPE_HULL  # reversed order
    assert status() == -1 and b"register table" in lib.nsl_last_error()
    q[0].a.type, q[0].b.type = capi.SHAPE_HULL, capi.SHAPE_HULL
    q[0].cache.kind, q[0].cache.index_a[0] = 2, 6  # FaceAxisA, face 6
    assert status() == -1 and b"face index" in lib.nsl_last_error()
    q[0].cache.kind, q[0].cache.index_a[0], q[0].cache.index_b[0] = 4, 23, 0  # edge 23
    assert status() == -1 and b"edge index" in lib.nsl_last_error()
    q[0].b.type = capi.SHAPE_CAPSULE
    q[0].cache.kind = 4  # simplex count 4
    assert status() == -1 and b"count" in lib.nsl_last_error()
    q[0].cache.kind, q[0].cache.index_b[0] = 1, 2  # capsule has support points 0 and 1
    assert status() == -1 and b"support points" in lib.nsl_last_error()
    assert lib.nsl_collide(0, q, out, -1) == -1
    assert ctypes.sizeof(capi.CollideQuery) == 168
