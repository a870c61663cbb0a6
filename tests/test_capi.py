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
