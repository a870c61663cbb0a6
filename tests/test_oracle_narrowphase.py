"""The oracle's narrow phase against the reference's own collision tests and closed forms.

Pins the sphere, capsule and hull-round routines, GJK (shape_distance) and the
separating-axis query with its edge route, which the golden world hashes (box stacks)
do not reach: the known answers of distance_tests.rs, shape_collision_tests.rs and
convex_manifold_tests.rs, the sweeps and brute-force oracles of
collision/manifold_robustness_tests.rs and collision/separating_axis_oracle_tests.rs,
and f64 closed forms (tests/narrowphase_cases.py)."""
import pytest

import narrowphase_cases as C
from oracle import pyoracle

_ALL = {**C.COLLIDE_CHECKS, **C.ORACLE_CHECKS,
        "rb_edge_axis_oracle": lambda n: C.rb_edge_axis_oracle(n, n.box_hull)}


@pytest.fixture(scope="module")
def narrow():
    return pyoracle.Narrow()


@pytest.mark.parametrize("name", sorted(_ALL))
def test_oracle_narrowphase(narrow, name):
    failures = _ALL[name](narrow)
    assert not failures, f"{name}: {failures[:3]}"
