"""Each planted mistake of the oracle's narrow phase fails at least one narrow-phase check.

oracle/ref_plant.h numbers twelve plausible slips (a radius term with the wrong sign or
dropped, a reversed normal, swapped GJK weights, an edge contact point off its midpoint,
an edge axis interpolated with the wrong sign, the reduction's fourth point searched on
the wrong side, an inverted cache tolerance, a sign slip in segment distance, a point past a
segment's end snapped to the wrong endpoint). The narrow phase is built once per slip with
only that slip on, and the checks of tests/narrowphase_cases.py must catch every one: the
pins test the arithmetic, not just that the code runs."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import narrowphase_cases as C
from oracle import pyoracle

PLANTS = list(range(1, 13))
ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.fixture(scope="module")
def planted_libs(tmp_path_factory):
    out = tmp_path_factory.mktemp("plants")

    def build(n):
        path = str(out / f"libnsl_narrow_{n}.so")
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "narrow", f"PLANT={n}", f"NARROW_OUT={path}"], check=True)
        return n, path

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return dict(ex.map(build, PLANTS))


@pytest.mark.parametrize("plant", PLANTS)
def test_planted_mistake_is_caught(planted_libs, plant):
    nar = pyoracle.Narrow(planted_libs[plant])
    results = C.run_all(nar)
    caught = sorted(name for name, failures in results.items() if failures)
    print(f"plant {plant} caught by {caught}")
    assert caught, f"planted mistake {plant} passes every narrow-phase check"
