"""The oracle against the reference's restitution contract (CPU only).

Each test asserts what the same-named test of src/tests/simulation/restitution_phase_tests.rs
asserts, with its tolerances. Together they pin the oracle's restitution stage, the speculative
contact, sphere/box manifolds and the relax stage: a dropped term or a sign error there moves
a measured coefficient far outside these bounds.
"""
import numpy as np
import pytest

import restitution_cases as rc


def _load(scene):
    from oracle import pyoracle

    return pyoracle.load_scene(scene)


@pytest.mark.parametrize("e", [0.25, 0.5, 0.9])
def test_phase(e):
    """phase (restitution_phase_tests.rs:101-132): spread and mean error at most 0.05."""
    m = [rc.ground_bounce(_load, e, g) for g in rc.phase_gaps()]
    assert max(m) - min(m) <= 0.05
    assert abs(float(np.mean(m)) - e) <= 0.05


@pytest.mark.parametrize("propagate", [False, True])
def test_supported(propagate):
    """supported (restitution_phase_tests.rs:137-162): at most e + 0.01, at least 0.5 e
    (0.75 e with propagation)."""
    floor = 0.75 if propagate else 0.5
    for e in (0.5, 0.9):
        for n in range(4):
            m = rc.supported_bounce(_load, e, n, propagate)
            assert floor * e <= m <= e + 0.01, (e, n, m)


def test_threshold():
    """threshold / threshold_above (restitution_phase_tests.rs:164-198)."""
    assert rc.threshold_bounce(_load, 10.0, 1.0) <= 0.01
    assert rc.threshold_bounce(_load, 1.0, rc.IMPACT) / rc.IMPACT > 0.85


def test_head_on():
    """head_on (restitution_phase_tests.rs:241-266): coefficient within 0.02 for every
    mass ratio, momentum conserved to 1e-4."""
    for density in (1.0, 10.0, 100.0):
        for e in (0.0, 0.25, 0.5, 0.75, 1.0):
            ratio, drift = rc.head_on(_load, e, density)
            assert abs(ratio - e) <= 0.02, (density, e, ratio)
            assert abs(drift) <= 1e-4, (density, e, drift)


def test_flat_landing():
    """flat_landing (restitution_phase_tests.rs:857-885): coefficient within 0.01, spin at
    most 0.03, at four substeps."""
    for e in (0.5, 0.9):
        ratio, spin = rc.flat_bounce(_load, e)
        assert abs(ratio - e) <= 0.01 and spin <= 0.03, (e, ratio, spin)


def test_spin_along_axis():
    """spin_along_axis (restitution_phase_tests.rs:921-933): coefficient within 0.05 of 0.9."""
    for spin in (0.0, 0.5, 1.0, 2.0):
        ratio, _ = rc.flat_bounce(_load, 0.9, spin=spin)
        assert abs(ratio - 0.9) <= 0.05, (spin, ratio)


@pytest.mark.parametrize("h", [40.0, 20.0, 10.0])
def test_drop_apexes(h):
    """drop_apexes (restitution_phase_tests.rs:962-996): six apexes, none above the first
    by more than 0.1 %, the first not above the drop height, the last at least 80 % of it.
    From 20 m and more the ball moves farther per step than its fat AABB margin, so the
    continuous sweep is what lands it on the ground."""
    rc.check_apexes(rc.drop_apexes(_load, h), h)


def test_overshoot():
    """overshoot (restitution_phase_tests.rs:382-434): an elastic cube never returns more
    speed than it received; swept onto the floor (safety factor 0.1) it stays under its start
    height and gains no energy."""
    rc.check_overshoot(rc.overshoot(_load, False), rc.overshoot(_load, True))


def test_resting_stack():
    """resting_stack (restitution_phase_tests.rs:1001-1049): drift under 1 cm, peak speed
    under 0.05 m/s."""
    drift, peak = rc.resting_stack(_load)
    assert drift < 0.01 and peak < 0.05, (drift, peak)
