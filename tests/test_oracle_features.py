"""The oracle honours every BodyDef and ShapeDef field the C ABI carries.

natsel_amd.scenes.def_features names one body per field; the checks are what the field
means in the reference: motion locks (integrate_positions_range, solver_integration.rs:
377-426, and the locked inertia of body_api.rs:792) keep a locked coordinate exactly fixed;
a filter group below zero or masks that exclude each other's category
(shape_runtime.rs:285-292) and sensors (contact_creation.rs:187) make no contact; a sensor
takes no part in continuous collision (continuous.rs:239-246, :398-401), so a 60 m/s ball
crosses a thin sensor wall; rolling resistance (contact_solver.rs:241) stops a rolling
sphere that a frictional sphere without it keeps rolling; allow_fast_rotation lifts the
angular clamp of 0.25 pi per substep."""
import numpy as np
import pytest

from natsel_amd import scenes
from oracle import pyoracle


@pytest.fixture(scope="module")
def run():
    sc = scenes.def_features()
    w = pyoracle.load_scene(sc)
    t = sc.tags
    hist = []
    pairs = set()
    for k in range(240):
        w.step(sc.dt, sc.substeps)
        pos, rot, v, om = w.body_arrays()
        hist.append((pos.copy(), rot.copy(), v.copy(), om.copy()))
        a, b = w.contact_shapes()
        pairs |= {(int(x), int(y)) for x, y in zip(a, b) if x >= 0}
    return sc, t, hist, pairs


def _shape_of(sc, body):
    return [i for i, s in enumerate(sc.shapes) if s.body == body]


def test_motion_locks(run):
    sc, t, hist, _ = run
    for k in range(3):
        b = t[f"planar_{k}"]
        # the locks zero the velocity before the positions integrate; relax (after) may leave a
        # locked velocity component non-zero, but the locked coordinates never move
        for pos, rot, v, om in hist:
            assert pos[b, 2] == 0.0 and rot[b, 0] == 0.0 and rot[b, 1] == 0.0
    b = t["fixed_rotation"]
    for pos, rot, v, om in hist:
        assert tuple(rot[b]) == (0.0, 0.0, 0.0, 1.0) and not om[b].any()
    assert hist[-1][0][b, 0] > -3.0 + 0.05  # it slid


def test_filters_and_sensors_make_no_contact(run):
    sc, t, hist, pairs = run
    for x, y in (("group_a", "group_b"), ("mask_a", "mask_b")):
        sx, sy = _shape_of(sc, t[x])[0], _shape_of(sc, t[y])[0]
        assert (min(sx, sy), max(sx, sy)) not in pairs, f"{x} and {y} touched"
        # the upper box fell through the lower one to the ground
        assert abs(hist[-1][0][t[y], 1] - 0.5) < 0.02
    sensors = [i for i, s in enumerate(sc.shapes) if s.is_sensor]
    assert not any(a in sensors or b in sensors for a, b in pairs)
    assert abs(hist[-1][0][t["sensor_drop"], 1] - 0.25) < 0.02  # fell through the sensor box to the ground
    # continuous collision is on and the ball is fast, but the wall is a sensor: it crosses
    assert hist[-1][0][t["bullet_ball"], 0] > 20.0


def test_rolling_resistance_stops_rolling(run):
    sc, t, hist, _ = run
    speed = lambda k, name: float(np.linalg.norm(hist[k][2][t[name]]))  # noqa: E731
    assert speed(179, "slide_sphere") > 1.5
    assert speed(179, "roll_sphere") < 0.5 * speed(179, "slide_sphere")
    assert speed(239, "roll_capsule") < 0.5


def test_fast_rotation(run):
    sc, t, hist, _ = run
    clamp = 0.25 * np.pi * 60.0 * 4  # 0.25 pi per substep of 1/240 s
    assert abs(hist[0][3][t["clamped_spin"], 1]) <= clamp * 1.001
    assert abs(hist[0][3][t["fast_spin"], 1]) > 1.2 * clamp
