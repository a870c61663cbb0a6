"""The oracle's continuous collision: time of impact and the swept pass of fast bodies.

time_of_impact (time_of_impact.rs:406-598) is pinned three ways: the reference's own known
answers (distance_tests.rs:182-284: state, normal and iteration counts of a box sweep, a
receding sweep and a rotating sweep); closed forms for translating spheres and boxes, where
the separation at the returned fraction is exact; and, for rotating boxes and capsules
approaching a large wall, the separation from the wall plane sampled over the sweep: the
separation must reach the target (linear slop short of touching) at the fraction and never
dip below it earlier. A wrong sign or a swapped operand in a sweep transform or a separation
function moves the fraction off those values.

The swept pass (continuous.rs:337-621) is pinned by the reference's world tests: fast boxes
and bullets do not tunnel through thin walls at any world offset, tunnelling returns with
continuous collision off, and the balls dropped from 20 m and 40 m keep their apexes
(test_oracle_restitution.py).
"""
import math

import numpy as np
import pytest

import ccd_cases as cc

I4 = (0.0, 0.0, 0.0, 1.0)
Z3 = (0.0, 0.0, 0.0)
SLOP = 0.005
TOL = 0.25 * SLOP


def _toi(*args, **kw):
    from oracle import pyoracle

    return pyoracle.time_of_impact(*args, **kw)


def _load(scene):
    from oracle import pyoracle

    return pyoracle.load_scene(scene)


def box_points(hx, hy=None, hz=None):
    """make_box (distance_tests.rs:14-23)."""
    hy = hx if hy is None else hy
    hz = hx if hz is None else hz
    return np.array([[-hx if i & 1 else hx, -hy if i & 2 else hy, -hz if i & 4 else hz] for i in range(8)], np.float32)


def quat_axis_angle(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    s, c = math.sin(angle / 2), math.cos(angle / 2)
    return tuple(float(np.float32(x)) for x in (a[0] * s, a[1] * s, a[2] * s, c))


def rot(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def nlerp(q1, q2, t):
    q1, q2 = np.asarray(q1, np.float64), np.asarray(q2, np.float64)
    if q1.dot(q2) < 0:
        q1 = -q1
    q = (1 - t) * q1 + t * q2
    return q / np.linalg.norm(q)


def world_points(points, sweep, t):
    """Points of a proxy at time t of its sweep (distance.rs:817-823), in float64."""
    lc, c1, c2, q1, q2 = (np.asarray(x, np.float64) for x in sweep)
    c = (1 - t) * c1 + t * c2
    return c + (np.asarray(points, np.float64) - lc) @ rot(nlerp(q1, q2, t)).T


# ---------------------------------------------------------------------------------
# known answers of the reference (distance_tests.rs:182-284)

STILL = (Z3, Z3, Z3, I4, I4)


def test_translate_known_answer():
    out = _toi(box_points(1.0), 0.0, STILL, box_points(1.0), 0.0, (Z3, (5.0, 0.0, 0.0), Z3, I4, I4))
    assert out["state"] == "hit"
    assert tuple(out["normal"]) == (1.0, 0.0, 0.0)
    assert (out["distance_iterations"], out["push_back_iterations"], out["root_iterations"]) == (2, 1, 2)
    # the faces meet at t = 3/5; the sweep stops the target (one linear slop) short of that
    assert abs(float(out["fraction"]) - (3.0 - SLOP) / 5.0) <= TOL / 5.0


def test_recede_known_answer():
    out = _toi(box_points(1.0), 0.0, STILL, box_points(1.0), 0.0, (Z3, (5.0, 0.0, 0.0), (10.0, 0.0, 0.0), I4, I4))
    assert out["state"] == "separated" and out["fraction"] == 1.0
    assert out["distance_iterations"] == 1


def test_rotate_known_answer():
    q = quat_axis_angle((0.0, 0.0, 1.0), 0.785398163)
    out = _toi(box_points(1.0), 0.0, STILL, box_points(1.0), 0.0, (Z3, (4.0, 0.0, 0.0), Z3, I4, q))
    assert out["state"] == "hit"
    assert tuple(out["normal"]) == (1.0, 0.0, 0.0)
    assert out["distance_iterations"] == 2 and out["root_iterations"] == 4


# ---------------------------------------------------------------------------------
# closed forms


@pytest.mark.parametrize("seed", range(8))
def test_spheres_closed_form(seed):
    """Two spheres, both translating: the centre distance at the fraction is the target
    r_a + r_b - slop within the tolerance, and no earlier time is closer."""
    rng = np.random.default_rng(seed)
    ra, rb = rng.uniform(0.05, 0.5, 2)
    ca1, ca2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
    cb2 = ca2 + rng.normal(size=3) * 0.1
    cb1 = cb2 + rng.normal(size=3) * 8.0
    sa = (Z3, ca1, ca2, I4, I4)
    sb = (Z3, cb1, cb2, I4, I4)
    out = _toi(np.zeros((1, 3)), ra, sa, np.zeros((1, 3)), rb, sb)
    target = max(SLOP, float(np.float32(ra) + np.float32(rb)) - SLOP)
    d = lambda t: float(np.linalg.norm(world_points([[0, 0, 0]], sb, t) - world_points([[0, 0, 0]], sa, t)))  # noqa: E731
    if d(1.0) > target + TOL and min(d(t) for t in np.linspace(0, 1, 2001)) > target + TOL:
        assert out["state"] == "separated"
        return
    assert out["state"] == "hit", out
    f = float(out["fraction"])
    assert abs(d(f) - target) <= TOL + 1e-5, (d(f), target)
    assert min(d(t) for t in np.linspace(0.0, f, 400)) >= target - TOL - 1e-5


@pytest.mark.parametrize("seed", range(8))
def test_boxes_translation_closed_form(seed):
    """Two axis-aligned boxes, the second translating: their separation is the length of
    the positive parts of |dc| - (h_a + h_b) per axis."""
    rng = np.random.default_rng(100 + seed)
    ha, hb = rng.uniform(0.2, 1.0, 3), rng.uniform(0.2, 1.0, 3)
    c2 = (ha + hb) * rng.choice([-1, 1], 3) * rng.uniform(0.3, 0.9, 3)
    c1 = c2 + rng.normal(size=3) * 6.0
    sb = (Z3, c1, c2, I4, I4)
    out = _toi(box_points(*ha), 0.0, STILL, box_points(*hb), 0.0, sb)

    def sep(t):
        c = (1 - t) * c1 + t * c2
        return float(np.linalg.norm(np.maximum(np.abs(c) - (ha.astype(np.float32) + hb.astype(np.float32)), 0.0)))

    if sep(0.0) <= 0.0:
        assert out["state"] == "overlapped"
        return
    assert out["state"] == "hit", out
    f = float(out["fraction"])
    assert abs(sep(f) - SLOP) <= TOL + 1e-5, (sep(f), f)
    assert min(sep(t) for t in np.linspace(0.0, f, 400)) >= SLOP - TOL - 1e-5


def _wall_gap(points, radius, sweep, t, face_x):
    return face_x - float(world_points(points, sweep, t)[:, 0].max()) - radius


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("kind", ["box", "capsule", "sphere"])
def test_rotating_against_wall(kind, seed):
    """A box, capsule or sphere spinning as it flies at a large static wall (face x = 0):
    its gap to the wall plane (core gap less its radius) reaches the target at the fraction
    and not before."""
    rng = np.random.default_rng(1000 * len(kind) + seed)
    wall = box_points(0.5, 50.0, 50.0)
    sa = (Z3, (0.5, 0.0, 0.0), (0.5, 0.0, 0.0), I4, I4)  # its face toward the mover is x = 0
    if kind == "box":
        pts, r = box_points(*rng.uniform(0.1, 0.5, 3)), 0.0
    elif kind == "capsule":
        half = rng.uniform(0.1, 0.6)
        pts, r = np.array([[0.0, -half, 0.0], [0.0, half, 0.0]], np.float32), float(rng.uniform(0.05, 0.3))
    else:
        pts, r = np.zeros((1, 3), np.float32), float(rng.uniform(0.05, 0.3))
    # the shape sits off the body origin and the body's centre of mass elsewhere again
    pts = (pts + rng.uniform(-0.3, 0.3, 3)).astype(np.float32)
    lc = tuple(float(np.float32(x)) for x in rng.uniform(-0.3, 0.3, 3))
    q1 = quat_axis_angle(rng.normal(size=3), rng.uniform(0, math.pi))
    q2 = quat_axis_angle(rng.normal(size=3), rng.uniform(0, math.pi))
    c1 = (-rng.uniform(2.0, 6.0), rng.uniform(-1, 1), rng.uniform(-1, 1))
    c2 = (rng.uniform(-0.2, 0.5), rng.uniform(-1, 1), rng.uniform(-1, 1))
    sb = (lc, c1, c2, q1, q2)
    out = _toi(wall, 0.0, sa, pts, r, sb)
    target = max(SLOP, r - SLOP)
    gap = lambda t: _wall_gap(pts, 0.0, sb, t, 0.0)  # noqa: E731
    samples = np.linspace(0.0, 1.0, 4001)
    if min(gap(t) for t in samples) > target + TOL:
        assert out["state"] == "separated", out
        return
    assert out["state"] == "hit", out
    f = float(out["fraction"])
    assert abs(gap(f) - target) <= TOL + 1e-4, (gap(f), target, f)
    assert min(gap(t) for t in samples[samples <= f]) >= target - TOL - 1e-4


def test_max_fraction_bounds_the_sweep():
    """A hit later than max_fraction reports separated at max_fraction; an earlier bound
    than the hit does not change the hit."""
    sb = (Z3, (5.0, 0.0, 0.0), Z3, I4, I4)
    hit = _toi(box_points(1.0), 0.0, STILL, box_points(1.0), 0.0, sb)
    f = float(hit["fraction"])
    early = _toi(box_points(1.0), 0.0, STILL, box_points(1.0), 0.0, sb, max_fraction=0.5 * f)
    assert early["state"] == "separated" and float(early["fraction"]) == np.float32(0.5 * f)
    late = _toi(box_points(1.0), 0.0, STILL, box_points(1.0), 0.0, sb, max_fraction=0.9)
    assert late["state"] == "hit" and abs(float(late["fraction"]) - f) <= TOL / 5.0


def test_overlapped_start():
    out = _toi(box_points(1.0), 0.0, STILL, box_points(1.0), 0.0, (Z3, (1.5, 0.0, 0.0), (5.0, 0.0, 0.0), I4, I4))
    assert out["state"] == "overlapped" and out["fraction"] == 0.0


# ---------------------------------------------------------------------------------
# the swept pass in the world (ccd_tests.rs, large_world_tests.rs, world_tuning_tests.rs)


@pytest.mark.parametrize("bullet", [False, True])
def test_fast_box_does_not_tunnel(bullet):
    cc.check_ccd_wall(cc.ccd_wall(_load, bullet))


@pytest.mark.parametrize("offset", [0.0, 1.0e9, -1.0e9, 1.0e12])
def test_bullet_stops(offset):
    assert cc.bullet_wall(_load, offset) < 5.0


def test_tunnels_without_continuous():
    assert cc.bullet_wall(_load, 0.0, enable_continuous=False) > 5.0


def test_fast_and_bullets_scene_runs_the_sweep():
    """Spheres at 300 m/s through a field of cubes between two walls: the sweep runs for
    fast bodies and bullets, finds impacts, and no sphere leaves through a wall."""
    from natsel_amd import scenes

    w = _load(scenes.fast_and_bullets())
    for _ in range(30):
        w.step(1.0 / 60.0, 4)
    s = w.stats()
    assert s["fast_bodies"] > 0 and s["toi_calls"] > 0 and s["toi_hits"] > 0, s
    pos = w.body_arrays()[0]
    behind = (np.abs(pos[3:, 0]) > 16.0) & (np.abs(pos[3:, 2]) < 20.0) & (pos[3:, 1] < 8.0)
    assert not behind.any(), pos[3:][behind]  # none passed through a wall (some go round its ends)
