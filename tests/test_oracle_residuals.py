"""The oracle's per-colour residual sweep (World::colour_residuals, the checker of the GPU's
nsl_world_get_residuals). The reference has no such diagnostic, so the sweep is pinned by
what the mathematics fixes:

- for a single scalar constraint, one Gauss-Seidel relax is idempotent: after relax the
  point's normal velocity plus bias is zero (or its impulse is clamped at zero), so a second
  relax changes nothing beyond rounding. A sphere falling onto the ground, touching it and
  resting (one point, no angular coupling of the normal under the centre) has a residual of
  rounding size at every step;
- two spheres stacked on the ground couple two constraints of different colours, where one
  sweep is not a fixed point: the residual is of the order of the impulses while they
  settle;
- the sweep leaves the world as the step made it (the hash with the sweep on equals the hash
  with it off), counts each colour's constraints, and its sum is at least its largest term.
"""
import numpy as np

from natsel_amd.scenes import BOX, DYNAMIC, SPHERE, STATIC, BodySpec, Scene, ShapeSpec
from natsel_amd import scenes
from oracle import pyoracle

RADIUS, DENSITY = 0.5, 1000.0
MASS = DENSITY * 4.0 / 3.0 * np.pi * RADIUS ** 3


def spheres_on_ground(count: int, drop: float) -> Scene:
    sc = Scene(name=f"spheres_{count}", enable_sleep=False, enable_continuous=False)
    g = sc.add_body(BodySpec(STATIC, (0.0, -1.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (20.0, 1.0, 20.0)))
    for i in range(count):
        b = sc.add_body(BodySpec(DYNAMIC, (0.0, RADIUS + drop + 2 * RADIUS * i, 0.0)))
        sc.add_shape(ShapeSpec(b, SPHERE, (0.0, 0.0, 0.0, RADIUS), density=DENSITY))
    return sc


def impulse_scale(sc: Scene) -> float:
    """A resting sphere's normal impulse per substep, m g h."""
    return MASS * 10.0 * sc.dt / sc.substeps


def run(sc: Scene, steps: int, residuals: bool = True):
    w = pyoracle.load_scene(sc)
    w.set_residuals(residuals)
    out = []
    for _ in range(steps):
        w.step(sc.dt, sc.substeps)
        out.append(w.colour_residuals() if residuals else None)
    return w, out


def test_single_constraint_sweep_is_idempotent():
    sc = spheres_on_ground(1, drop=0.3)
    _, res = run(sc, 120)
    touched = [r for r in res if sum(c["constraints"] for c in r) > 0]
    assert len(touched) > 60  # it lands early and rests
    tol = 1e-5 * impulse_scale(sc)
    for r in touched:
        for c in r:
            assert c["max_normal"] <= tol and c["sum_normal"] <= tol
            assert c["max_friction"] <= tol


def test_coupled_constraints_leave_a_residual():
    sc = spheres_on_ground(2, drop=0.2)
    _, res = run(sc, 90)
    worst = max(c["max_normal"] for r in res for c in r)
    assert worst > 1e-3 * impulse_scale(sc)
    for r in res:
        for c in r:
            assert c["sum_normal"] >= c["max_normal"] >= 0.0
            assert c["max_friction"] >= 0.0
            if c["constraints"] == 0:
                assert c["sum_normal"] == 0.0 and c["max_normal"] == 0.0


def test_sweep_leaves_the_world_unchanged_and_counts_colours():
    sc = scenes.pile(300)
    w_on, res = run(sc, 150, True)
    w_off, _ = run(sc, 150, False)
    assert w_on.world_state_hash() == w_off.world_state_hash()
    color, _, alive = w_on.contact_colors()
    touching = color[(alive != 0) & (color >= 0) & (color < 23)]
    counts = np.bincount(touching, minlength=23)
    assert [c["constraints"] for c in res[-1]] == counts.tolist()
    assert sum(counts) > 100
