"""The reference's continuous-collision tests as measurements over any stepper.

ccd_tests.rs fires a box at 200 m/s into a thin wall (with and without the bullet flag);
large_world_tests.rs fires a sphere bullet at a thin wall at several world offsets;
world_tuning_tests.rs shows the same bullet tunnelling with continuous collision off.
`load(scene)` returns a stepper with step(dt, substeps), body_arrays() and contact_count():
the oracle for the CPU tests, or the HIP world twinned with the oracle for the GPU tests.
"""
import numpy as np

from natsel_amd import scenes


def ccd_wall(load, bullet):
    """run (ccd_tests.rs:30-80): 30 steps of one substep; (final x, all finite, contacts)."""
    w = load(scenes.ccd_wall(bullet))
    for _ in range(30):
        w.step(1.0 / 60.0, 1)
    pos, rot, v, av = w.body_arrays()
    finite = bool(np.isfinite(v[1]).all() and np.isfinite(av[1]).all())
    return float(pos[1, 0]), finite, w.contact_count()


def check_ccd_wall(result):
    """The assertions of bullet / fast_body (ccd_tests.rs:60-116)."""
    x, finite, contacts = result
    assert x < 0.0 and finite and contacts == 0, result


def bullet_wall(load, offset, enable_continuous=True):
    """bullet_stops (large_world_tests.rs:169-200): the bullet's x after 30 steps of four
    substeps, relative to the offset."""
    w = load(scenes.bullet_wall(offset, enable_continuous))
    for _ in range(30):
        w.step(1.0 / 60.0, 4)
    return float(w.body_arrays()[0][1, 0] - offset)
