"""The oracle's revolute joint against closed forms and the reference's joint tests (CPU).

revolute_joint_tests.rs only checks replay, so the joint's arithmetic is pinned here by what
mechanics fixes: a dropped term or wrong sign in the point, collinearity, motor, limit or
spring rows moves one of these measurements far outside its bound.
"""
import math

import numpy as np

import joint_cases as jc
from natsel_amd import scenes

I_CM_OVER_I_PIVOT = (0.4 * 0.4 / 6.0) / (0.4 * 0.4 / 6.0 + 0.25)


def _load(scene):
    from oracle import pyoracle

    return pyoracle.load_scene(scene)


def test_pendulum_period_pivot_and_energy():
    """Released level, the hinged box swings with the compound pendulum's period (elliptic
    integral, within 0.5 %), keeps its pivot within 2 mm and climbs back to its start height
    without gaining energy."""
    _, r = jc.run_pendulum(_load, 600)
    t = np.arange(1, 601) * jc.TIME_STEP
    period = jc.crossing_period(t, r["pos"][:, 0] - 1.0)
    assert abs(period / jc.pendulum_period(math.pi / 2) - 1.0) < 0.005, period
    assert r["pivot_err"].max() < 0.002
    later = r["pos"][100:, 1]
    assert 2.99 < later.max() <= 3.0 + 1e-3


def test_hinge_removes_off_axis_spin():
    """Spin about x and y decays away under the soft collinearity rows (below 1e-3 rad/s in
    10 steps, 1e-6 by the end); spin about z about the box's own centre becomes rotation
    about the pivot with angular momentum conserved (I_cm / I_pivot)."""
    _, r = jc.run_pendulum(_load, 120, scenes.hinged_pendulum(gravity=(0.0, 0.0, 0.0), spin=(3.0, 2.0, 1.0)))
    assert np.abs(r["w"][10:, :2]).max() < 1e-3 and np.abs(r["w"][-1, :2]).max() < 1e-6
    assert abs(r["w"][-1, 2] / I_CM_OVER_I_PIVOT - 1.0) < 0.01
    assert np.abs(r["pos"][:, 2]).max() < 1e-3 and r["pivot_err"].max() < 1e-3


def test_motor_reaches_speed():
    _, r = jc.run_pendulum(_load, 60, scenes.hinged_pendulum(gravity=(0.0, 0.0, 0.0), motor=(2.0, 1e4)))
    assert np.abs(r["w"][30:, 2] - 2.0).max() < 1e-3


def test_limit_stops_the_swing():
    _, r = jc.run_pendulum(_load, 240, scenes.hinged_pendulum(limit=(-0.5, 0.5)))
    angle = jc.twist_angle(r["rot"])
    assert angle.min() > -0.52 and abs(angle[-1] + 0.5) < 0.005


def test_spring_frequency():
    """The spring acts through the axial mass at the box's centre, so the hinged box
    oscillates at hertz * sqrt(I_cm / I_pivot)."""
    _, r = jc.run_pendulum(_load, 600, scenes.hinged_pendulum(gravity=(0.0, 0.0, 0.0), spring=(1.0, 0.0, 0.0),
                                                              spin=(0.0, 0.0, 2.0)))
    t = np.arange(1, 601) * jc.TIME_STEP
    period = jc.crossing_period(t, jc.twist_angle(r["rot"]))
    assert abs(period * math.sqrt(I_CM_OVER_I_PIVOT) - 1.0) < 0.01, period


def test_chain_holds_together_without_self_contacts():
    worst, contacts, last = jc.chain_stretch(_load)
    assert worst.max() < 0.02
    assert contacts.max() == 0  # adjacent links overlap; collide_connected = false filters them
    assert last[1] < 10.0 - 1.0  # it swung down


def test_joint_clears_contacts_unless_collide_connected():
    """create_joint_clears_contacts / create_joint_keeps_contacts
    (joint_inspection_tests.rs:460-567)."""
    for collide_connected, expect in ((False, 0), (True, 1)):
        w = _load(scenes.touching_boxes())
        w.step(jc.TIME_STEP, jc.SUB_STEPS)
        assert w.contact_count() == 1
        w.create_revolute_joint(scenes.JointSpec(0, 1, collide_connected=collide_connected))
        assert w.contact_count() == expect
        w.step(jc.TIME_STEP, jc.SUB_STEPS)
        assert w.contact_count() == expect


def test_hinged_pendulum_replays():
    """hinged_pendulum (revolute_joint_tests.rs:103-172): two runs agree bit for bit,
    body state and joint impulses."""
    runs = []
    for _ in range(2):
        w = _load(scenes.hinged_pendulum())
        for _ in range(120):
            w.step(jc.TIME_STEP, jc.SUB_STEPS)
        runs.append((w.world_state_hash(), w.joint(0)[0].tobytes(), w.joint(0)[1]))
    assert runs[0] == runs[1]
    assert runs[0][2] == 22  # a joint with one dynamic body takes the highest free colour


def test_chain_joint_colours_alternate():
    """Dynamic-dynamic joints take the first free colour (constraint_graph.rs:256-276): a
    chain's joints alternate 0, 1, 0, 1, ...; the anchor joint (static) takes 22."""
    w = _load(scenes.hanging_chain(6))
    colours = [w.joint(j)[1] for j in range(6)]
    assert colours == [0, 1, 0, 1, 0, 22]


def test_joint_hub_overflows_joints_and_contacts():
    """A hub jointed to 24 bodies: joints 0..19 take colours 0..19, the rest overflow; the
    two boxes resting on the hub overflow too (constraint_graph.rs:113-197, :256-298)."""
    w = _load(scenes.joint_hub())
    assert [w.joint(j)[1] for j in range(24)] == list(range(20)) + [23] * 4
    for _ in range(60):
        w.step(jc.TIME_STEP, jc.SUB_STEPS)
    color, _, alive = w.contact_colors()
    assert int(((color == 23) & (alive == 1)).sum()) == 2
