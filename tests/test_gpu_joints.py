"""Revolute joints on the HIP stepper, bit for bit with the oracle.

Each scene runs on the GPU world with the oracle beside it: state hashes (body_queries.rs:432)
agree after every step, the joints' accumulated impulses and graph colours agree at the end,
and the closed-form bounds of the oracle's joint tests hold on the GPU's own values.
"""
import math

import numpy as np
import pytest

import joint_cases as jc
from natsel_amd import scenes

pytestmark = pytest.mark.gpu

I_CM_OVER_I_PIVOT = (0.4 * 0.4 / 6.0) / (0.4 * 0.4 / 6.0 + 0.25)


class Twin:
    """GPU world and oracle on one scene; every step compares their hashes."""

    def __init__(self, scene, solver_blocks=0):
        from natsel_amd.world import load_scene
        from oracle import pyoracle

        self.name = scene.name
        self.n_joints = len(scene.joints)
        self.gpu = load_scene(scene, solver_blocks=solver_blocks)
        self.ref = pyoracle.load_scene(scene)
        self.steps = 0

    def step(self, dt, substeps):
        self.gpu.step(dt, substeps)
        self.ref.step(dt, substeps)
        self.steps += 1
        g, r = self.gpu.world_state_hash(), self.ref.world_state_hash()
        assert g == r, f"{self.name}: step {self.steps} hash {g:016x} != oracle {r:016x}"

    def create_revolute_joint(self, spec):
        a = self.gpu.create_revolute_joint(spec)
        b = self.ref.create_revolute_joint(spec)
        assert a == b
        self.n_joints += 1
        return a

    def check_joints(self):
        for j in range(self.n_joints):
            gi, gc = self.gpu.joint(j)
            ri, rc = self.ref.joint(j)
            assert gc == rc, f"{self.name}: joint {j} colour {gc} != oracle {rc}"
            assert gi.tobytes() == ri.tobytes(), f"{self.name}: joint {j} impulses {gi} != oracle {ri}"

    def body_arrays(self):
        return self.gpu.body_arrays()

    def world_state_hash(self):
        return self.gpu.world_state_hash()

    def contact_count(self):
        n = self.gpu.contact_count()
        assert n == self.ref.contact_count()
        return n

    def contact_shapes(self):
        a, b = self.gpu.contact_shapes()
        ra, rb = self.ref.contact_shapes()
        live, rlive = np.nonzero(a >= 0)[0], np.nonzero(ra >= 0)[0]
        assert np.array_equal(live, rlive) and np.array_equal(a[live], ra[live]) and np.array_equal(b[live], rb[live])
        return a, b


def _twin_loader(made):
    def load(scene):
        t = Twin(scene)
        made.append(t)
        return t
    return load


def test_pendulum_period_and_pivot(gpu_available):
    made = []
    _, r = jc.run_pendulum(_twin_loader(made), 600)
    made[0].check_joints()
    t = np.arange(1, 601) * jc.TIME_STEP
    period = jc.crossing_period(t, r["pos"][:, 0] - 1.0)
    assert abs(period / jc.pendulum_period(math.pi / 2) - 1.0) < 0.005
    assert r["pivot_err"].max() < 0.002


def test_off_axis_spin_motor_limit_spring(gpu_available):
    made = []
    load = _twin_loader(made)
    _, r = jc.run_pendulum(load, 120, scenes.hinged_pendulum(gravity=(0.0, 0.0, 0.0), spin=(3.0, 2.0, 1.0)))
    assert np.abs(r["w"][-1, :2]).max() < 1e-6 and abs(r["w"][-1, 2] / I_CM_OVER_I_PIVOT - 1.0) < 0.01
    _, r = jc.run_pendulum(load, 60, scenes.hinged_pendulum(gravity=(0.0, 0.0, 0.0), motor=(2.0, 1e4)))
    assert np.abs(r["w"][30:, 2] - 2.0).max() < 1e-3
    _, r = jc.run_pendulum(load, 240, scenes.hinged_pendulum(limit=(-0.5, 0.5)))
    angle = jc.twist_angle(r["rot"])
    assert angle.min() > -0.52 and abs(angle[-1] + 0.5) < 0.005
    _, r = jc.run_pendulum(load, 600, scenes.hinged_pendulum(gravity=(0.0, 0.0, 0.0), spring=(1.0, 0.0, 0.0),
                                                             spin=(0.0, 0.0, 2.0)))
    t = np.arange(1, 601) * jc.TIME_STEP
    assert abs(jc.crossing_period(t, jc.twist_angle(r["rot"])) * math.sqrt(I_CM_OVER_I_PIVOT) - 1.0) < 0.01
    for tw in made:
        tw.check_joints()


def test_hanging_chain(gpu_available):
    made = []
    worst, contacts, last = jc.chain_stretch(_twin_loader(made))
    made[0].check_joints()
    assert worst.max() < 0.02 and contacts.max() == 0 and last[1] < 9.0


def test_joint_clears_contacts_unless_collide_connected(gpu_available):
    """joint_inspection_tests.rs:460-567 on the GPU world: the joint is created after a
    step, while the boxes touch."""
    for collide_connected, expect in ((False, 0), (True, 1)):
        tw = Twin(scenes.touching_boxes())
        tw.step(jc.TIME_STEP, jc.SUB_STEPS)
        assert tw.contact_count() == 1
        tw.create_revolute_joint(scenes.JointSpec(0, 1, collide_connected=collide_connected))
        assert tw.contact_count() == expect
        for _ in range(30):
            tw.step(jc.TIME_STEP, jc.SUB_STEPS)
        assert tw.contact_count() == expect
        tw.check_joints()


def test_chain_pile_every_step_matches_oracle(gpu_available):
    """400 mixed bodies and 40 chains of 16 capsule links (600 joints) falling into a
    container together: joints and contacts share the colours."""
    tw = Twin(scenes.chain_pile())
    for _ in range(300):
        tw.step(jc.TIME_STEP, jc.SUB_STEPS)
    tw.check_joints()


@pytest.mark.parametrize("blocks", [1, 3, 17])
def test_chain_pile_workgroup_count_does_not_change_results(gpu_available, blocks):
    tw = Twin(scenes.chain_pile(count=200, chains=20), solver_blocks=blocks)
    for _ in range(120):
        tw.step(jc.TIME_STEP, jc.SUB_STEPS)
    tw.check_joints()


def test_joint_hub_overflow_every_step(gpu_available):
    """Overflow joints and overflow contacts sharing the hub body: the dependency levels
    must reproduce the sequential overflow walk bit for bit."""
    tw = Twin(scenes.joint_hub())
    for _ in range(120):
        tw.step(jc.TIME_STEP, jc.SUB_STEPS)
    tw.check_joints()
