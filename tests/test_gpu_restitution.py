"""The reference's restitution contract on the HIP stepper, bit for bit with the oracle.

Each case of src/tests/simulation/restitution_phase_tests.rs runs on the GPU world with the
oracle stepping the same scene beside it; the state hashes (body_queries.rs:432) must agree
after every step, and the reference test's own thresholds are asserted on the GPU's values.
The phase and head-on sweeps also run as one batch of worlds each.
"""
import numpy as np
import pytest

import restitution_cases as rc
from natsel_amd import scenes

pytestmark = pytest.mark.gpu


class Twin:
    """The GPU world and the oracle on one scene; every step compares their hashes."""

    def __init__(self, scene):
        from natsel_amd.world import load_scene
        from oracle import pyoracle

        self.name = scene.name
        self.gpu = load_scene(scene)
        self.ref = pyoracle.load_scene(scene)
        self.steps = 0

    def step(self, dt, substeps):
        self.gpu.step(dt, substeps)
        self.ref.step(dt, substeps)
        self.steps += 1
        g, r = self.gpu.world_state_hash(), self.ref.world_state_hash()
        assert g == r, f"{self.name}: step {self.steps} hash {g:016x} != oracle {r:016x}"

    def body_arrays(self):
        return self.gpu.body_arrays()

    def world_state_hash(self):
        return self.gpu.world_state_hash()


def test_phase(gpu_available):
    for e in (0.25, 0.5, 0.9):
        m = [rc.ground_bounce(Twin, e, g) for g in rc.phase_gaps()]
        assert max(m) - min(m) <= 0.05 and abs(float(np.mean(m)) - e) <= 0.05, (e, m)


def test_supported_and_threshold(gpu_available):
    for propagate in (False, True):
        floor = 0.75 if propagate else 0.5
        for e in (0.5, 0.9):
            for n in range(4):
                m = rc.supported_bounce(Twin, e, n, propagate)
                assert floor * e <= m <= e + 0.01, (propagate, e, n, m)
    assert rc.threshold_bounce(Twin, 10.0, 1.0) <= 0.01
    assert rc.threshold_bounce(Twin, 1.0, rc.IMPACT) / rc.IMPACT > 0.85


def test_head_on(gpu_available):
    for density in (1.0, 10.0, 100.0):
        for e in (0.0, 0.25, 0.5, 0.75, 1.0):
            ratio, drift = rc.head_on(Twin, e, density)
            assert abs(ratio - e) <= 0.02 and abs(drift) <= 1e-4, (density, e, ratio, drift)


def test_flat_landing_and_spin(gpu_available):
    for e in (0.5, 0.9):
        ratio, spin = rc.flat_bounce(Twin, e)
        assert abs(ratio - e) <= 0.01 and spin <= 0.03, (e, ratio, spin)
    for spin in (0.0, 0.5, 1.0, 2.0):
        ratio, _ = rc.flat_bounce(Twin, 0.9, spin=spin)
        assert abs(ratio - 0.9) <= 0.05, (spin, ratio)


def test_drop_apexes_10m(gpu_available):
    rc.check_apexes(rc.drop_apexes(Twin, 10.0), 10.0)


def test_resting_stack(gpu_available):
    drift, peak = rc.resting_stack(Twin)
    assert drift < 0.01 and peak < 0.05, (drift, peak)


def _batch_vs_oracles(scene_list, steps):
    from natsel_amd.batch import load_batch
    from oracle import pyoracle

    b = load_batch(scene_list)
    refs = [pyoracle.load_scene(s) for s in scene_list]
    for k in range(steps):
        b.step(rc.TIME_STEP, rc.SUB_STEPS)
        for r in refs:
            r.step(rc.TIME_STEP, rc.SUB_STEPS)
        want = np.array([r.world_state_hash() for r in refs], dtype=np.uint64)
        bad = np.nonzero(b.state_hashes() != want)[0]
        assert len(bad) == 0, f"step {k + 1}: worlds {bad.tolist()} differ from the oracle"
    assert b.stats()["error_worlds"] == 0
    return b


def test_batched_phase_sweep(gpu_available):
    """All 48 impact phases of phase() as one batch."""
    sweep = [(e, g) for e in (0.25, 0.5, 0.9) for g in rc.phase_gaps()]
    b = _batch_vs_oracles([scenes.ground_bounce(e, g) for e, g in sweep], 60)
    for e in (0.25, 0.5, 0.9):
        m = [float(b.body_arrays(k)[2][1, 1]) / rc.IMPACT for k, (ek, _) in enumerate(sweep) if ek == e]
        assert max(m) - min(m) <= 0.05 and abs(float(np.mean(m)) - e) <= 0.05, (e, m)


def test_batched_head_on(gpu_available):
    """The fifteen head-on collisions of head_on() as one batch."""
    cases = [(e, d) for d in (1.0, 10.0, 100.0) for e in (0.0, 0.25, 0.5, 0.75, 1.0)]
    b = _batch_vs_oracles([scenes.head_on(e, d) for e, d in cases], 120)
    for k, (e, _) in enumerate(cases):
        v = b.body_arrays(k)[2]
        assert abs((float(v[1, 0]) - float(v[0, 0])) / rc.IMPACT - e) <= 0.02, (k, e)
