"""The reference's restitution and resting-contact tests as measurements over any stepper.

src/tests/simulation/restitution_phase_tests.rs holds the reference's bounce contract: a
measured coefficient of restitution within a stated tolerance of the material's, with no
energy from nowhere and no creep of a resting stack. Each measure_* function below rebuilds
one of those tests' scenes (natsel_amd.scenes), steps it through `load(scene)` and returns
the quantity the reference test asserts on; the thresholds live in the test files.

`load` returns a stepper with step(dt, substeps), body_arrays() and world_state_hash(): the
oracle for the CPU tests, or the HIP world twinned with the oracle (hashes compared after
every step) for the GPU tests.
"""
import math

import numpy as np

from natsel_amd import scenes

TIME_STEP = 1.0 / 60.0
SUB_STEPS = 4
IMPACT = float(scenes.IMPACT_SPEED)


def _run(w, steps, substeps=SUB_STEPS):
    for _ in range(steps):
        w.step(TIME_STEP, substeps)


def ground_bounce(load, restitution, gap):
    """measure_ground_bounce (restitution_phase_tests.rs:62-72): v_y / 5 after 60 steps."""
    w = load(scenes.ground_bounce(restitution, gap))
    _run(w, 60)
    return float(w.body_arrays()[2][1, 1]) / IMPACT


def phase_gaps(samples=16):
    """The impact phases of phase() (restitution_phase_tests.rs:108-113), in f32."""
    return [np.float32(i / samples) * scenes.IMPACT_SPEED * np.float32(TIME_STEP) for i in range(samples)]


def supported_bounce(load, restitution, count, propagate):
    """measure_supported_bounce (restitution_phase_tests.rs:76-96)."""
    w = load(scenes.supported_bounce(restitution, count, propagate))
    _run(w, 60)
    return float(w.body_arrays()[2][count + 1, 1]) / IMPACT


def threshold_bounce(load, threshold, speed):
    """threshold / threshold_above (restitution_phase_tests.rs:160-193): v_y after 60 steps."""
    w = load(scenes.threshold_bounce(threshold, speed))
    _run(w, 60)
    return float(w.body_arrays()[2][1, 1])


def head_on(load, restitution, density_b):
    """measure_head_on (restitution_phase_tests.rs:202-238): (relative speed ratio, momentum
    drift per unit mass) after 120 steps. Both spheres have radius 0.5, so their masses are
    in the ratio of their densities."""
    w = load(scenes.head_on(restitution, density_b))
    m_a = 1.0 * (4.0 / 3.0) * math.pi * 0.125
    m_b = density_b * (4.0 / 3.0) * math.pi * 0.125
    v = w.body_arrays()[2]
    p0 = m_a * float(v[0, 0]) + m_b * float(v[1, 0])
    _run(w, 120)
    v = w.body_arrays()[2]
    va, vb = float(v[0, 0]), float(v[1, 0])
    return (vb - va) / IMPACT, (m_a * va + m_b * vb - p0) / (m_a + m_b)


def flat_bounce(load, restitution, sub_steps=SUB_STEPS, spin=0.0):
    """measure_flat_bounce / measure_spinning_bounce (restitution_phase_tests.rs:824-918):
    (v_y / 5, |angular velocity|) after 60 steps."""
    w = load(scenes.flat_landing(restitution, sub_steps, spin))
    _run(w, 60, sub_steps)
    _, _, v, av = w.body_arrays()
    return float(v[1, 1]) / IMPACT, float(np.linalg.norm(av[1].astype(np.float64)))


def drop_apexes(load, drop_height, capacity=6):
    """measure_apexes (restitution_phase_tests.rs:937-958): the heights at which the ball's
    vertical velocity turns from up to down."""
    w = load(scenes.drop_ball(drop_height))
    apexes, previous = [], 0.0
    step = 0
    while step < 4000 and len(apexes) < capacity:
        w.step(TIME_STEP, SUB_STEPS)
        pos, _, v, _ = w.body_arrays()
        speed = float(v[1, 1])
        if previous > 0.0 and speed <= 0.0:
            apexes.append(float(np.float32(pos[1, 1])))
        previous = speed
        step += 1
    return apexes


def check_apexes(a, h):
    """The assertions of drop_apexes (restitution_phase_tests.rs:965-995)."""
    assert len(a) >= 6, (h, a)
    assert max(a) <= 1.001 * a[0] and a[0] <= 1.001 * h and a[-1] >= 0.8 * a[0], (h, a)


def resting_stack(load, box_count=10):
    """resting_stack (restitution_phase_tests.rs:1001-1049): settle 240 steps, then over 300
    steps the top box's largest height change and the largest body speed."""
    w = load(scenes.resting_stack(box_count))
    _run(w, 240)
    top = box_count
    settled = np.float32(w.body_arrays()[0][top, 1])
    drift, peak = 0.0, 0.0
    for _ in range(300):
        w.step(TIME_STEP, SUB_STEPS)
        pos, _, v, _ = w.body_arrays()
        drift = max(drift, abs(float(np.float32(pos[top, 1]) - settled)))
        peak = max(peak, float(np.linalg.norm(v[1:top + 1].astype(np.float64), axis=1).max()))
    return drift, peak


def overshoot(load, continuous):
    """measure_overshoot (restitution_phase_tests.rs:311-380): (first apex, energy ratio at
    the first apex, peak energy ratio, rebound/impact speed ratio, bounce count) of an
    elastic unit cube (density 1000, so mass 1000 kg and inertia m/6) dropped from 10 m."""
    w = load(scenes.overshoot(continuous))
    mass = 1000.0
    inertia = mass / 6.0
    g = np.array([0.0, -10.0, 0.0], np.float32)
    h = np.float32(TIME_STEP) / np.float32(SUB_STEPS)

    def energy():
        pos, rot, v, av = w.body_arrays()
        vel = v[1].astype(np.float64)
        omega = av[1].astype(np.float64)
        kinetic = 0.5 * mass * vel.dot(vel) + 0.5 * inertia * omega.dot(omega)
        potential = -mass * g.astype(np.float64).dot(pos[1])
        phase = 0.5 * float(h) * mass * vel.dot(g.astype(np.float64))
        return kinetic + potential + phase

    start = energy()
    first_apex = first_ratio = peak = 0.0
    bounces = 0
    impact = rebound = previous = 0.0
    step = 0
    while step < 600 and bounces < 2:
        before = float(w.body_arrays()[2][1, 1])
        w.step(TIME_STEP, SUB_STEPS)
        speed = float(w.body_arrays()[2][1, 1])
        if rebound == 0.0 and before < 0.0 and speed > 0.0:
            impact, rebound = -before, speed
        if rebound == 0.0:
            previous = speed
            step += 1
            continue
        ratio = energy() / start
        peak = max(peak, ratio)
        if previous > 0.0 and speed <= 0.0:
            if bounces == 0:
                first_apex = float(np.float32(w.body_arrays()[0][1, 1]))
                first_ratio = ratio
            bounces += 1
        previous = speed
        step += 1
    return first_apex, first_ratio, peak, (rebound / impact if impact else 0.0), bounces


def check_overshoot(discrete, continuous):
    """The assertions of overshoot (restitution_phase_tests.rs:384-434)."""
    assert discrete[3] <= 1.0 + 1e-5 and continuous[3] <= 1.0 + 1e-5, (discrete, continuous)
    assert continuous[0] <= 10.0 and continuous[2] <= 1.001, continuous
    assert discrete[0] <= 10.0 + 0.15 and discrete[1] <= 1.02, discrete
    assert discrete[4] >= 2 and continuous[4] >= 1, (discrete, continuous)
