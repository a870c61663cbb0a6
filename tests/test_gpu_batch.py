"""Batched worlds on the GPU against the oracle stepping each world alone, bit for bit.

A batch world must reach the state hash the same world reaches as a World of its own
(World::step, solver_integration.rs:807; world_state_hash, body_queries.rs:432).
"""
import numpy as np
import pytest

from natsel_amd import scenes

pytestmark = pytest.mark.gpu


def _oracles(scene_list):
    from oracle import pyoracle

    return [pyoracle.load_scene(s) for s in scene_list]


def _check(batch, oracles, step):
    got = batch.state_hashes()
    want = np.array([o.world_state_hash() for o in oracles], dtype=np.uint64)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, f"step {step}: worlds {bad[:8].tolist()} differ from the oracle"


def test_batch_worlds_match_oracle_every_step(gpu_available):
    """Piles and stacks of 8..40 bodies (the 4096-world batch generator, scaled down), the
    210-box pyramid and the overflow hub in one batch."""
    from natsel_amd.batch import load_batch

    worlds = scenes.batched(count=14, seed=1234, lo=8, hi=40)
    worlds += [scenes.pyramid(20, density=1000.0), scenes.overflow_hub(enable_sleep=False)]
    for w in worlds:  # a batch shares its options
        w.enable_sleep = False
        w.enable_continuous = False
    b = load_batch(worlds)
    o = _oracles(worlds)
    for k in range(90):
        b.step(1.0 / 60.0, 4)
        for w in o:
            w.step(1.0 / 60.0, 4)
        _check(b, o, k)
    st = b.stats()
    assert st["error_worlds"] == 0


def test_batch_multi_step_launch_equals_single_steps(gpu_available):
    """Thirty steps in one call equal thirty calls of one step, and the oracle."""
    from natsel_amd.batch import load_batch

    worlds = scenes.batched(count=8, seed=99, lo=8, hi=64)
    a = load_batch(worlds)
    b = load_batch(worlds)
    a.step(1.0 / 60.0, 4, 30)
    for _ in range(30):
        b.step(1.0 / 60.0, 4)
    assert np.array_equal(a.state_hashes(), b.state_hashes())
    o = _oracles(worlds)
    for w in o:
        for _ in range(30):
            w.step(1.0 / 60.0, 4)
    _check(a, o, 29)


def test_batch_body_arrays_match_single_world(gpu_available):
    """Body poses and velocities read from a batch world equal the single-world stepper's."""
    from natsel_amd.batch import load_batch
    from natsel_amd.world import load_scene

    sc = scenes.batched(count=3, seed=7, lo=20, hi=30)
    b = load_batch(sc)
    singles = [load_scene(s) for s in sc]
    b.step(1.0 / 60.0, 4, 40)
    for w in singles:
        for _ in range(40):
            w.step(1.0 / 60.0, 4)
    for k, w in enumerate(singles):
        for x, y in zip(b.body_arrays(k), w.body_arrays()):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_batch_graph_replay_equals_direct_launches(gpu_available):
    """Steps replayed from the captured HIP graph (8 steps a replay, the rest launched
    directly) equal direct launches call after call, with the program's own eager work
    (hash readback, stats, a one-step call) between the replays, and match the oracle."""
    from natsel_amd.batch import load_batch

    worlds = scenes.batched(count=12, seed=2468, lo=8, hi=96)
    a = load_batch(worlds)
    b = load_batch(worlds)
    a.set_tuning(use_graph=True)
    assert a.tuning()["use_graph"] and not b.tuning()["use_graph"]
    o = _oracles(worlds)
    done = 0
    for steps in (8, 16, 1, 8, 21, 8):
        a.step(1.0 / 60.0, 4, steps)
        b.step(1.0 / 60.0, 4, steps)
        done += steps
        assert np.array_equal(a.state_hashes(), b.state_hashes()), f"after {done} steps"
        assert a.stats()["error_worlds"] == 0
    for w in o:
        for _ in range(done):
            w.step(1.0 / 60.0, 4)
    _check(a, o, done)


def test_batch_autotune_keeps_state_and_results(gpu_available):
    """The batch autotuner times its candidates from a snapshot and restores it: the hashes
    after the call equal those before, and the batch then steps as an untuned copy does."""
    from natsel_amd.batch import load_batch

    worlds = scenes.batched(count=16, seed=1357, lo=8, hi=128)
    a = load_batch(worlds)
    b = load_batch(worlds)
    a.step(1.0 / 60.0, 4, 20)
    b.step(1.0 / 60.0, 4, 20)
    before = a.state_hashes()
    t = a.autotune(1.0 / 60.0, 4, 8)
    assert t["candidates"] == 10 and t["step_ms"] > 0.0 and t["step_ms"] <= t["previous_ms"]
    assert np.array_equal(a.state_hashes(), before)
    assert a.tuning()["collide_blocks"] == t["collide_blocks"]
    for _ in range(3):
        a.step(1.0 / 60.0, 4, 10)
        b.step(1.0 / 60.0, 4, 10)
        assert np.array_equal(a.state_hashes(), b.state_hashes())
