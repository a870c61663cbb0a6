"""Parity at the benchmark sizes: the GPU world against the oracle on the full workloads.

The 10k pile (sleep and continuous collision off, the headline workload) by state hash
at steps 100, 300, 600 and 1000; the large world (100,000 mixed bodies and 2,000
revolute-joint chains) by hash over its first 60 steps; and a seeded sample of 64 worlds
of the full 4096-world batch after its 240 steps. The hash is World::world_state_hash
(body_queries.rs:432) over every body's position, rotation and velocities, so equal
hashes mean every body's state agrees bit for bit.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from natsel_amd import scenes

pytestmark = pytest.mark.gpu


def test_pile10k_matches_oracle_through_settling(gpu_available):
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = scenes.pile(10_000)
    assert not sc.enable_sleep and not sc.enable_continuous
    pyoracle.set_threads(pyoracle.max_threads())
    g, o = load_scene(sc), pyoracle.load_scene(sc)
    done = 0
    for mark in (100, 300, 600, 1000):
        for _ in range(mark - done):
            g.step(sc.dt, sc.substeps)
            o.step(sc.dt, sc.substeps)
        done = mark
        hg, ho = g.world_state_hash(), o.world_state_hash()
        print(f"pile10k step {mark}: gpu {hg:016x} oracle {ho:016x}", flush=True)
        assert hg == ho, f"10k pile: step {mark}: gpu {hg:x} != oracle {ho:x}"
    # the settled pile is the contact-rich workload the headline times
    ks = g.kernel_stats()
    assert ks["points"] / ks["steps"] > 20_000
    g.close()


def test_large_world_matches_oracle_first_60_steps(gpu_available):
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = scenes.large_world()
    pyoracle.set_threads(pyoracle.max_threads())
    g, o = load_scene(sc), pyoracle.load_scene(sc)
    for k in range(1, 61):
        g.step(sc.dt, sc.substeps)
        o.step(sc.dt, sc.substeps)
        if k in (1, 10, 20, 30, 40, 50, 60):
            hg, ho = g.world_state_hash(), o.world_state_hash()
            print(f"large world step {k}: gpu {hg:016x} oracle {ho:016x}", flush=True)
            assert hg == ho, f"large world: step {k}: gpu {hg:x} != oracle {ho:x}"
    assert g.kernel_stats()["joints"] > 0
    g.close()


def test_batch4096_sample_matches_oracle(gpu_available):
    """All 4096 worlds step together for 240 steps; 64 seeded worlds are stepped alone on
    the oracle (one thread each, worlds in parallel) and their hashes compared."""
    from natsel_amd.batch import load_batch
    from oracle import pyoracle

    worlds = scenes.batched(4096)
    b = load_batch(worlds)
    b.step(worlds[0].dt, worlds[0].substeps, 240)
    got = b.state_hashes()
    assert b.stats()["error_worlds"] == 0
    b.close()
    sample = np.sort(np.random.default_rng(4096).choice(len(worlds), 64, replace=False))

    def run(i):
        pyoracle.set_threads(1)  # per calling thread: each worker runs its world on one thread
        o = pyoracle.load_scene(worlds[i])
        for _ in range(240):
            o.step(worlds[i].dt, worlds[i].substeps)
        return o.world_state_hash()

    with ThreadPoolExecutor(max_workers=16) as ex:
        want = list(ex.map(run, sample))
    bad = [int(i) for i, h in zip(sample, want) if int(got[i]) != h]
    assert not bad, f"batch worlds {bad[:8]} differ from the oracle after 240 steps"
