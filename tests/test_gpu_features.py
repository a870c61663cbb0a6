"""Every BodyDef and ShapeDef field of the C ABI on the GPU against the oracle, every step.

natsel_amd.scenes.def_features: motion locks, fixed rotation, filter groups and masks,
sensors (including one crossed by a continuous-collision ball), rolling resistance on
spheres and a capsule, allow_fast_rotation, contact recycling off, a per-body sleep
threshold, speculative contact off and hit events on. The single-world stepper runs it with
sleep and continuous collision on; the batched stepper with both off."""
import numpy as np
import pytest

from natsel_amd import scenes

pytestmark = pytest.mark.gpu


def test_def_features_single_world_every_step(gpu_available):
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = scenes.def_features()
    g, o = load_scene(sc), pyoracle.load_scene(sc)
    for k in range(300):
        g.step(sc.dt, sc.substeps)
        o.step(sc.dt, sc.substeps)
        hg, ho = g.world_state_hash(), o.world_state_hash()
        assert hg == ho, f"step {k}: gpu {hg:x} != oracle {ho:x}"
    ga, gb = g.contact_shapes()
    oa, ob = o.contact_shapes()
    n = min(len(ga), len(oa))
    assert np.array_equal(ga[:n], oa[:n]) and np.array_equal(gb[:n], ob[:n])


def test_def_features_batch_every_step(gpu_available):
    from natsel_amd.batch import load_batch
    from oracle import pyoracle

    worlds = [scenes.def_features(enable_sleep=False, enable_continuous=False), scenes.pyramid(6, density=1000.0)]
    for w in worlds:
        w.enable_sleep, w.enable_continuous = False, False
    b = load_batch(worlds)
    o = [pyoracle.load_scene(w) for w in worlds]
    for k in range(200):
        b.step(worlds[0].dt, worlds[0].substeps)
        for w in o:
            w.step(worlds[0].dt, worlds[0].substeps)
        got = b.state_hashes()
        want = [w.world_state_hash() for w in o]
        assert [int(x) for x in got] == want, f"step {k}"
    assert b.stats()["error_worlds"] == 0
