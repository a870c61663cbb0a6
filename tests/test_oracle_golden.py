"""The oracle against the reference's own golden state hashes.

Tools/Benchmarks/reference-hashes.json holds the end-state world_state_hash of each
benchmark scene (FNV-1a over every body's f64 position, rotation and awake
velocities). Matching it pins the oracle's whole box path bit for bit: pair list,
contact ids, graph colouring, SAT/clipping/reduction, feature ids, the soft-step
solver, finalize and fat AABBs.
"""
import json
import os

from natsel_amd import scenes
from oracle import pyoracle

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_hashes.json")))


def _run(scene, steps):
    w = pyoracle.load_scene(scene)
    for _ in range(steps):
        w.step(1.0 / 60.0, 4)
    return w


def test_empty_world_hash():
    w = pyoracle.OracleWorld()
    assert w.world_state_hash() == int(GOLDEN["empty_world_hash"], 16)


def test_large_pyramid_golden_hash():
    g = GOLDEN["scenes"]["large_pyramid"]
    w = _run(scenes.pyramid(100), g["steps"])
    assert w.stats()["fast_bodies"] == 0
    assert w.world_state_hash() == int(g["state_hash"], 16)


def test_many_pyramids_golden_hash():
    g = GOLDEN["scenes"]["many_pyramids"]
    w = _run(scenes.many_pyramids(), g["steps"])
    assert w.stats()["fast_bodies"] == 0
    assert w.world_state_hash() == int(g["state_hash"], 16)


def test_rollback_golden_hash():
    """rollback (reference-hashes.json:90-94): the large pyramid saved every step, every 30
    steps rolled back 8 states and replayed, ends on the large pyramid's hash."""
    g = GOLDEN["scenes"]["rollback"]
    w = scenes.run_rollback(pyoracle.load_scene(scenes.pyramid(100)), g["steps"])
    assert w.world_state_hash() == int(g["state_hash"], 16)


def test_planted_mistake_changes_hash():
    """A one-ulp change of one body's start position must change the end hash: the
    hash check is sensitive to single-bit differences, not just coarse motion."""
    import numpy as np

    sc = scenes.pyramid(20, density=1000.0)
    a = _run(sc, 30).world_state_hash()
    b0 = sc.bodies[5]
    x = np.float32(b0.position[0])
    sc.bodies[5].position = (float(np.nextafter(x, np.float32(1e9))), b0.position[1], b0.position[2])
    b = _run(sc, 30).world_state_hash()
    assert a != b
