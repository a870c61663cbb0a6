"""Islands and sleep on the HIP stepper, bit for bit with the oracle.

Each scene steps on the GPU world with the oracle beside it; after every step they agree on
the state hash (body_queries.rs:432), every body's solver set, index and island, the island
counts, every live contact's colour and list index, and both pass their record cross-checks.
The scenes cover a body falling asleep (sleep_tests.rs:33-99), island bookkeeping with sleep
off (island_split_tests.rs:46-117), waking by a falling body, a split followed by sleep,
joints sleeping and waking with their bodies, a pyramid sleeping whole and toppling towers
whose islands merge, split, sleep and wake; and a snapshot taken with sleeping sets.
"""
import numpy as np
import pytest

from natsel_amd import scenes

pytestmark = pytest.mark.gpu

DT, SUB = 1.0 / 60.0, 4


class SleepTwin:
    def __init__(self, scene):
        from natsel_amd.world import load_scene
        from oracle import pyoracle

        self.name = scene.name
        self.n_bodies = len(scene.bodies)
        self.n_joints = len(scene.joints)
        self.gpu = load_scene(scene)
        self.ref = pyoracle.load_scene(scene)
        self.steps = 0

    def compare(self):
        tag = f"{self.name} step {self.steps}"
        g, r = self.gpu.world_state_hash(), self.ref.world_state_hash()
        assert g == r, f"{tag}: hash {g:016x} != oracle {r:016x}"
        for b in range(self.n_bodies):
            gl, rl = self.gpu.body_location(b), self.ref.body_location(b)
            assert gl == tuple(rl), f"{tag}: body {b} (set, index, island) {gl} != oracle {tuple(rl)}"
        assert self.gpu.island_summary() == self.ref.island_summary(), tag
        gc, rc = self.gpu.contact_colors(), self.ref.contact_colors()
        alive = np.nonzero(rc[2])[0]
        assert np.array_equal(np.nonzero(gc[2])[0], alive), f"{tag}: live contacts differ"
        for k, what in ((0, "colour"), (1, "index")):
            assert np.array_equal(gc[k][alive], rc[k][alive]), f"{tag}: contact {what}s differ"
        for j in range(self.n_joints):
            gi, gcol = self.gpu.joint(j)
            ri, rcol = self.ref.joint(j)
            assert gcol == rcol and gi.tobytes() == ri.tobytes(), f"{tag}: joint {j}"
        assert self.gpu.check_islands() == 0, tag
        assert self.ref.check_islands() == 0, tag

    def _hash_only(self):
        g, r = self.gpu.world_state_hash(), self.ref.world_state_hash()
        assert g == r, f"{self.name} step {self.steps}: hash {g:016x} != oracle {r:016x}"
        assert self.gpu.island_summary() == self.ref.island_summary(), f"{self.name} step {self.steps}"

    def step(self, n=1):
        for _ in range(n):
            self.gpu.step(DT, SUB)
            self.ref.step(DT, SUB)
            self.steps += 1
            self.compare()


def _sleeping_sets(t):
    return t.gpu.island_summary()["sleeping_sets"]


def test_box_sleeps():
    t = SleepTwin(scenes.box_sleeps())
    t.step(180)
    assert t.gpu.body_location(1)[:2] == (3, 0)
    assert not t.gpu.body_arrays()[2][1].any()


def test_island_bookkeeping_without_sleep():
    t = SleepTwin(scenes.island_split())
    t.step(60)
    s = t.gpu.island_summary()
    assert s["island_slots"] == 2 and s["islands"] == 1


def test_sleep_and_wake():
    t = SleepTwin(scenes.sleep_and_wake())
    seen = []
    for _ in range(200):
        t.step()
        seen.append(t.gpu.body_location(1)[0])
    assert 3 in seen and seen[seen.index(3):].count(2) > 0 and seen[-1] >= 3


def test_roll_off_splits_and_sleeps():
    t = SleepTwin(scenes.roll_off())
    t.step(120)
    assert t.gpu.body_location(1)[0] >= 3 and t.gpu.body_location(2)[0] == 2


def test_jointed_bodies_sleep_and_wake_together():
    t = SleepTwin(scenes.jointed_sleepers())
    colours = []
    for _ in range(200):
        t.step()
        colours.append(t.gpu.joint(0)[1])
    assert -1 in colours and colours[colours.index(-1):].count(-1) < len(colours) - colours.index(-1)


def test_pyramid_sleeps_entirely():
    sc = scenes.pyramid(6, density=1000.0, ground_half=(40.0, 1.0, 40.0))
    sc.enable_sleep = True
    t = SleepTwin(sc)
    for _ in range(400):
        t.step()
        if t.gpu.island_summary()["awake_islands"] == 0:
            break
    assert t.gpu.island_summary()["awake_islands"] == 0
    t.step(10)  # nothing awake: the step does no work and the state holds


@pytest.mark.parametrize("count,seed", [(64, 7), (200, 3)])
def test_sleepy_stacks(count, seed):
    t = SleepTwin(scenes.sleepy_stacks(count, seed))
    most = 0
    for _ in range(600):
        t.step()
        most = max(most, _sleeping_sets(t))
    assert most >= 3


def test_snapshot_with_sleeping_sets():
    """save_state / restore_state (world_state.rs:340-560) with bodies asleep: the steps after a
    restore repeat the steps after the save bit for bit."""
    t = SleepTwin(scenes.sleepy_stacks(64, 7))
    t.step(250)
    assert _sleeping_sets(t) > 0
    snap = t.gpu.save_state()
    hashes = []
    for _ in range(120):
        t.gpu.step(DT, SUB)
        hashes.append(t.gpu.world_state_hash())
    t.gpu.restore_state(snap)
    for k in range(120):
        t.gpu.step(DT, SUB)
        assert t.gpu.world_state_hash() == hashes[k], k
    assert t.gpu.check_islands() == 0


def test_pile_with_reference_defaults():
    """1000 mixed bodies dropped into a container with the reference's defaults (sleep and
    continuous collision on): falling bodies are swept, islands form and fall asleep."""
    t = SleepTwin(scenes.pile_defaults(1000, size_x=10, size_y=20, size_z=10))
    full = t.compare
    for k in range(600):
        t.compare = full if k % 20 == 19 else t._hash_only
        t.step()
    t.compare = full
    s = t.gpu.island_summary()
    assert s["sleeping_sets"] >= 4 and t.gpu.kernel_stats()["toi_hits"] > 0, s
