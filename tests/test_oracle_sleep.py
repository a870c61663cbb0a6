"""The oracle's islands and sleep (island.rs, sleep_island.rs, solver_set_ops.rs).

Pinned by the reference's own tests — a resting box is in a sleeping set after 60 steps and
ends in set 3 at index 0 (sleep_tests.rs:33-99); two stacked boxes leave two island slots,
the first touching contact is contact 0 in island 0 (island_split_tests.rs:46-117) — and by
invariants a wrong index or a lost link breaks: a sleeping body does not move and has no
velocity; bodies are woken by an awake body touching them and fall asleep again; an island
whose contacts were removed is split, so the part at rest sleeps while the rest moves; every
island's bodies, contacts and indices agree with the bodies' and contacts' own records.
"""
import numpy as np

from natsel_amd import scenes

DT, SUB = 1.0 / 60.0, 4


def _load(scene):
    from oracle import pyoracle

    return pyoracle.load_scene(scene)


def test_box_sleeps():
    w = _load(scenes.box_sleeps())
    for i in range(180):
        w.step(DT, SUB)
        assert w.check_islands() == 0
        if i == 59:
            assert w.body_location(1)[0] >= 3
    assert w.body_location(1)[:2] == (3, 0)
    assert not w.body_arrays()[2][1].any()


def test_island_split_known_answer():
    w = _load(scenes.island_split())
    for _ in range(30):
        w.step(DT, SUB)
    s = w.island_summary()
    assert s["island_slots"] == 2 and s["islands"] == 1
    assert w.first_island_contact() == (0, 0)
    before = w.island_hash()
    w.unlink_and_split(0, 0)
    assert w.island_hash() != before
    assert w.check_islands() == 0


def test_sleep_and_wake():
    """The box sleeps before the ball arrives, the ball's landing wakes it into one island
    with the ball, and both sleep again; while asleep the box does not move."""
    w = _load(scenes.sleep_and_wake())
    asleep_pos, woke, slept_again = None, None, None
    for i in range(400):
        w.step(DT, SUB)
        assert w.check_islands() == 0, i
        box, ball = w.body_location(1)[0], w.body_location(2)[0]
        pos = w.body_arrays()[0][1].copy()
        if box >= 3 and woke is None:
            if asleep_pos is None:
                asleep_pos = pos
            assert np.array_equal(pos, asleep_pos)
        if box == 2 and asleep_pos is not None and woke is None:
            woke = i
            assert w.island_summary()["islands"] == 1
        if woke is not None and box >= 3 and ball >= 3:
            slept_again = i
            break
    assert asleep_pos is not None and woke is not None and slept_again is not None, (asleep_pos, woke, slept_again)
    assert w.body_location(1)[0] == w.body_location(2)[0]  # one island, one sleeping set


def test_roll_off_splits_and_sleeps():
    w = _load(scenes.roll_off())
    split_seen = False
    for _ in range(120):
        w.step(DT, SUB)
        assert w.check_islands() == 0
        if w.island_summary()["islands"] == 2:
            split_seen = True
        if w.body_location(1)[0] >= 3:
            break
    assert split_seen and w.body_location(1)[0] >= 3 and w.body_location(2)[0] == 2
    assert np.linalg.norm(w.body_arrays()[2][2]) > 0.1  # the ball still rolls


def test_pyramid_sleeps_entirely():
    """A settled pyramid falls asleep as one island; sleeping changes no position."""
    sc = scenes.pyramid(6, density=1000.0, ground_half=(40.0, 1.0, 40.0))
    sc.enable_sleep = True
    w = _load(sc)
    for _ in range(400):
        w.step(DT, SUB)
        assert w.check_islands() == 0
        if w.island_summary()["awake_islands"] == 0:
            break
    s = w.island_summary()
    assert s["awake_islands"] == 0 and s["sleeping_sets"] >= 1, s
    pos = w.body_arrays()[0].copy()
    for _ in range(30):
        w.step(DT, SUB)
    assert np.array_equal(w.body_arrays()[0], pos)


def test_jointed_bodies_sleep_and_wake_together():
    """A joint joins its bodies' islands: both boxes fall asleep in one set with the joint out
    of the graph (colour -1) and its impulses held, and the ball wakes both and returns the
    joint to the graph."""
    w = _load(scenes.jointed_sleepers())
    asleep = woke = held = None
    for i in range(600):
        w.step(DT, SUB)
        assert w.check_islands() == 0, i
        a, b, ball = w.body_location(1), w.body_location(2), w.body_location(3)
        if asleep is None and a[0] >= 3:
            asleep = i
            assert b[0] == a[0] and ball[0] == 2 and w.joint(0)[1] == -1
            held = w.joint(0)[0]
        if asleep is not None and woke is None:
            if a[0] >= 3:
                assert np.array_equal(w.joint(0)[0], held)
            else:
                woke = i
                assert b[0] == 2 and w.joint(0)[1] >= 0
        if woke is not None and a[0] >= 3 and ball[0] >= 3:
            break
    assert asleep is not None and woke is not None and a[0] == b[0] == ball[0] >= 3, (asleep, woke, a, b, ball)
    assert held.any()


def test_sleepy_stacks_keep_their_records():
    """Toppling towers with sleep on: islands merge, split, sleep and wake (all of it happens
    within 600 steps) while every island, set and list record stays consistent."""
    w = _load(scenes.sleepy_stacks(64, 7))
    slept = woke = split = 0
    prev = w.island_summary()
    for i in range(600):
        w.step(DT, SUB)
        assert w.check_islands() == 0, i
        s = w.island_summary()
        slept += s["sleeping_sets"] > prev["sleeping_sets"]
        woke += s["sleeping_sets"] < prev["sleeping_sets"]
        split += s["split_island"] >= 0
        prev = s
    assert slept and woke and split, (slept, woke, split)
