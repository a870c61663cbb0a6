"""The oracle against the reference's own known-answer tests (CPU only).

Each test rebuilds the reference test's scene through natsel_amd.scenes and checks the
values the reference test asserts.
"""
import numpy as np

from natsel_amd import scenes


def _oracle(scene):
    from oracle import pyoracle

    return pyoracle.load_scene(scene)


def test_overflow_hub_fills_overflow_colour():
    """src/tests/simulation/overflow_color_tests.rs:82-103: after 10 steps the overflow
    colour holds 5 contacts, with 27 bodies and 65 contacts."""
    sc = scenes.overflow_hub()
    assert len(sc.bodies) == 27
    w = _oracle(sc)
    for _ in range(10):
        w.step(1.0 / 60.0, 4)
    color, _, alive = w.contact_colors()
    assert int(alive.sum()) == 65
    assert int(((color == 23) & (alive == 1)).sum()) == 5


def test_graph_colours_and_local_indices():
    """src/tests/graph_tests.rs:152-172: colour and local index of each touching contact
    after the first pair update and narrow phase. (The reference test adds contacts to
    the graph without removing them from the non-touching list, so only the touching
    contacts' indices carry over to a full step.)"""
    w = _oracle(scenes.graph_scene())
    w.step(1.0 / 60.0, 4)
    color, local, alive = w.contact_colors()
    assert len(color) == 8 and alive.all()
    expected = {0: (22, 0), 1: (22, 1), 2: (22, 2), 3: (0, 0), 5: (1, 0)}
    for cid, (c, l) in expected.items():
        assert (int(color[cid]), int(local[cid])) == (c, l), cid
    for cid in (4, 6, 7):
        assert color[cid] == -1
    used = np.unique(color[color >= 0])
    assert list(used) == [0, 1, 22]


def test_cos_sin_matches_closed_form():
    """compute_cos_sin is a rational approximation: within 1e-3 of cos/sin, exact at the
    quadrant points, unit length."""
    for t in np.linspace(-7.0, 7.0, 141, dtype=np.float32):
        c, s = scenes.compute_cos_sin(np.float32(t))
        assert abs(float(c) - np.cos(float(t))) < 2e-3 and abs(float(s) - np.sin(float(t))) < 2e-3
        assert abs(float(c) * float(c) + float(s) * float(s) - 1.0) < 1e-6
    assert scenes.compute_cos_sin(np.float32(0.0)) == (1.0, 0.0)


def test_batched_scene_generator_is_seeded():
    """The batched-world generator: log-uniform counts within bounds, deterministic."""
    a = scenes.batched(count=64, seed=5)
    b = scenes.batched(count=64, seed=5)
    counts = [s.dynamic_count for s in a]
    assert counts == [s.dynamic_count for s in b]
    assert min(counts) >= 8 and max(counts) <= 256
    assert [s.bodies[1].position for s in a] == [s.bodies[1].position for s in b]
    assert sum(1 for s in a if s.name.startswith("pile")) == 32


def test_pyramid_settles_and_rests():
    """The 210-box pyramid (scenes.rs:105-139, base 20, density 1000) settles within 60
    steps and then rests: over steps 240-600 the top box moves less than 1 cm and no body
    exceeds 0.05 m/s (the resting-stack pattern of restitution_phase_tests.rs:1036-1049).
    The 20 soft contact layers compress the stack by 2.76 cm while it settles."""
    sc = scenes.pyramid(20, density=1000.0)
    w = _oracle(sc)
    top = len(sc.bodies) - 1
    assert sc.bodies[top].position[1] == 19.5
    heights, speeds = [], []
    for _ in range(600):
        w.step(1.0 / 60.0, 4)
        pos, _, v, _ = w.body_arrays()
        heights.append(pos[top, 1])
        speeds.append(float(np.linalg.norm(v[1:].astype(np.float64), axis=1).max()))
    settled = heights[239]
    assert max(abs(h - settled) for h in heights[240:]) < 0.01
    assert max(speeds[240:]) < 0.05
    assert 0.02 < 19.5 - settled < 0.03


def test_pushed_box_worlds_only_touch_the_ground():
    """The differentiable-rollout worlds: 32 boxes each, every contact a box-ground contact
    in colour 22 (no box reaches another within 64 steps), all at rest at the end."""
    for seed in (3, 77):
        sc = scenes.pushed_boxes(seed)
        assert sc.dynamic_count == 32
        w = _oracle(sc)
        for _ in range(64):
            w.step(1.0 / 60.0, 4)
        color, _, alive = w.contact_colors()
        assert sorted(set(color[alive == 1].tolist())) == [22] and int(alive.sum()) == 32
    t = scenes.pushed_box_targets(16)
    assert t.shape == (16, 3) and np.array_equal(t, scenes.pushed_box_targets(16))
