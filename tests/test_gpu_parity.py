"""GPU step vs the oracle, bit for bit, through the C ABI.

The reference engine is deterministic, and every float operation of the HIP path
follows its operand order, so each step's world_state_hash must equal the oracle's.
The golden reference hashes are also checked directly on the GPU.
"""
import json
import os

import numpy as np
import pytest

from natsel_amd import scenes

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_hashes.json")))


def _pair(scene):
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    return load_scene(scene), pyoracle.load_scene(scene)


def _first_difference(g, o):
    gp, gr, gv, gw = g.body_arrays()
    op, orr, ov, ow = o.body_arrays()
    for name, a, b in (("pos", gp, op), ("rot", gr, orr), ("v", gv, ov), ("w", gw, ow)):
        bad = np.nonzero(np.any(a.view(np.uint8).reshape(len(a), -1) != b.view(np.uint8).reshape(len(b), -1), axis=1))[0]
        if len(bad):
            i = int(bad[0])
            return f"{name} of body {i} ({len(bad)} bodies differ): gpu {a[i]} oracle {b[i]}"
    return "bodies equal (hash covers the same data)"


def test_pyramid20_every_step_matches_oracle(gpu_available):
    """The 210-box pyramid over 600 steps: the GPU's hash equals the oracle's after every
    step, and the GPU's own top box settles and stays put (the resting-stack pattern of
    restitution_phase_tests.rs:1036-1049: drift under 1 cm, speeds under 0.05 m/s once
    settled)."""
    sc = scenes.pyramid(20, density=1000.0)
    g, o = _pair(sc)
    top = len(sc.bodies) - 1
    heights, speeds = [], []
    for k in range(600):
        g.step(1.0 / 60.0, 4)
        o.step(1.0 / 60.0, 4)
        hg, ho = g.world_state_hash(), o.world_state_hash()
        assert hg == ho, f"step {k}: {hg:x} != {ho:x}; {_first_difference(g, o)}"
        pos, _, v, _ = g.body_arrays()
        heights.append(pos[top, 1])
        speeds.append(float(np.linalg.norm(v[1:].astype(np.float64), axis=1).max()))
    settled = heights[239]
    assert max(abs(h - settled) for h in heights[240:]) < 0.01
    assert max(speeds[240:]) < 0.05
    assert abs(settled - sc.bodies[top].position[1]) < 0.03  # 20 soft contact layers compress 2.76 cm


def test_pyramid20_colors_match_oracle(gpu_available):
    sc = scenes.pyramid(20, density=1000.0)
    g, o = _pair(sc)
    for _ in range(10):
        g.step(1.0 / 60.0, 4)
        o.step(1.0 / 60.0, 4)
    gc, gl, ga = g.contact_colors()
    oc, ol, oa = o.contact_colors()
    n = min(len(gc), len(oc))
    assert np.array_equal(ga[:n], oa[:n])
    live = oa[:n] == 1
    assert np.array_equal(gc[:n][live], oc[:n][live])
    assert np.array_equal(gl[:n][live], ol[:n][live])


@pytest.mark.parametrize("name,builder", [("large_pyramid", lambda: scenes.pyramid(100)), ("many_pyramids", scenes.many_pyramids)])
def test_golden_hash_on_gpu(gpu_available, name, builder):
    from natsel_amd.world import load_scene

    g = GOLDEN["scenes"][name]
    w = load_scene(builder())
    w.step_n(1.0 / 60.0, 4, g["steps"])
    assert w.profile()["fast_bodies"] == 0
    assert w.world_state_hash() == int(g["state_hash"], 16)


def test_rollback_golden_hash_on_gpu(gpu_available):
    """rollback (reference-hashes.json:90-94) on the GPU world: snapshots saved every step,
    restored and replayed every 30 steps; the end hash is the large pyramid's."""
    from natsel_amd.world import load_scene

    g = GOLDEN["scenes"]["rollback"]
    w = scenes.run_rollback(load_scene(scenes.pyramid(100)), g["steps"])
    assert w.world_state_hash() == int(g["state_hash"], 16)


def test_snapshot_restores_a_pile_with_joints(gpu_available):
    """A restored snapshot replays bit for bit: the chain pile (contacts, overflow, joints)
    saved at step 60, stepped 40 more, restored and stepped 40 again ends on the same hash,
    which the oracle reaches too."""
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = scenes.chain_pile(count=200, chains=20)
    w = load_scene(sc)
    o = pyoracle.load_scene(sc)
    for _ in range(60):
        w.step(1.0 / 60.0, 4)
        o.step(1.0 / 60.0, 4)
    s = w.save_state()
    for _ in range(40):
        w.step(1.0 / 60.0, 4)
        o.step(1.0 / 60.0, 4)
    first = w.world_state_hash()
    assert first == o.world_state_hash()
    w.restore_state(s)
    for _ in range(40):
        w.step(1.0 / 60.0, 4)
    assert w.world_state_hash() == first


def test_small_mixed_pile_every_step_matches_oracle(gpu_available):
    """Boxes, spheres and capsules dropped into a container: exercises every pair type
    (hull-hull SAT, GJK hull-sphere and hull-capsule, capsule-capsule clipping,
    sphere-sphere) and contact begin/end transitions."""
    sc = scenes.pile(300, seed=7, size_x=6, size_y=10, size_z=6)
    g, o = _pair(sc)
    for k in range(240):
        g.step(1.0 / 60.0, 4)
        o.step(1.0 / 60.0, 4)
        hg, ho = g.world_state_hash(), o.world_state_hash()
        assert hg == ho, f"step {k}: {hg:x} != {ho:x}; {_first_difference(g, o)}"


def test_overflow_hub_every_step_matches_oracle(gpu_available):
    """The overflow colour (solved in list order by one lane) against the oracle, and the
    reference's known counts after 10 steps (overflow_color_tests.rs:82-103)."""
    sc = scenes.overflow_hub(enable_sleep=False)
    g, o = _pair(sc)
    for k in range(60):
        g.step(1.0 / 60.0, 4)
        o.step(1.0 / 60.0, 4)
        hg, ho = g.world_state_hash(), o.world_state_hash()
        assert hg == ho, f"step {k}: {hg:x} != {ho:x}; {_first_difference(g, o)}"
        if k == 9:
            color, _, alive = g.contact_colors()
            assert int(alive.sum()) == 65
            assert int(((color == 23) & (alive == 1)).sum()) == 5


def test_graph_scene_colours_on_gpu(gpu_available):
    """graph_tests.rs:152-172 colours of the touching contacts after the first step."""
    from natsel_amd.world import load_scene

    w = load_scene(scenes.graph_scene())
    w.step(1.0 / 60.0, 4)
    color, local, alive = w.contact_colors()
    assert len(color) == 8 and alive.all()
    for cid, (c, l) in {0: (22, 0), 1: (22, 1), 2: (22, 2), 3: (0, 0), 5: (1, 0)}.items():
        assert (int(color[cid]), int(local[cid])) == (c, l), cid


def test_step_kernel_workgroup_count_does_not_change_results(gpu_available):
    """The step kernel's grid-stride loops and grid barriers give the same bits for any
    number of workgroups, equal to the oracle's."""
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = scenes.pile(300, seed=7, size_x=6, size_y=10, size_z=6)
    counts = (1, 2, 7, 32, 48)
    worlds = [load_scene(sc, solver_blocks=b) for b in counts]
    o = pyoracle.load_scene(sc)
    for k in range(90):
        o.step(1.0 / 60.0, 4)
        hashes = set()
        for w in worlds:
            w.step(1.0 / 60.0, 4)
            hashes.add(w.world_state_hash())
        assert hashes == {o.world_state_hash()}, f"step {k}: {sorted(hashes)}"
    for b, w in zip(counts, worlds):
        assert w.kernel_stats()["solver_blocks"] == b


def test_small_colour_groups_do_not_change_results(gpu_available):
    """Runs of small colours on one workgroup (workgroup barriers between them) give the
    same bits as a grid barrier after every colour, for any size threshold, with joints in
    the colours; fewer grid barriers as the threshold grows."""
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = scenes.chain_pile(120, chains=8, links=8, size_x=6, size_y=6, size_z=6)
    limits = (0, 16, 256, 100000)
    worlds = [load_scene(sc, small_colour_items=s) for s in limits]
    o = pyoracle.load_scene(sc)
    for k in range(120):
        o.step(1.0 / 60.0, 4)
        hashes = set()
        for w in worlds:
            w.step(1.0 / 60.0, 4)
            hashes.add(w.world_state_hash())
        assert hashes == {o.world_state_hash()}, f"step {k}: {sorted(hashes)}"
    barriers = [w.kernel_stats()["barriers"] for w in worlds]
    assert barriers[0] > barriers[2] >= barriers[3], barriers


def test_kinematic_platform_every_step_matches_oracle(gpu_available):
    """A kinematic slab carrying boxes: the step keeps warm starting colour by colour when an
    awake body is not dynamic (a kinematic body can sit in several items of one colour)."""
    sc = scenes.kinematic_platform()
    g, o = _pair(sc)
    for k in range(300):
        g.step(1.0 / 60.0, 4)
        o.step(1.0 / 60.0, 4)
        hg, ho = g.world_state_hash(), o.world_state_hash()
        assert hg == ho, f"step {k}: {hg:x} != {ho:x}; {_first_difference(g, o)}"
    assert abs(g.body_arrays()[0][1][0] - 2.5) < 1e-4  # the platform moved at its own speed
