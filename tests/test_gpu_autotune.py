"""The step kernel's autotuner (nsl_world_autotune) leaves the world's state as it was and
never changes results: after tuning, the world continues bit for bit as the oracle does."""
import pytest

from natsel_amd import scenes

pytestmark = pytest.mark.gpu


def test_autotune_keeps_state_and_results(gpu_available):
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = scenes.pile(1500, seed=11, size_x=10, size_y=20, size_z=10)
    g, o = load_scene(sc), pyoracle.load_scene(sc)
    for _ in range(150):
        g.step(sc.dt, sc.substeps)
        o.step(sc.dt, sc.substeps)
    before = g.world_state_hash()
    assert before == o.world_state_hash()
    t = g.autotune(sc.dt, sc.substeps, 3)
    assert g.world_state_hash() == before
    assert t["candidates"] >= 6 and t["step_kernel_us"] <= t["previous_us"]
    assert g.options.solver_blocks == t["solver_blocks"]
    for k in range(60):
        g.step(sc.dt, sc.substeps)
        o.step(sc.dt, sc.substeps)
        assert g.world_state_hash() == o.world_state_hash(), f"step {k} after tuning"


def test_plain_step_launch_matches_oracle(gpu_available):
    """The ordinary launch of the step kernel (plain_step_launch) runs the same grid as the
    cooperative one and ends every step on the oracle's hash."""
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = scenes.pile(1500, seed=12, size_x=10, size_y=20, size_z=10)
    g, o = load_scene(sc, plain_step_launch=True), pyoracle.load_scene(sc)
    assert g.options.plain_step_launch
    for k in range(120):
        g.step(sc.dt, sc.substeps)
        o.step(sc.dt, sc.substeps)
        if k % 20 == 19:
            assert g.world_state_hash() == o.world_state_hash(), f"step {k}"

