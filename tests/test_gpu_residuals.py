"""The GPU's per-colour residual sweep (nsl_world_set_residuals / nsl_world_get_residuals)
equals the oracle's (World::colour_residuals) field for field: the wavefront butterfly sums,
the maxima and the constraint counts, on a pile and a pyramid; and turning it on changes no
result (the state hash equals the oracle's every step)."""
import pytest

from natsel_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,make,steps", [
    ("pile1000", lambda: scenes.pile(1000), 200),
    ("pyramid20", lambda: scenes.pyramid(20), 120),
])
def test_residuals_match_oracle(gpu_available, name, make, steps):
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = make()
    g, o = load_scene(sc), pyoracle.load_scene(sc)
    g.set_residuals(True)
    o.set_residuals(True)
    compared = 0
    for k in range(steps):
        g.step(sc.dt, sc.substeps)
        o.step(sc.dt, sc.substeps)
        hg, ho = g.world_state_hash(), o.world_state_hash()
        assert hg == ho, f"{name} step {k + 1}: gpu {hg:x} != oracle {ho:x}"
        rg, ro = g.colour_residuals(), o.colour_residuals()
        for ci, (a, b) in enumerate(zip(rg, ro)):
            assert a == b, f"{name} step {k + 1} colour {ci}: gpu {a} != oracle {b}"
            compared += a["constraints"]
    print(f"{name}: {compared} constraint residuals compared", flush=True)
    assert compared > steps * 10
    g.close()
