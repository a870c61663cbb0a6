"""The device narrow phase (nsl_collide) against the reference's collision tests and the oracle.

Every collide check of tests/narrowphase_cases.py (the reference's known answers, sweeps and
brute-force oracles, and the closed forms) runs on the HIP path; and seeded random queries
of all six pair types, first with empty caches and then with the caches the first round
returned (the way a contact carries them from step to step), equal the oracle's outputs bit
for bit: normal, points, separations, feature ids and every cache field."""
import numpy as np
import pytest

import narrowphase_cases as C
from oracle import pyoracle

pytestmark = pytest.mark.gpu

PAIRS = [("box", "box"), ("box", "sphere"), ("box", "capsule"), ("capsule", "sphere"), ("capsule", "capsule"),
         ("sphere", "sphere")]


def _gpu():
    from natsel_amd import queries

    return queries.Narrow(0)


@pytest.mark.parametrize("name", sorted(C.COLLIDE_CHECKS) + ["rb_edge_axis_oracle"])
def test_gpu_narrowphase_checks(gpu_available, name):
    nar = _gpu()
    if name == "rb_edge_axis_oracle":
        failures = C.rb_edge_axis_oracle(nar, pyoracle.Narrow().box_hull)
    else:
        failures = C.COLLIDE_CHECKS[name](nar)
    assert not failures, f"{name}: {failures[:3]}"


def _unit_quat(rng):
    q = rng.normal(size=4).astype(np.float32)
    return tuple(float(x) for x in q / np.float32(np.linalg.norm(q)))


def _shape(rng, kind):
    if kind == "box":
        half = rng.uniform(0.1, 0.7, 3)
        if rng.random() < 0.5:
            return pyoracle.shape_box(*half)
        return pyoracle.shape_box(*half, p=tuple(rng.uniform(-0.2, 0.2, 3)), q=_unit_quat(rng))
    if kind == "sphere":
        return pyoracle.shape_sphere(tuple(rng.uniform(-0.2, 0.2, 3)), float(rng.uniform(0.05, 0.5)))
    c1 = rng.uniform(-0.4, 0.4, 3)
    c2 = c1 if rng.random() < 0.1 else rng.uniform(-0.4, 0.4, 3)
    return pyoracle.shape_capsule(tuple(c1), tuple(c2), float(rng.uniform(0.05, 0.4)))


def _queries(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        ka, kb = PAIRS[k % len(PAIRS)]
        a, b = _shape(rng, ka), _shape(rng, kb)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        xf = (tuple(float(x) for x in d * rng.uniform(0.0, 1.4)), _unit_quat(rng) if rng.random() < 0.8 else (0, 0, 0, 1))
        out.append((a, b, xf))
    return out


def _perturb(xf, rng):
    p = np.asarray(xf[0], np.float32) + rng.uniform(-0.01, 0.01, 3).astype(np.float32)
    return (tuple(float(x) for x in p), xf[1])


def _same(g, o, kind_pair):
    """Bitwise equality of a GPU and an oracle result, or the first difference."""
    f32 = lambda x: np.asarray(x, np.float32).view(np.uint32)  # noqa: E731
    if len(g["points"]) != len(o["points"]):
        return f"point count {len(g['points'])} vs {len(o['points'])}"
    if g["points"] and not np.array_equal(f32(g["normal"]), f32(o["normal"])):
        return f"normal {g['normal']} vs {o['normal']}"
    for j, (gp, op) in enumerate(zip(g["points"], o["points"])):
        if not np.array_equal(f32(gp[0]), f32(op[0])) or f32(gp[1]) != f32(op[1]) or gp[3] != op[3]:
            return f"point {j}: {gp} vs {op}"
    gc, oc = g["cache"], o["cache"]
    if kind_pair == ("box", "box"):
        gs, os_ = gc.sat, oc.sat
        if f32(gs["separation"]) != f32(os_["separation"]) or any(gs[k] != os_[k] for k in ("type", "index_a", "index_b", "hit")):
            return f"SAT cache {gs} vs {os_}"
    elif kind_pair[0] == "box":
        gs, os_ = gc.simplex, oc.simplex
        n = os_["count"]
        if f32(gs["metric"]) != f32(os_["metric"]) or gs["count"] != n or gs["index_a"][:n] != os_["index_a"][:n] or \
                gs["index_b"][:n] != os_["index_b"][:n]:
            return f"simplex cache {gs} vs {os_}"
    return None


def test_gpu_collide_equals_oracle_bitwise(gpu_available):
    from natsel_amd import queries

    rng = np.random.default_rng(77)
    qs = _queries(2024, 3000)
    orc = pyoracle.Narrow()
    caches_g = [queries.CollideCache() for _ in qs]
    caches_o = [orc.cache() for _ in qs]
    touching = 0
    for rnd in range(2):
        if rnd == 1:  # the second round carries the first round's caches over a small motion
            qs = [(a, b, _perturb(xf, rng)) for a, b, xf in qs]
        got = queries.collide([(a, b, xf, c) for (a, b, xf), c in zip(qs, caches_g)])
        bad = []
        for k, ((a, b, xf), g, co) in enumerate(zip(qs, got, caches_o)):
            o = orc.collide(a, b, xf, co)
            g = dict(g, points=[(p, s, None, fid) for p, s, fid in g["points"]])
            touching += bool(o["points"])
            why = _same(g, o, (a[0], b[0]))
            if why:
                bad.append(f"round {rnd} query {k} {a[0]}-{b[0]}: {why}")
        assert not bad, f"{len(bad)} of {len(qs)} differ: {bad[:4]}"
    assert touching > 1500  # most queries produce a manifold
