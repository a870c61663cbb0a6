"""Differentiable rollout of batched worlds: the recorded forward is the stepper's, and the
adjoint is the gradient of the model it records.

Worlds of 32 cubes pushed across the ground (natsel_amd.scenes.pushed_boxes), 64 steps of
4 substeps; loss |com(T) - target|^2 per world; gradients w.r.t. every body's initial
velocity, angular velocity and friction.

- Recording does not change the stepping, and the recorded worlds are the oracle's
  (world_state_hash, body_queries.rs:432).
- The adjoint's float recompute of each step lands bit for bit on the recorded next state
  (status 0).
- Against central differences (epsilon 1e-5) of the same model in double precision with the
  recorded active set (frozen): per world max |adjoint - fd| / max |fd| below 1e-4.
- Against central differences of the double model deciding its own clamps (free): the
  tolerance is 1e-3 of the world's largest gradient for 99 % of the components; the rest
  sit at contact-branch switches, where the function has kinks.
"""
import numpy as np
import pytest

from natsel_amd import scenes

pytestmark = pytest.mark.gpu

W, T, EPS = 64, 64, 1e-5


@pytest.fixture(scope="module")
def rollout(gpu_available):
    from natsel_amd.batch import load_batch

    sc = [scenes.pushed_boxes(500 + k) for k in range(W)]
    targets = scenes.pushed_box_targets(W, seed=11)
    a = load_batch(sc)
    a.record(1.0 / 60.0, 4, T)
    r = a.adjoint(targets)
    yield sc, targets, a, r
    a.close()


def test_recorded_forward_is_the_stepper(rollout):
    from natsel_amd.batch import load_batch
    from oracle import pyoracle

    sc, _, a, _ = rollout
    b = load_batch(sc)
    b.step(1.0 / 60.0, 4, T)
    assert np.array_equal(a.state_hashes(), b.state_hashes())
    for k in (0, 17, 63):
        o = pyoracle.load_scene(sc[k])
        for _ in range(T):
            o.step(1.0 / 60.0, 4)
        assert int(a.state_hashes()[k]) == o.world_state_hash(), k
    b.close()


def test_adjoint_recompute_is_exact(rollout):
    _, _, _, r = rollout
    assert (r["status"] == 0).all()
    assert np.isfinite(r["loss"]).all() and np.isfinite(r["grad_v"]).all()


def _fd(rollout, sample, frozen):
    sc, targets, a, r = rollout
    offs = np.cumsum([0] + [s.dynamic_count for s in sc])
    worlds = np.repeat(sample, 32 * 7)
    params = np.tile(np.arange(32 * 7), len(sample))
    g = np.concatenate([r["grad_v"], r["grad_w"], r["grad_friction"][:, None]], axis=1)
    adj = np.array([g[offs[w] + p // 7, p % 7] for w, p in zip(worlds, params)])
    lp = a.model_losses(np.concatenate([worlds, worlds]), np.concatenate([params, params]),
                        np.concatenate([np.full(len(worlds), EPS), np.full(len(worlds), -EPS)]), targets, frozen=frozen)
    fd = (lp[:len(worlds)] - lp[len(worlds):]) / (2 * EPS)
    return worlds, adj, fd


def test_adjoint_matches_frozen_model_differences(rollout):
    sample = np.array([0, 5, 22, 41, 63])
    worlds, adj, fd = _fd(rollout, sample, True)
    for w in sample:
        m = worlds == w
        assert np.abs(adj[m] - fd[m]).max() <= 1e-4 * np.abs(fd[m]).max(), w


def test_adjoint_against_free_model_differences(rollout):
    sample = np.array([1, 9, 30, 47, 58])
    worlds, adj, fd = _fd(rollout, sample, False)
    ok = np.zeros(len(adj), bool)
    for w in sample:
        m = worlds == w
        ok[m] = np.abs(adj[m] - fd[m]) <= 1e-3 * np.abs(fd[m]).max()
    assert ok.mean() >= 0.99, ok.mean()


def test_model_loss_matches_the_recorded_loss(rollout):
    """The double-precision model at the nominal parameters reproduces the float loss."""
    _, targets, a, r = rollout
    base = a.model_losses(np.arange(8), np.zeros(8, np.int32), np.zeros(8), targets)
    assert np.allclose(base, r["loss"][:8], rtol=1e-5)


def test_adjoint_matches_cpu_reference_adjoint(rollout):
    """The GPU adjoint (float) against the oracle's reference gradients of the same model
    (oracle/ref_adjoint.cpp: the oracle's own recording of each world, the model repeated in
    float bit for bit and differentiated in double by reverse-mode automatic
    differentiation) on 16 seeded worlds: per world max |gpu - ref| / max |ref| below 1e-4,
    and the losses within 1e-5 relative."""
    from oracle import pyoracle

    sc, targets, _, r = rollout
    offs = np.cumsum([0] + [s.dynamic_count for s in sc])
    worlds = np.sort(np.random.default_rng(3).choice(W, 16, replace=False))
    errs = []
    for k in worlds:
        ref = pyoracle.Rollout(sc[k], T)
        g = ref.gradient(targets[k])
        ref.close()
        assert g["exact"] and g["status"] == 0, k
        a, b = offs[k], offs[k + 1]
        gpu = np.concatenate([r["grad_v"][a:b], r["grad_w"][a:b], r["grad_friction"][a:b, None]], axis=1)
        cpu = np.concatenate([g["grad_v"], g["grad_w"], g["grad_friction"][:, None]], axis=1)
        errs.append(np.abs(gpu - cpu).max() / np.abs(cpu).max())
        assert abs(float(r["loss"][k]) - g["loss"]) <= 1e-5 * g["loss"] + 1e-9, k
    print(f"GPU adjoint vs CPU reference adjoint: max rel error {max(errs):.2e}, median {np.median(errs):.2e}")
    assert max(errs) < 1e-4
