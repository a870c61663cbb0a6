"""Benchmark: body-substeps/s of the rigid-body step on the 10,000-body random pile.

Workload (one GPU): 10,000 dynamic bodies (1/3 boxes, 1/3 spheres, 1/3 capsules,
sizes U[0.25, 0.75] m, density 1000, friction U[0.4, 0.8]) dropped from seeded
lattice cells of a 20 x 60 x 20 m region into a static 5-wall container; dt = 1/60,
4 substeps, f32, continuous collision and sleeping off; the pile's run is 1000 steps. A
"step" is one World::step (pair update, narrow phase, constraint preparation, 4 soft-step
substeps, impulse storage, finalize). The timed K steps are the last K of the 1000-step run
(the settled pile, ~33k constraints and ~57k contact points per step): the first 1000 - K steps (at least W)
run untimed, then K steps are timed with a device synchronise on both sides.

value = dynamic bodies x substeps x K / seconds (inputs resident in HBM: the scene is
uploaded before the timed region). The roofline object reports the dominant kernel (the
step kernel): algorithmic bytes (the byte model below) / its summed launch time from HIP
events. cpu_baseline is the CPU oracle (a restatement of the reference step) run through
the same first 1000 - K steps on this host and timed on the same K steps from the same
state (minimum of 3 repeats), with its hashes compared with the GPU's.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
TRIAL_STEPS = 5         # steps per launch-shape trial of the pile's step kernel
BATCH_TRIAL_STEPS = 60  # steps per launch-shape trial of the batch
BATCH_GRAPH_STEPS = 8   # steps of one captured batch graph (kBatchGraphSteps)

# Algorithmic bytes per item: distinct fields the reference formula reads plus writes,
# f32 = 4 B, each body-state gather counted per constraint touch. Coloured constraint
# stages: per constraint (one manifold) + per manifold point.
# Revolute joints (per joint, same rule): prepare reads two bodies' pose, centre, mass and
# world inertia (2 x 92) and the definition (104) and writes the prepared rows (224); warm
# start reads masses, inertias, anchors, axes and impulses (184) and both states (128) and
# writes two velocities (48); push and relax read the prepared rows and impulses (224) and
# both states and write the impulses and axes (60) and two velocities.
BYTE_MODEL = {
    "prepare": {"per_constraint": 268 + 452, "per_point": 0, "per_joint": 184 + 104 + 224},
    "integrate_velocities": {"per_body": 100 + 24},
    "warm_start": {"per_constraint": 64 + 72 + 56 + 48, "per_point": 28, "per_joint": 184 + 128 + 48},
    "push": {"per_constraint": 68 + 12 + 112 + 48, "per_point": 40, "per_joint": 224 + 60 + 128 + 48},
    "integrate_positions": {"per_body": 56 + 56},
    "relax": {"per_constraint": 68 + 108 + 112 + 48, "per_point": 52, "per_joint": 224 + 60 + 128 + 48},
    "store": {"per_constraint": 76, "per_point": 16},
    "finalize": {"per_body": 400},
    # contact record read and written (~300 B each way), both bodies' poses (2 x 64 B),
    # both fat AABBs (2 x 24 B)
    "collide": {"per_contact": 600 + 128 + 48},
}
PER_SUBSTEP = ("integrate_velocities", "warm_start", "push", "integrate_positions", "relax")


def stage_bytes(stage: str, ks: dict, substeps: int) -> float:
    """Algorithmic bytes of one stage summed over the measured steps."""
    m = BYTE_MODEL[stage]
    reps = substeps if stage in PER_SUBSTEP else 1
    if "per_body" in m:
        return reps * m["per_body"] * ks["bodies"]
    if "per_contact" in m:
        return m["per_contact"] * ks["collided"]
    return reps * (m["per_constraint"] * ks["constraints"] + m["per_point"] * ks["points"] +
                   m.get("per_joint", 0) * ks.get("joints", 0))


def step_bytes(counts: dict, substeps: int) -> float:
    """Algorithmic bytes of whole steps (collide + prepare .. finalize) for the given work
    counts (bodies, collided, constraints, points summed over the steps)."""
    return sum(stage_bytes(k, counts, substeps) for k in BYTE_MODEL)


def solve_bytes(ks: dict, substeps: int) -> float:
    """Algorithmic bytes of the step kernel (the substep stages and restitution; prepare runs
    before it and store/finalize after it as launches of their own) over the measured steps."""
    return sum(stage_bytes(k, ks, substeps) for k in BYTE_MODEL if k not in ("collide", "prepare", "store", "finalize"))


TRAFFIC_PROFILE = os.path.join("profiles", "k_solve_step_pile10k_pmc_bytes.md")


BATCH_TRAFFIC_PROFILE = os.path.join("profiles", "k_batch_step_4096_trace_pmc.md")


def measured_traffic(profile: str = TRAFFIC_PROFILE, key: str = "traffic_bytes_per_launch"):
    """Fabric-side bytes (per step-kernel launch, or per batch step) from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by a copy kernel of known size), or None."""
    try:
        text = open(os.path.join(ROOT, profile)).read()
    except OSError:
        return None
    m = re.search(r"^" + key + r": (\d+)", text, re.M)
    return float(m.group(1)) if m else None


def _progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


PILE_SCHEDULE = 1000  # the pile's run: 1000 steps from the drop


def settle_steps(steps: int, warmup: int) -> int:
    """Untimed steps before the timed window: the timed K steps are the last K of the pile's
    1000-step run (the settled pile), and never fewer than the W warmup steps precede them."""
    return max(warmup, PILE_SCHEDULE - steps)


def _timed_steps(w, scene, k: int) -> float:
    t0 = time.perf_counter()
    for _ in range(k):
        w.step(scene.dt, scene.substeps)
    return time.perf_counter() - t0


def _oracle_rate(w, scene, snap, k: int, threads: int, repeats: int = 3) -> float:
    """Seconds for `k` oracle steps from snapshot `snap` on `threads` threads, the minimum of
    `repeats` runs (each restored from the snapshot)."""
    from oracle import pyoracle

    pyoracle.set_threads(threads)
    best = float("inf")
    for _ in range(repeats):
        w.restore_state(snap)
        best = min(best, _timed_steps(w, scene, k))
    return best


def cpu_baseline(scene, settle: int, k: int, seconds: float, gpu_hashes: dict) -> dict:
    """The oracle (a C++ restatement of the reference step, OpenMP over the loops whose
    items are independent) on this host, timed on the same steps from the same state as the
    GPU's timed window: it runs the pile's first `settle` steps untimed on all threads (its
    hash there is compared with the GPU's), saves its state, then times the window's `k`
    steps three times from that state (minimum reported), on all threads and on one thread.
    The one-thread sample is cut to fewer steps when `k` would not fit in its share of
    `seconds`."""
    from oracle import pyoracle

    threads = pyoracle.max_threads()
    pyoracle.set_threads(threads)
    w = pyoracle.load_scene(scene)
    settle_s = _timed_steps(w, scene, settle)
    settle_equal = w.world_state_hash() == gpu_hashes["settle"]
    snap = w.save_state()
    reps = []
    for _ in range(3):
        w.restore_state(snap)
        p0 = w.phase_ms()
        reps.append(_timed_steps(w, scene, k))
    end_equal = w.world_state_hash() == gpu_hashes["end"]
    phases = {key: v - p0[key] for key, v in w.phase_ms().items()}
    pyoracle.set_threads(1)
    w.restore_state(snap)
    one = _timed_steps(w, scene, 1)
    k1 = int(max(1, min(k, 0.3 * seconds / 3 / max(one, 1e-6))))
    reps1 = []
    for _ in range(3):
        w.restore_state(snap)
        reps1.append(_timed_steps(w, scene, k1))
    pyoracle.set_threads(threads)
    n = scene.dynamic_count * scene.substeps
    return {
        "value": n * k / min(reps),
        "unit": "body-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"CPU oracle (C++ restatement of the reference step, g++ -O3, OpenMP, {threads} threads): steps "
                  f"{settle + 1}..{settle + k} of the pile from the same settled state as the GPU's timed steps, "
                  "minimum of 3 repeats",
        "steps": k,
        "seconds_min": min(reps),
        "seconds_repeats": reps,
        "phase_ms_per_step": {key: round(v / k, 3) for key, v in phases.items()},
        "single_thread": {"value": n * k1 / min(reps1), "steps": k1, "seconds_min": min(reps1), "seconds_repeats": reps1,
                          "sample": f"steps {settle + 1}..{settle + k1} from the same state, minimum of 3 repeats"},
        "full_schedule_settle": {"steps": settle, "seconds": settle_s,
                                 "value": n * settle / settle_s if settle_s > 0 else None},
        "hash_equal_to_gpu_after_settle": settle_equal,
        "hash_equal_to_gpu_after_window": end_equal,
    }


def batch_cpu_baseline(scene_list: list, steps: int, gpu_hashes, gpu_body_substeps: float, every: int = 64) -> dict:
    """The oracle on a seeded sample of the batch's worlds (every `every`-th): each world
    stepped once untimed, as the GPU's first step, then the same `steps` steps timed, on all
    host threads (one world per thread, worlds in parallel) and on one thread, the minimum
    of 3 repeats from per-world snapshots; the sample's hashes after the steps against the
    GPU's. Rates in body-substeps/s (the sample's worlds differ in size from the whole
    batch's mean) and world-steps/s."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import pyoracle

    sample = list(range(0, len(scene_list), every))
    threads = pyoracle.max_threads()
    pyoracle.set_threads(1)  # one OpenMP thread per world; the worlds run on the pool's threads
    ws = [pyoracle.load_scene(scene_list[i]) for i in sample]
    for i, w in zip(sample, ws):
        w.step(scene_list[i].dt, scene_list[i].substeps)
    snaps = [w.save_state() for w in ws]

    def run(j):
        sc = scene_list[sample[j]]
        pyoracle.set_threads(1)  # the OpenMP thread count is per calling thread
        ws[j].restore_state(snaps[j])
        for _ in range(steps):
            ws[j].step(sc.dt, sc.substeps)

    def timed(pool_threads):
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            if pool_threads == 1:
                for j in range(len(ws)):
                    run(j)
            else:
                with ThreadPoolExecutor(max_workers=pool_threads) as ex:
                    list(ex.map(run, range(len(ws))))
            best = min(best, time.perf_counter() - t0)
        return best

    t_n = timed(threads)
    equal = all(int(ws[j].world_state_hash()) == int(gpu_hashes[sample[j]]) for j in range(len(ws)))
    t_1 = timed(1)
    pyoracle.set_threads(threads)
    body_substeps = sum(scene_list[i].dynamic_count for i in sample) * scene_list[0].substeps * steps
    return {
        "value": body_substeps / t_n, "unit": "body-substeps/s", "cores": threads, "kind": "port",
        "world_steps_per_s": len(sample) * steps / t_n,
        "sample": f"CPU oracle on {len(sample)} of the {len(scene_list)} worlds (every {every}th), {steps} steps after one "
                  f"untimed step, the worlds spread over {threads} host threads (one OpenMP thread each), minimum of 3 repeats",
        "single_thread": {"value": body_substeps / t_1, "world_steps_per_s": len(sample) * steps / t_1},
        "hash_equal_to_gpu_for_sample": equal,
        "gpu_over_cpu": gpu_body_substeps / (body_substeps / t_n),
    }


def batched_measurement(worlds: int, steps: int, autotune: bool = True, cpu: bool = True) -> dict:
    """Batched worlds (the log-uniform 8..256-body piles and stacks of
    natsel_amd.scenes.batched): every world stepped together, `steps` steps in one call
    after one untimed step; device time from HIP events around the call. With `autotune`,
    the launch shape is first chosen by nsl_batch_autotune (60 steps per candidate from a
    snapshot of the batch after the untimed step, which is then restored)."""
    from natsel_amd import scenes
    from natsel_amd.batch import load_batch

    scene_list = scenes.batched(worlds)
    dt, sub = scene_list[0].dt, scene_list[0].substeps
    b = load_batch(scene_list)
    b.step(dt, sub, 1)  # uploads the worlds
    tuning = b.autotune(dt, sub, BATCH_TRIAL_STEPS) if autotune else None
    # each candidate runs a warm (capture) call and a timed call from the snapshot
    trial_steps = tuning["candidates"] * (min(BATCH_TRIAL_STEPS, BATCH_GRAPH_STEPS) + BATCH_TRIAL_STEPS) if tuning else 0
    s0 = b.stats()
    b.step(dt, sub, steps)
    s1 = b.stats()
    gpu_hashes = b.state_hashes()
    b.close()
    seconds = s1["last_step_ms"] / 1e3
    bodies = s1["bodies"]
    counts = {"bodies": bodies * steps, "steps": steps, "collided": s1["contacts_collided"] - s0["contacts_collided"],
              "constraints": s1["constraints"] - s0["constraints"], "points": s1["points"] - s0["points"]}
    total = step_bytes(counts, sub)
    achieved = total / seconds / 1e9
    return {
        "metric": "world-steps/s (batched worlds)",
        "value": worlds * steps / seconds,
        "unit": "world-steps/s",
        "body_substeps_per_s": bodies * sub * steps / seconds,
        "ms_per_step": 1e3 * seconds / steps,
        "steps": steps,
        "config": {"workload": f"{worlds} independent worlds, {bodies} dynamic bodies (8..256 per world, log-uniform; "
                               "piles and stacks of boxes/spheres/capsules on ground boxes), dt 1/60, 4 substeps",
                   "worlds": worlds, "bodies": bodies},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(BATCH_TRAFFIC_PROFILE, "traffic_bytes_per_step"),
                     "traffic_unit": "bytes per step (the three launches)", "traffic_source": BATCH_TRAFFIC_PROFILE,
                     "kernel": "k_batch_pairs + k_batch_collide + k_batch_solve (one step)",
                     "algorithmic_bytes_per_step": total / steps},
        "contacts_collided_per_step": counts["collided"] / steps,
        "constraints_per_step": counts["constraints"] / steps,
        "points_per_step": counts["points"] / steps,
        "error_worlds": s1["error_worlds"],
        "autotune": tuning,
        "cpu_baseline": batch_cpu_baseline(scene_list, steps, gpu_hashes, bodies * sub * steps / seconds) if cpu else None,
        "dispatch_window": [2 + trial_steps, 1 + trial_steps + steps],
        "dispatch_window_note": "each batch kernel runs once per step: one untimed step, the launch-shape trials' "
                                f"{trial_steps} steps, then the timed steps (this process's first batch)",
    }


# World options every measured world shares (set from the command line).
WORLD_OPTS: dict = {}


def pyramid_measurement(steps: int = 600, cpu: bool = True) -> dict:
    """The pyramid of base 20 (210 boxes) with sleep and continuous collision off: one
    untimed step, then `steps` steps timed on the GPU and on the oracle (all host threads and
    one thread, minimum of 3 repeats from the snapshot), hashes compared."""
    import dataclasses

    import torch

    from natsel_amd import scenes
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = dataclasses.replace(scenes.pyramid(20), enable_sleep=False, enable_continuous=False)
    g = load_scene(sc, **WORLD_OPTS)
    g.step(sc.dt, sc.substeps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.step(sc.dt, sc.substeps)
    torch.cuda.synchronize()
    t_g = time.perf_counter() - t0
    gpu_hash = g.world_state_hash()
    g.close()
    n = sc.dynamic_count * sc.substeps * steps
    out = {"metric": "body-substeps/s (pyramid, base 20)", "value": n / t_g, "unit": "body-substeps/s",
           "ms_per_step": 1e3 * t_g / steps, "steps": steps,
           "config": {"workload": f"pyramid of base 20 ({sc.dynamic_count} boxes), dt 1/60, 4 substeps, sleep and "
                                  "continuous collision off; one untimed step first", "bodies": sc.dynamic_count}}
    if cpu:
        threads = pyoracle.max_threads()
        w = pyoracle.load_scene(sc)
        pyoracle.set_threads(threads)
        w.step(sc.dt, sc.substeps)
        snap = w.save_state()
        t_n = _oracle_rate(w, sc, snap, steps, threads)
        equal = w.world_state_hash() == gpu_hash
        t_1 = _oracle_rate(w, sc, snap, steps, 1)
        pyoracle.set_threads(threads)
        out["cpu_baseline"] = {"value": n / t_n, "unit": "body-substeps/s", "cores": threads, "kind": "port",
                               "sample": f"CPU oracle, the same {steps} steps, minimum of 3 repeats",
                               "single_thread": {"value": n / t_1}, "hash_equal_to_gpu": equal,
                               "gpu_over_cpu": (n / t_g) / (n / t_n)}
    return out


def large_world_cpu_baseline(sc, seconds: float) -> dict:
    """The oracle on the large world: one untimed step, then a bounded sample of k steps
    (sized to about a third of `seconds` per repeat) timed on all host threads and fewer on one
    thread, minimum of 3 repeats from the snapshot; the GPU over the same steps 2..k + 1 from
    the scene's start, and its hash after them against the oracle's."""
    import torch

    from natsel_amd.world import load_scene
    from oracle import pyoracle

    threads = pyoracle.max_threads()
    pyoracle.set_threads(threads)
    w = pyoracle.load_scene(sc)
    w.step(sc.dt, sc.substeps)
    snap = w.save_state()
    one = _timed_steps(w, sc, 1)
    k = int(max(2, min(60, seconds / 3 / max(one, 1e-6))))
    t_n = _oracle_rate(w, sc, snap, k, threads)
    cpu_hash = w.world_state_hash()
    w.restore_state(snap)
    pyoracle.set_threads(1)
    one1 = _timed_steps(w, sc, 1)
    k1 = int(max(1, min(k, seconds / 6 / max(one1, 1e-6))))
    t_1 = _oracle_rate(w, sc, snap, k1, 1)
    pyoracle.set_threads(threads)
    del w, snap
    g = load_scene(sc, **WORLD_OPTS)
    g.step(sc.dt, sc.substeps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        g.step(sc.dt, sc.substeps)
    torch.cuda.synchronize()
    t_g = time.perf_counter() - t0
    equal = g.world_state_hash() == cpu_hash
    g.close()
    n = sc.dynamic_count * sc.substeps
    return {
        "value": n * k / t_n, "unit": "body-substeps/s", "cores": threads, "kind": "port",
        "sample": f"CPU oracle, steps 2..{k + 1} from the scene's start on {threads} threads (minimum of 3 repeats); "
                  f"one thread: steps 2..{k1 + 1}",
        "single_thread": {"value": n * k1 / t_1, "steps": k1},
        "gpu_same_steps": {"value": n * k / t_g, "gpu_over_cpu": (n * k / t_g) / (n * k / t_n)},
        "hash_equal_to_gpu_after_sample": equal,
    }


def large_world_measurement(steps: int, cpu_seconds: float = 0.0) -> dict:
    """The large single world: 100,000 mixed bodies in a seeded pile plus 2,000 chains of 16
    capsule links (30,000 revolute joints), steps 1..`steps` from the scene's start. One run
    for the rate, a second with the step kernel's phase clock on for the per-stage
    algorithmic bytes / time (HBM GB/s by the byte model)."""
    import torch

    from natsel_amd import scenes
    from natsel_amd.world import load_scene

    sc = scenes.large_world()
    n = sc.dynamic_count
    runs = []
    for timing in (False, True):
        world = load_scene(sc, **WORLD_OPTS)
        world.step(sc.dt, sc.substeps)  # uploads the world
        world.reset_kernel_stats()
        world.set_kernel_timing(timing)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            world.step(sc.dt, sc.substeps)
        torch.cuda.synchronize()
        runs.append((time.perf_counter() - t0, world.kernel_stats()))
        _progress(f"large world: {steps} steps, phase clock {'on' if timing else 'off'}")
        world.close()
    elapsed, _ = runs[0]
    ks = runs[1][1]
    stage_ms = {
        "prepare": ks["ms"]["prepare"],
        "integrate_velocities": ks["phase_ms"]["integrate_velocities"],
        "warm_start": ks["phase_ms"]["warm_start"],
        "push": ks["phase_ms"]["push"],
        "integrate_positions": ks["phase_ms"]["integrate_positions"],
        "relax": ks["phase_ms"]["relax"],
        "collide": ks["ms"]["collide"],
    }
    per_stage = {}
    for st, ms in stage_ms.items():
        b = stage_bytes(st, ks, sc.substeps)
        per_stage[st] = {"ms_per_step": ms / steps, "gb_per_step": b / steps / 1e9, "gbs": b / (ms / 1e3) / 1e9 if ms > 0 else None}
    sf_bytes = stage_bytes("store", ks, sc.substeps) + stage_bytes("finalize", ks, sc.substeps)
    sf_ms = ks["ms"]["finalize"]
    per_stage["store_finalize"] = {"ms_per_step": sf_ms / steps, "gb_per_step": sf_bytes / steps / 1e9,
                                   "gbs": sf_bytes / (sf_ms / 1e3) / 1e9 if sf_ms > 0 else None}
    return {
        "metric": "body-substeps/s (large world)",
        "value": n * sc.substeps * steps / elapsed,
        "unit": "body-substeps/s",
        "ms_per_step": 1e3 * elapsed / steps,
        "steps": steps,
        "config": {"workload": f"{n} dynamic bodies: a 100,000-body pile (boxes/spheres/capsules, 50 x 60 x 50 m) and "
                               f"{len(sc.joints) // 15} chains of 16 capsule links ({len(sc.joints)} revolute joints), dt 1/60, "
                               "4 substeps, sleep and continuous collision off", "bodies": n, "joints": len(sc.joints)},
        "per_stage_hbm": per_stage,
        "overflow_ms_per_step": ks["phase_ms"]["overflow"] / steps,
        "host_ms_per_step": {k: v / steps for k, v in ks["host_ms"].items()},
        "points_per_step": ks["points"] / max(ks["steps"], 1),
        "constraints_per_step": ks["constraints"] / max(ks["steps"], 1),
        "note": "per-stage figures from the run with the phase clock on; value from the run without it",
        "cpu_baseline": large_world_cpu_baseline(sc, cpu_seconds) if cpu_seconds > 0 else None,
    }


def defaults_measurement(steps: int, cpu_seconds: float) -> dict:
    """The 10k pile with the reference's defaults (sleep and continuous collision on for the
    world and every body): body-substeps/s counting awake bodies, steps 1..`steps` after one
    warmup step; the oracle on the host threads from step 1 for a bounded sample, counting
    the same way; and the GPU's hash after the sample's steps against the oracle's."""
    import torch

    from natsel_amd import scenes
    from natsel_amd.world import load_scene
    from oracle import pyoracle

    sc = scenes.pile_defaults(10_000)
    world = load_scene(sc, **WORLD_OPTS)
    world.step(sc.dt, sc.substeps)
    world.reset_kernel_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        world.step(sc.dt, sc.substeps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ks = world.kernel_stats()
    summary = world.island_summary()
    world.close()
    awake_substeps = ks["bodies"] * sc.substeps

    pyoracle.set_threads(pyoracle.max_threads())
    ref = pyoracle.load_scene(sc)
    awake, n = 0, 0
    c0 = time.perf_counter()
    while n < steps and time.perf_counter() - c0 < cpu_seconds:
        awake += ref.awake_body_count()  # awake bodies as the step starts (the GPU counts them as solved)
        ref.step(sc.dt, sc.substeps)
        n += 1
    cpu_s = time.perf_counter() - c0
    # the GPU over the oracle's steps 1..n (like for like with the oracle's rate; the first
    # step includes the upload), ending on the hash compared with the oracle's
    check = load_scene(sc, **WORLD_OPTS)
    check.reset_kernel_stats()
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    for _ in range(n):
        check.step(sc.dt, sc.substeps)
    torch.cuda.synchronize()
    gpu_sample_s = time.perf_counter() - g0
    gpu_sample_substeps = check.kernel_stats()["bodies"] * sc.substeps
    same = check.world_state_hash() == ref.world_state_hash()
    check.close()
    return {
        "metric": "body-substeps/s (10k pile, reference defaults)",
        "value": awake_substeps / elapsed,
        "unit": "body-substeps/s",
        "ms_per_step": 1e3 * elapsed / steps,
        "steps": steps,
        "config": {"workload": "the 10k random pile with sleep and continuous collision on (the reference's defaults), "
                               "dt 1/60, 4 substeps; awake bodies counted per step", "bodies": sc.dynamic_count},
        "awake_bodies_per_step": ks["bodies"] / max(ks["steps"], 1),
        "toi_calls": ks["toi_calls"],
        "toi_hits": ks["toi_hits"],
        "islands_at_end": summary,
        "host_ms_per_step": {k: v / steps for k, v in ks["host_ms"].items()},
        "cpu_baseline": {"value": awake * sc.substeps / cpu_s, "unit": "body-substeps/s", "cores": pyoracle.max_threads(),
                         "kind": "port", "sample": f"CPU oracle, steps 1..{n} of the same scene", "steps": n, "seconds": cpu_s},
        "hash_equal_to_oracle_after_sample": same,
        "gpu_same_steps": {"value": gpu_sample_substeps / gpu_sample_s, "steps": n,
                           "gpu_over_cpu": (gpu_sample_substeps / gpu_sample_s) / (awake * sc.substeps / cpu_s),
                           "note": "the GPU over the oracle's steps 1..n, awake body-substeps counted the same way"},
    }


def adjoint_measurement(worlds: int, steps: int, fd_worlds: int = 16) -> dict:
    """Differentiable rollout: `worlds` worlds of 32 cubes pushed across the ground
    (natsel_amd.scenes.pushed_boxes), `steps` steps; loss |com(T) - target|^2; gradients
    w.r.t. every body's initial velocity, angular velocity and friction. Forward = the
    batched stepper's device time for the steps; forward+adjoint = recording the same steps
    plus the adjoint. Gradient check: central differences (epsilon 1e-5) of the same model in
    double precision on `fd_worlds` seeded worlds, with the recorded active set (frozen) and
    with the model deciding its own branches (free)."""
    import numpy as np

    from natsel_amd import scenes
    from natsel_amd.batch import load_batch

    sc = [scenes.pushed_boxes(1000 + k) for k in range(worlds)]
    targets = scenes.pushed_box_targets(worlds)
    dt, sub = sc[0].dt, sc[0].substeps
    fwd = load_batch(sc)
    fwd.step(dt, sub, steps)
    fwd_ms = fwd.stats()["last_step_ms"]
    fwd.close()
    rb = load_batch(sc)
    rb.record(dt, sub, steps)
    r = rb.adjoint(targets)
    rec_ms, adj_ms = rb.rollout_times()
    offs = np.cumsum([0] + [s.dynamic_count for s in sc])
    sample = np.sort(np.random.default_rng(5).choice(worlds, min(fd_worlds, worlds), replace=False))
    wl = np.repeat(sample, 32 * 7)
    pl = np.tile(np.arange(32 * 7), len(sample))
    g = np.concatenate([r["grad_v"], r["grad_w"], r["grad_friction"][:, None]], axis=1)
    adj = np.array([g[offs[w] + p // 7, p % 7] for w, p in zip(wl, pl)])
    eps = 1e-5
    check = {"epsilon": eps, "worlds_sampled": len(sample), "parameters_per_world": 32 * 7}
    for mode in ("frozen", "free"):
        lp = rb.model_losses(np.concatenate([wl, wl]), np.concatenate([pl, pl]),
                             np.concatenate([np.full(len(wl), eps), np.full(len(wl), -eps)]), targets, frozen=mode == "frozen")
        fd = (lp[:len(wl)] - lp[len(wl):]) / (2 * eps)
        errs, within = [], []
        for w in sample:
            m = wl == w
            scale = np.abs(fd[m]).max()
            errs.append(float(np.abs(adj[m] - fd[m]).max() / scale))
            within.append(np.abs(adj[m] - fd[m]) <= 1e-3 * scale)
        check[mode] = {"max_rel_error": max(errs), "median_rel_error": float(np.median(errs)),
                       "components_within_1e-3": float(np.concatenate(within).mean())}
    rb.close()
    check["cpu_reference"] = adjoint_cpu_reference(sc, targets, sample, steps, r, offs)
    fs, fas = fwd_ms / 1e3, (rec_ms + adj_ms) / 1e3
    return {
        "metric": "world-steps/s (differentiable rollout: forward vs forward+adjoint)",
        "forward": worlds * steps / fs,
        "forward_adjoint": worlds * steps / fas,
        "unit": "world-steps/s",
        "fwd_adjoint_over_fwd": fas / fs,
        "ms": {"forward": fwd_ms, "record": rec_ms, "adjoint": adj_ms},
        "config": {"workload": f"{worlds} worlds x 32 cubes pushed across the ground, {steps} steps, dt 1/60, 4 substeps; "
                               "loss |com - target|^2; gradients w.r.t. v0, w0, friction", "worlds": worlds, "steps": steps},
        "status_nonzero_worlds": int((r["status"] != 0).sum()),
        "gradient_check": check,
    }


def adjoint_cpu_reference(sc: list, targets, sample, steps: int, r: dict, offs) -> dict:
    """CPU baseline and check of the differentiable rollout: the oracle's reference
    (oracle/ref_adjoint.cpp: each sampled world recorded on the oracle, the model repeated in
    float and differentiated in double by reverse-mode automatic differentiation, one host
    thread) timed on the sampled worlds, its gradients against the GPU adjoint's, and central
    differences (epsilon 1e-3) of the oracle's own float step (contacts re-detected) for a few
    velocity components of two worlds against the GPU adjoint."""
    import dataclasses

    import numpy as np

    from oracle import pyoracle

    errs, exact, t_ref = [], True, 0.0
    for k in sample:
        t0 = time.perf_counter()
        ref = pyoracle.Rollout(sc[k], steps)
        g = ref.gradient(targets[k])
        t_ref += time.perf_counter() - t0
        ref.close()
        exact = exact and g["exact"] and g["status"] == 0
        a, b = offs[k], offs[k + 1]
        gpu = np.concatenate([r["grad_v"][a:b], r["grad_w"][a:b], r["grad_friction"][a:b, None]], axis=1)
        cpu = np.concatenate([g["grad_v"], g["grad_w"], g["grad_friction"][:, None]], axis=1)
        errs.append(float(np.abs(gpu - cpu).max() / np.abs(cpu).max()))
    step_fd = []
    for k in sample[:2]:
        s = sc[k]
        dyn = [i for i, bd in enumerate(s.bodies) if bd.body_type == 2]
        half = np.array([sh.params[0] for sh in s.shapes if sh.body in dyn])
        mass = 1000.0 * (2.0 * half) ** 3

        def loss_of(scene):
            w = pyoracle.load_scene(scene)
            for _ in range(steps):
                w.step(scene.dt, scene.substeps)
            pos = w.body_arrays()[0][dyn]
            w.close()
            d = (pos * mass[:, None]).sum(0) / mass.sum() - np.asarray(targets[k], np.float64)
            return float(d @ d)

        def perturbed(bi, c, e):
            bodies = list(s.bodies)
            v = list(bodies[dyn[bi]].linear_velocity)
            v[c] = float(np.float32(v[c] + e))
            bodies[dyn[bi]] = dataclasses.replace(bodies[dyn[bi]], linear_velocity=tuple(v))
            return dataclasses.replace(s, bodies=bodies)

        gv = r["grad_v"][offs[k]:offs[k + 1]]
        scale = float(np.abs(gv).max())
        for bi in (0, 11, 27):
            for c in (0, 2):
                fd = (loss_of(perturbed(bi, c, 1e-3)) - loss_of(perturbed(bi, c, -1e-3))) / 2e-3
                step_fd.append(abs(float(gv[bi][c]) - fd) / scale)
    return {"value": len(sample) * steps / t_ref, "unit": "world-steps/s (recording + reference gradient)", "cores": 1,
            "kind": "port", "sample": f"{len(sample)} sampled worlds x {steps} steps, one host thread",
            "model_repeats_recorded_steps_bit_for_bit": exact,
            "gpu_vs_cpu_reference_max_rel_error": max(errs), "gpu_vs_cpu_reference_median_rel_error": float(np.median(errs)),
            "gpu_vs_oracle_step_fd": {"epsilon": 1e-3, "components": len(step_fd), "max_rel_error": max(step_fd),
                                      "median_rel_error": float(np.median(step_fd)),
                                      "note": "error / the world's largest velocity gradient; the step re-detects contacts "
                                              "(the model holds them fixed), in float"}}


def summarize_profile(path: str, kernel: str, window: str) -> dict:
    """Summary of a rocprofv3 CSV (a kernel trace or a --pmc counter collection) over the
    dispatches `window` ("a:b", 1-based, inclusive, counted over this kernel's dispatches in
    dispatch order, or "all") of the kernel named `kernel` (namespace-qualified or not): average duration (us) for
    a trace, average value per counter for a counter collection, and the same over all of
    the kernel's dispatches."""
    import csv

    name = re.compile(r"(^|[\s:])" + re.escape(kernel) + r"($|[(<\s])")  # plain or namespace-qualified
    rows = [r for r in csv.DictReader(open(path)) if name.search(r["Kernel_Name"])]
    if not rows:
        raise SystemExit(f"no dispatch of {kernel} in {path}")
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    n = len({r["Dispatch_Id"] for r in rows})
    a, b = (1, n) if window == "all" else (int(x) for x in window.split(":"))
    out = {"file": os.path.basename(path), "kernel": kernel, "window": [a, b]}
    if "Counter_Name" in rows[0]:
        per_dispatch: dict = {}
        for r in rows:
            per_dispatch.setdefault(int(r["Dispatch_Id"]), {}).setdefault(r["Counter_Name"], 0.0)
            per_dispatch[int(r["Dispatch_Id"])][r["Counter_Name"]] += float(r["Counter_Value"])
        ids = sorted(per_dispatch)
        sel = ids[a - 1:b]
        names = sorted(per_dispatch[ids[0]])
        out["dispatches"] = len(ids)
        out["window_dispatches"] = len(sel)
        out["window_mean"] = {c: sum(per_dispatch[i][c] for i in sel) / len(sel) for c in names} if sel else None
        out["all_mean"] = {c: sum(per_dispatch[i][c] for i in ids) / len(ids) for c in names}
    else:
        dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in rows]
        sel = dur[a - 1:b]
        out["dispatches"] = len(dur)
        out["window_dispatches"] = len(sel)
        out["window_avg_us"] = sum(sel) / len(sel) if sel else None
        out["window_min_us"] = min(sel) if sel else None
        out["window_max_us"] = max(sel) if sel else None
        out["all_avg_us"] = sum(dur) / len(dur)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100,
                    help="timed steps: the last K of the pile's 1000-step run")
    ap.add_argument("--warmup", type=int, default=5,
                    help="untimed steps before the window (at least; the run's first 1000 - K steps are untimed)")
    ap.add_argument("--bodies", type=int, default=10_000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--blocks", type=int, default=0, help="workgroups of the step kernel (0: sized from the largest phase)")
    ap.add_argument("--small-colours", type=int, default=128,
                    help="colours of at most this many items run on one workgroup (0: a grid barrier after every colour)")
    ap.add_argument("--plain-launch", action="store_true",
                    help="launch the step kernel without the cooperative API (same grid, same results): rocprofv3 "
                         "ends before writing its output for a process that made cooperative launches")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default launch shape of the step kernel instead of autotuning it after the settle")
    ap.add_argument("--batched-worlds", type=int, default=4096, help="worlds of the batched measurement (0: skip it)")
    ap.add_argument("--batched-steps", type=int, default=240)
    ap.add_argument("--large-world-steps", type=int, default=500, help="steps of the large-world measurement (0: skip it)")
    ap.add_argument("--pyramid-steps", type=int, default=600, help="steps of the base-20 pyramid measurement (0: skip it)")
    ap.add_argument("--adjoint-worlds", type=int, default=1024, help="worlds of the differentiable rollout (0: skip it)")
    ap.add_argument("--adjoint-steps", type=int, default=64)
    ap.add_argument("--defaults-steps", type=int, default=1000,
                    help="steps of the pile with sleep and continuous collision on (0: skip it)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only copy this many MiB device to device 5 times (reference kernel for rocprofv3 --pmc byte counters)")
    ap.add_argument("--summarize-profile", default="",
                    help="only summarise this rocprofv3 CSV (kernel trace or counter collection) over --kernel's "
                         "dispatches --window a:b, then exit (no GPU used)")
    ap.add_argument("--kernel", default="k_solve_step", help="kernel name prefix (several: comma-separated)")
    ap.add_argument("--window", default="", help="a:b, 1-based dispatches of --kernel (default: the timed window "
                                                 "of a run with these --steps/--warmup)")
    args = ap.parse_args()
    if args.plain_launch:
        WORLD_OPTS["plain_step_launch"] = True
    if args.summarize_profile:
        settle = settle_steps(args.steps, args.warmup)
        window = args.window or f"{settle + 1}:{settle + args.steps}"
        res = [summarize_profile(args.summarize_profile, k, window) for k in args.kernel.split(",")]
        print(json.dumps(res[0] if len(res) == 1 else res, indent=1))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)", file=sys.stderr)
        return 2

    import torch

    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_mib": args.calibrate_copy_mib, "launches": 5}))
        return 0

    from natsel_amd import scenes
    from natsel_amd.world import load_scene

    scene = scenes.pile(args.bodies)
    settle = settle_steps(args.steps, args.warmup)
    world = load_scene(scene, device=0, solver_blocks=args.blocks, small_colour_items=args.small_colours, **WORLD_OPTS)
    # untimed: the pile's first `settle` steps (the drop and the settling), which include the
    # W warmup steps; the timed window is the run's last K steps
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    for k in range(settle):
        world.step(scene.dt, scene.substeps)
        if (k + 1) % 250 == 0:
            _progress(f"pile: untimed step {k + 1}")
    torch.cuda.synchronize()
    settle_s = time.perf_counter() - s0
    # the step kernel's launch shape for the settled pile, by timing (results do not change;
    # the world is restored to the settled state)
    tuning = world.autotune(scene.dt, scene.substeps, TRIAL_STEPS) if args.autotune else None
    # every trial (the previous shape, then each candidate) runs TRIAL_STEPS steps
    trial_steps = (tuning["candidates"] + 1) * TRIAL_STEPS if tuning else 0
    hashes = {"settle": world.world_state_hash()}
    _progress(f"pile: {settle} untimed steps in {settle_s:.2f} s")
    snap = world.save_state()
    world.reset_kernel_stats()
    world.set_kernel_timing(True, phases=False)  # launch events only: the step kernel's HIP-event time
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        world.step(scene.dt, scene.substeps)  # returns after the step's device work
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    world.set_kernel_timing(False)
    ks = world.kernel_stats()
    hashes["end"] = world.world_state_hash()
    # the same steps again from the saved state with the step kernel's phase clock on, for the
    # per-phase breakdown; the restored run must end on the same hash
    world.restore_state(snap)
    world.reset_kernel_stats()
    world.set_kernel_timing(True, phases=True)
    for _ in range(args.steps):
        world.step(scene.dt, scene.substeps)
    world.set_kernel_timing(False)
    ks_phases = world.kernel_stats()
    rerun_equal = world.world_state_hash() == hashes["end"]
    # the solver's per-colour residuals on the first timed step (a diagnostic sweep after the
    # step kernel; untimed, results unchanged)
    world.restore_state(snap)
    world.set_residuals(True)
    world.step(scene.dt, scene.substeps)
    world.set_residuals(False)
    colour_res = [r for r in world.colour_residuals() if r["constraints"] > 0]
    residuals = {"step": settle + 1, "colours": len(colour_res),
                 "max_normal_impulse_change": max((r["max_normal"] for r in colour_res), default=0.0),
                 "sum_normal_impulse_change": sum(r["sum_normal"] for r in colour_res),
                 "max_friction_impulse_change": max((r["max_friction"] for r in colour_res), default=0.0),
                 "per_colour_max_normal": [r["max_normal"] for r in colour_res],
                 "note": "one more relax sweep over the coloured contacts after the step (N s impulses)"}
    world_opts = {"small_colour_items": world.options.small_colour_items, "stage_records": int(world.options.stage_records)}
    _progress(f"pile: {args.steps} timed steps, {1e3 * elapsed / args.steps:.3f} ms per step")
    del snap
    world.close()  # release device memory now, not at interpreter exit
    n = scene.dynamic_count
    value = n * scene.substeps * args.steps / elapsed

    ms = ks["ms"]
    # dominant launch: the step kernel (one per step); its HIP-event time over the timed region
    launches = max(ks["launches"]["solve"], 1)
    dom_s = ms["solve"] / 1e3
    dom_bytes = solve_bytes(ks, scene.substeps)
    achieved = dom_bytes / dom_s / 1e9 if dom_s > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": measured_traffic(),
        "traffic_source": TRAFFIC_PROFILE,
        "kernel": "k_solve_step",
        "launches": ks["launches"]["solve"],
        "avg_launch_us": 1e3 * ms["solve"] / launches,
        "algorithmic_bytes_per_launch": dom_bytes / launches,
        "grid_barriers_per_launch": ks["barriers"] / launches,
        "small_colour_items": world_opts["small_colour_items"],
        "stage_records": world_opts["stage_records"],
        "step_launch": "plain" if args.plain_launch else "cooperative",
        "workgroups": ks["solver_blocks"],
        "autotune": tuning,
    }
    out = {
        "metric": "body-substeps/s (10k pile)",
        "value": value,
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded pile generator)",
        "config": {"workload": f"random pile, {n} dynamic bodies (boxes/spheres/capsules), 5-wall container, dt 1/60, 4 substeps, "
                               "continuous collision and sleep off", "bodies": n, "substeps": scene.substeps,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "kernel_ms": {k: round(v, 3) for k, v in ms.items()},
        "step_kernel_phase_ms": {k: round(v, 3) for k, v in ks_phases["phase_ms"].items()},
        "step_kernel_phase_count": ks_phases["phase_count"],
        "step_kernel_phase_note": "from a second run of the same steps with the step kernel's phase clock on",
        "host_phase_ms": {k: round(v, 3) for k, v in ks["host_ms"].items()},
        "contacts_collided_per_step": ks["collided"] / max(ks["steps"], 1),
        "constraints_per_step": ks["constraints"] / max(ks["steps"], 1),
        "points_per_step": ks["points"] / max(ks["steps"], 1),
        "pair_tests_per_step": ks["pair_tests"] / max(ks["steps"], 1),
        "fast_bodies": ks["fast_bodies"],
        "timed_window": {"first_step": settle + 1, "last_step": settle + args.steps, "untimed_steps_before": settle,
                         "step_kernel_dispatches": [settle + trial_steps + 1, settle + trial_steps + args.steps],
                         "note": "the step kernel, k_collide, k_prepare and k_store_finalize run once per step; after "
                                 "the untimed steps the launch-shape trials run " + str(trial_steps) + " steps, so the "
                                 "timed window is each kernel's dispatches step_kernel_dispatches of this process "
                                 "(rocprofv3 summaries average over those)"},
        "full_schedule": {"steps": settle + args.steps, "seconds": settle_s + elapsed,
                          "value": n * scene.substeps * (settle + args.steps) / (settle_s + elapsed),
                          "note": "the whole run from the drop (first step includes the upload), untimed steps included"},
        "hash_after_settle": f"{hashes['settle']:016x}",
        "hash_after_window": f"{hashes['end']:016x}",
        "rerun_from_snapshot_same_hash": rerun_equal,
        "solver_residuals": residuals,
    }
    if not args.no_cpu:
        cb = cpu_baseline(scene, settle, args.steps, args.cpu_seconds, hashes)
        cb["gpu_over_cpu"] = value / cb["value"]
        out["cpu_baseline"] = cb
        _progress(f"cpu baseline: {cb['value'] / 1e6:.2f} M body-substeps/s on {cb['cores']} threads")
    if args.batched_worlds > 0:
        out["batched"] = batched_measurement(args.batched_worlds, args.batched_steps, args.autotune, not args.no_cpu)
        _progress(f"batched: {out['batched']['ms_per_step']:.3f} ms per step")
    if args.adjoint_worlds > 0:
        out["adjoint"] = adjoint_measurement(args.adjoint_worlds, args.adjoint_steps)
        _progress("adjoint done")
    if args.pyramid_steps > 0:
        out["pyramid"] = pyramid_measurement(args.pyramid_steps, not args.no_cpu)
        _progress(f"pyramid: {out['pyramid']['ms_per_step']:.3f} ms per step")
    if args.large_world_steps > 0:
        out["large_world"] = large_world_measurement(args.large_world_steps, 0.0 if args.no_cpu else args.cpu_seconds)
        _progress(f"large world: {out['large_world']['ms_per_step']:.3f} ms per step")
    if args.defaults_steps > 0:
        out["reference_defaults"] = defaults_measurement(args.defaults_steps, 0.4 * args.cpu_seconds)
        _progress("reference defaults done")
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
