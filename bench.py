"""Benchmark: body-substeps/s of the rigid-body step on the 10,000-body random pile.

Workload (one GPU): 10,000 dynamic bodies (1/3 boxes, 1/3 spheres, 1/3 capsules,
sizes U[0.25, 0.75] m, density 1000, friction U[0.4, 0.8]) dropped from seeded
lattice cells of a 20 x 60 x 20 m region into a static 5-wall container; dt = 1/60,
4 substeps, f32, continuous collision and sleeping off. A "step" is one World::step
(pair update, narrow phase, constraint preparation, 4 soft-step substeps, impulse
storage, finalize). W warmup steps run untimed, then K steps are timed with a device
synchronise on both sides, as the reference harness times steps 1..N.

value = dynamic bodies x substeps x K / seconds (inputs resident in HBM: the scene is
uploaded before the timed region). The roofline object reports the dominant kernel
class: algorithmic bytes (the byte model below) / its summed launch time from HIP
events. cpu_baseline is the CPU oracle (a restatement of the reference step) timed on a
bounded sample of the same workload on this host, one thread.
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

# Algorithmic bytes per item: distinct fields the reference formula reads plus writes,
# f32 = 4 B, each body-state gather counted per constraint touch. Coloured constraint
# stages: per constraint (one manifold) + per manifold point.
BYTE_MODEL = {
    "prepare": {"per_constraint": 268 + 452, "per_point": 0},
    "integrate_velocities": {"per_body": 100 + 24},
    "warm_start": {"per_constraint": 64 + 72 + 56 + 48, "per_point": 28},
    "push": {"per_constraint": 68 + 12 + 112 + 48, "per_point": 40},
    "integrate_positions": {"per_body": 56 + 56},
    "relax": {"per_constraint": 68 + 108 + 112 + 48, "per_point": 52},
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
    return reps * (m["per_constraint"] * ks["constraints"] + m["per_point"] * ks["points"])


def step_bytes(counts: dict, substeps: int) -> float:
    """Algorithmic bytes of whole steps (collide + prepare .. finalize) for the given work
    counts (bodies, collided, constraints, points summed over the steps)."""
    return sum(stage_bytes(k, counts, substeps) for k in BYTE_MODEL)


def solve_bytes(ks: dict, substeps: int) -> float:
    """Algorithmic bytes of the step kernel (prepare .. finalize) over the measured steps."""
    return sum(stage_bytes(k, ks, substeps) for k in BYTE_MODEL if k != "collide")


TRAFFIC_PROFILE = os.path.join("profiles", "k_solve_step_pile10k_pmc_bytes.md")


def measured_traffic():
    """HBM-side bytes per step-kernel launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (scaled by a copy kernel of known size), or None."""
    try:
        text = open(os.path.join(ROOT, TRAFFIC_PROFILE)).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch: (\d+)", text, re.M)
    return float(m.group(1)) if m else None


def _oracle_rate(scene, threads: int, seconds: float, max_steps: int):
    from oracle import pyoracle

    pyoracle.set_threads(threads)
    w = pyoracle.load_scene(scene)
    t0 = time.perf_counter()
    steps = 0
    while steps < max_steps and time.perf_counter() - t0 < seconds:
        w.step(scene.dt, scene.substeps)
        steps += 1
    dt = time.perf_counter() - t0
    return scene.dynamic_count * scene.substeps * steps / dt, steps, dt, w.phase_ms()


def cpu_baseline(scene, seconds: float, max_steps: int) -> dict:
    """The oracle (a C++ restatement of the reference step, OpenMP over the loops whose
    items are independent) on this host: all threads for ~70% of the budget, then one
    thread for the rest, each from step 1 of a fresh world."""
    from oracle import pyoracle

    threads = pyoracle.max_threads()
    value, steps, dt, phases = _oracle_rate(scene, threads, 0.7 * seconds, max_steps)
    value1, steps1, dt1, _ = _oracle_rate(scene, 1, 0.3 * seconds, max_steps)
    return {
        "value": value,
        "unit": "body-substeps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"CPU oracle (C++ restatement of the reference step, g++ -O3, OpenMP, {threads} threads), "
                  f"steps 1..{steps} of the same pile",
        "steps": steps,
        "seconds": dt,
        "phase_ms_per_step": {k: round(v / max(steps, 1), 3) for k, v in phases.items()},
        "single_thread": {"value": value1, "steps": steps1, "seconds": dt1},
    }


def batched_measurement(worlds: int, steps: int) -> dict:
    """Batched worlds (the log-uniform 8..256-body piles and stacks of
    natsel_amd.scenes.batched): every world stepped together, `steps` steps in one call
    after one untimed step; device time from HIP events around the call."""
    from natsel_amd import scenes
    from natsel_amd.batch import load_batch

    scene_list = scenes.batched(worlds)
    dt, sub = scene_list[0].dt, scene_list[0].substeps
    b = load_batch(scene_list)
    b.step(dt, sub, 1)  # uploads the worlds
    s0 = b.stats()
    b.step(dt, sub, steps)
    s1 = b.stats()
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
                     "traffic": None, "kernel": "k_batch_pairs + k_batch_collide + k_batch_solve (one step)",
                     "algorithmic_bytes_per_step": total / steps},
        "contacts_collided_per_step": counts["collided"] / steps,
        "constraints_per_step": counts["constraints"] / steps,
        "points_per_step": counts["points"] / steps,
        "error_worlds": s1["error_worlds"],
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--bodies", type=int, default=10_000)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--blocks", type=int, default=0, help="workgroups of the step kernel (0: sized from the largest phase)")
    ap.add_argument("--batched-worlds", type=int, default=4096, help="worlds of the batched measurement (0: skip it)")
    ap.add_argument("--batched-steps", type=int, default=240)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only copy this many MiB device to device 5 times (reference kernel for rocprofv3 --pmc byte counters)")
    args = ap.parse_args()
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
    world = load_scene(scene, device=0, solver_blocks=args.blocks)
    for _ in range(args.warmup):
        world.step(scene.dt, scene.substeps)
    world.reset_kernel_stats()
    world.set_kernel_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    marks = []
    for _ in range(args.steps):
        world.step(scene.dt, scene.substeps)  # returns after the step's device work
        marks.append(time.perf_counter())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    world.set_kernel_timing(False)
    ks = world.kernel_stats()
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
        "workgroups": ks["solver_blocks"],
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
        "step_kernel_phase_ms": {k: round(v, 3) for k, v in ks["phase_ms"].items()},
        "step_kernel_phase_count": ks["phase_count"],
        "host_phase_ms": {k: round(v, 3) for k, v in ks["host_ms"].items()},
        "contacts_collided_per_step": ks["collided"] / max(ks["steps"], 1),
        "constraints_per_step": ks["constraints"] / max(ks["steps"], 1),
        "points_per_step": ks["points"] / max(ks["steps"], 1),
        "pair_tests_per_step": ks["pair_tests"] / max(ks["steps"], 1),
        "fast_bodies": ks["fast_bodies"],
    }
    if not args.no_cpu:
        cb = cpu_baseline(scene, args.cpu_seconds, args.steps)
        k = min(cb["steps"], len(marks))
        # GPU rate over the same leading steps as the CPU sample, for a like-for-like ratio
        cb["gpu_value_same_steps"] = n * scene.substeps * k / (marks[k - 1] - t0)
        cb["gpu_over_cpu_same_steps"] = cb["gpu_value_same_steps"] / cb["value"]
        out["cpu_baseline"] = cb
    if args.batched_worlds > 0:
        out["batched"] = batched_measurement(args.batched_worlds, args.batched_steps)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
