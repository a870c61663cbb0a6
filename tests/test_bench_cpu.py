"""bench.py's host-side logic (no GPU): the settled window, the rocprofv3 CSV summaries over a
dispatch window, the committed traffic figures and the byte model's arithmetic."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_timed_window_is_the_end_of_the_pile_run():
    assert bench.settle_steps(20, 5) == 980
    assert bench.settle_steps(100, 5) == 900
    assert bench.settle_steps(1000, 5) == 5  # never fewer untimed steps than the warmup


def _trace(tmp_path, rows):
    p = tmp_path / "run_kernel_trace.csv"
    lines = ["Kind,Dispatch_Id,Kernel_Name,Start_Timestamp,End_Timestamp"]
    for i, (name, us) in enumerate(rows, 1):
        lines.append(f'KERNEL_DISPATCH,{i},"{name}",{1000 * i},{1000 * i + int(us * 1000)}')
    p.write_text("\n".join(lines) + "\n")
    return str(p)


def test_trace_summary_over_a_window(tmp_path):
    solve = "void nsl::(anonymous namespace)::k_solve_step(nsl::DeviceWorld, nsl::SolvePlan, nsl::StepParams)"
    collide = "nsl::(anonymous namespace)::k_collide(nsl::DeviceWorld, int)"
    wide = "nsl::(anonymous namespace)::k_collide_wide(nsl::DeviceWorld, int)"
    rows = []
    for k, us in enumerate([10.0, 20.0, 30.0, 40.0]):
        rows += [(collide, 1.0 + k), (wide, 99.0), (solve, us)]
    path = _trace(tmp_path, rows)
    s = bench.summarize_profile(path, "k_solve_step", "2:3")
    assert s["dispatches"] == 4 and s["window_dispatches"] == 2
    assert s["window_avg_us"] == pytest.approx(25.0)
    assert s["window_min_us"] == pytest.approx(20.0) and s["window_max_us"] == pytest.approx(30.0)
    c = bench.summarize_profile(path, "k_collide", "all")  # the prefix does not take k_collide_wide
    assert c["dispatches"] == 4 and c["window_avg_us"] == pytest.approx(2.5)


def test_counter_summary_sums_a_dispatch_counters(tmp_path):
    p = tmp_path / "run_counter_collection.csv"
    lines = ["Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value"]
    for d in range(1, 5):
        for part in range(3):  # one counter reported per XCD slice, summed per dispatch
            lines.append(f"{d},k_batch_solve(nsl::BatchDevice),FETCH_SIZE,{10.0 * d + part}")
    p.write_text("\n".join(lines) + "\n")
    s = bench.summarize_profile(str(p), "k_batch_solve", "3:4")
    assert s["window_dispatches"] == 2
    assert s["window_mean"]["FETCH_SIZE"] == pytest.approx(((30 + 31 + 32) + (40 + 41 + 42)) / 2)


def test_committed_traffic_figures_are_read():
    t = bench.measured_traffic()
    tb = bench.measured_traffic(bench.BATCH_TRAFFIC_PROFILE, "traffic_bytes_per_step")
    assert t is not None and t > 1e8  # above the step kernel's 144.9 MB of algorithmic bytes
    assert tb is not None and tb > 2e9


def test_byte_model_arithmetic():
    ks = {"bodies": 10, "collided": 7, "constraints": 3, "points": 5, "joints": 2}
    m = bench.BYTE_MODEL
    assert bench.stage_bytes("finalize", ks, 4) == 400 * 10
    assert bench.stage_bytes("integrate_positions", ks, 4) == 4 * 112 * 10
    assert bench.stage_bytes("collide", ks, 4) == m["collide"]["per_contact"] * 7
    push = m["push"]
    assert bench.stage_bytes("push", ks, 4) == 4 * (push["per_constraint"] * 3 + push["per_point"] * 5 + push["per_joint"] * 2)
    assert bench.step_bytes(ks, 4) == sum(bench.stage_bytes(k, ks, 4) for k in m)
    assert bench.solve_bytes(ks, 4) == sum(bench.stage_bytes(k, ks, 4) for k in bench.PER_SUBSTEP)
